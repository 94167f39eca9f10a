// Double-precision projector pair, for gradient checks and FP64 solvers.
//
// The reference is float32-only (projectors.cu:37); these kernels evaluate the
// same exact-trapezoid operator with every position and weight in FP64.  They
// favour simplicity over speed (one thread per output element):
//   A : thread per sinogram element, gathering the <= 5 voxels per line of the
//       volume whose footprint can reach the pixel;
//   AT: thread per volume element, gathering the 3 taps per angle.
// Both layouts are handled through element strides; outputs accumulate.  A
// per-angle detector translation du shifts every position u by -du.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vp_launch.hpp"

namespace vp {
namespace {

struct Strides {
    int64_t vy, vx, vh;  // volume: iy, ix, h
    int64_t pa, pp, ph;  // sinogram: angle, pixel, h
};

struct Taps64 {
    double r, wm, w0, wp;
};

__device__ inline double tail64(double z, double a, double b, double h, double cd)
{
    if (z <= a) return 0.5 - h * z;
    const double e = fmax(b - z, 0.0);
    return cd * e * e;
}

__device__ inline Taps64 taps64(double X, double Y, double pc, double c, double s, double a,
                                double b, double h, double cd)
{
    const double u = X * c + Y * s + pc;
    Taps64 t;
    t.r = floor(u + 0.5);
    const double d = u - t.r;
    t.wm = tail64(0.5 + d, a, b, h, cd);
    t.wp = tail64(0.5 - d, a, b, h, cd);
    t.w0 = 1.0 - t.wm - t.wp;
    return t;
}

__device__ inline void angle64(float phi, double& c, double& s, double& a, double& b, double& h,
                               double& cd)
{
    c = cos((double)phi);
    s = sin((double)phi);
    const double ac = fabs(c), as = fabs(s);
    a = 0.5 * fabs(ac - as);
    b = 0.5 * (ac + as);
    h = 1.0 / fmax(ac, as);
    cd = (b > a) ? 0.5 * h / (b - a) : 0.0;
}

__global__ void forward_f64_kernel(const double* __restrict__ x, double* __restrict__ y,
                                   const float* __restrict__ angles,
                                   const float* __restrict__ du, int A, int SY, int SX,
                                   int H, int PX, Strides st)
{
    const int64_t n = (int64_t)A * PX * H;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int h = (int)(i % H);
    const int p = (int)((i / H) % PX);
    const int ia = (int)(i / ((int64_t)H * PX));
    double c, s, a, b, hh, cd;
    angle64(angles[ia], c, s, a, b, hh, cd);
    const double xc = 0.5 * (SX - 1), yc = 0.5 * (SY - 1);
    const double pc = 0.5 * (PX - 1) - (du ? (double)du[ia] : 0.0);
    const bool modeX = fabs(c) >= fabs(s);
    const int nrows = modeX ? SY : SX;
    const int nalong = modeX ? SX : SY;
    const double kap = modeX ? c : s;
    const double oth = modeX ? s : c;
    const double ca = modeX ? xc : yc;
    const double cr = modeX ? yc : xc;
    double acc = 0.0;
    for (int row = 0; row < nrows; ++row) {
        const double R = row - cr;
        // u(j) = (j - ca) kap + R oth + pc; pixel p is reached by centres in [p-1, p+1]
        const double j0 = (p - 1.5 - R * oth - pc) / kap + ca;
        const double j1 = (p + 1.5 - R * oth - pc) / kap + ca;
        const int jlo = max(0, (int)floor(fmin(j0, j1)) - 1);
        const int jhi = min(nalong - 1, (int)ceil(fmax(j0, j1)) + 1);
        for (int j = jlo; j <= jhi; ++j) {
            const double X = modeX ? (j - xc) : R;
            const double Y = modeX ? R : (j - yc);
            const Taps64 t = taps64(X, Y, pc, c, s, a, b, hh, cd);
            const double dq = p - t.r;
            const double w = dq == -1.0 ? t.wm : dq == 0.0 ? t.w0 : dq == 1.0 ? t.wp : 0.0;
            if (w != 0.0) {
                const int ix = modeX ? j : row;
                const int iy = modeX ? row : j;
                acc += w * x[iy * st.vy + ix * st.vx + h * st.vh];
            }
        }
    }
    y[ia * st.pa + p * st.pp + h * st.ph] += acc;
}

__global__ void backward_f64_kernel(const double* __restrict__ y, double* __restrict__ x,
                                    const float* __restrict__ angles,
                                    const float* __restrict__ du, int A, int SY, int SX,
                                    int H, int PX, Strides st)
{
    const int64_t n = (int64_t)SY * SX * H;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int h = (int)(i % H);
    const int ix = (int)((i / H) % SX);
    const int iy = (int)(i / ((int64_t)H * SX));
    const double xc = 0.5 * (SX - 1), yc = 0.5 * (SY - 1);
    double acc = 0.0;
    for (int ia = 0; ia < A; ++ia) {
        double c, s, a, b, hh, cd;
        angle64(angles[ia], c, s, a, b, hh, cd);
        const double pc = 0.5 * (PX - 1) - (du ? (double)du[ia] : 0.0);
        const Taps64 t = taps64(ix - xc, iy - yc, pc, c, s, a, b, hh, cd);
        const double w[3] = {t.wm, t.w0, t.wp};
        for (int k = 0; k < 3; ++k) {
            const int64_t p = (int64_t)t.r - 1 + k;
            if (p >= 0 && p < PX) acc += w[k] * y[ia * st.pa + p * st.pp + h * st.ph];
        }
    }
    x[iy * st.vy + ix * st.vx + h * st.vh] += acc;
}

Strides strides(int z_order, const Geom& g)
{
    Strides st;
    if (z_order == 0) {
        st.vy = g.SX * g.H; st.vx = g.H; st.vh = 1;
        st.pa = g.PX * g.H; st.pp = g.H; st.ph = 1;
    } else {
        st.vy = g.SX; st.vx = 1; st.vh = g.SY * g.SX;
        st.pa = g.PX; st.pp = 1; st.ph = (int64_t)g.A * g.PX;
    }
    return st;
}

}  // namespace

hipError_t launch_forward_f64(int z_order, const double* x, double* y, const float* angles,
                              const Geom& g, hipStream_t s)
{
    const int64_t n = (int64_t)g.A * g.PX * g.H;
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(forward_f64_kernel, dim3(nb), dim3(256), 0, s, x, y, angles, g.du, g.A,
                       (int)g.SY, (int)g.SX, (int)g.H, (int)g.PX, strides(z_order, g));
    return hipGetLastError();
}

hipError_t launch_backward_f64(int z_order, const double* y, double* x, const float* angles,
                               const Geom& g, hipStream_t s)
{
    const int64_t n = g.SY * g.SX * g.H;
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(backward_f64_kernel, dim3(nb), dim3(256), 0, s, y, x, angles, g.du, g.A,
                       (int)g.SY, (int)g.SX, (int)g.H, (int)g.PX, strides(z_order, g));
    return hipGetLastError();
}

}  // namespace vp
