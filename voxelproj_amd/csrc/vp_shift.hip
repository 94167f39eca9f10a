// Detector translation along H (dv, in slices), as linear interpolation of
// the sinogram along H.  The operator acts slice by slice in the plane, so
// translating the detector by dv_a at angle a is the same as computing the
// untranslated sinogram and sampling it at h + dv_a:
//     y[a, p, h] = (1 - f) y0[a, p, h + k] + f y0[a, p, h + k + 1],
//     k = floor(dv_a), f = dv_a - k, with y0 = 0 outside [0, H).
// The backprojector applies the exact adjoint to y before AT:
//     g[a, p, j] = (1 - f) y[a, p, j - k] + f y[a, p, j - k - 1].
// Both are gathers (one thread per output element), so no atomics.  Slabs:
// `in` holds global slices [in_off, in_off + Hin), `out` [out_off, out_off +
// Hout), which lets the H-slab streaming path process a slab with its halo.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "vp_launch.hpp"

namespace vp {
namespace {

template <typename T>
__global__ void vshift_kernel(int z_order, int transpose, const T* __restrict__ in,
                              T* __restrict__ out, int A, int64_t PX, int64_t Hin, int64_t in_off,
                              int64_t Hout, int64_t out_off, const float* __restrict__ dv,
                              int accumulate)
{
    const int64_t n = (int64_t)A * PX * Hout;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t h, p;
    int a;
    if (z_order == 0) {  // (A, PX, H): h fastest
        h = i % Hout;
        p = (i / Hout) % PX;
        a = (int)(i / (Hout * PX));
    } else {             // (H, A, PX): p fastest
        p = i % PX;
        a = (int)((i / PX) % A);
        h = i / (PX * (int64_t)A);
    }
    const double d = fmin(fmax((double)dv[a], -1e15), 1e15);  // keeps k in range
    const double kf = floor(d);
    const T f = (T)(d - kf);
    const int64_t k = (int64_t)kf;
    const int64_t gh = h + out_off;  // global slice of this output element
    // the two input slices (global) and their weights
    int64_t j0, j1;
    T w0, w1;
    if (!transpose) {
        j0 = gh + k;
        j1 = gh + k + 1;
        w0 = (T)1 - f;
        w1 = f;
    } else {
        j0 = gh - k;
        j1 = gh - k - 1;
        w0 = (T)1 - f;
        w1 = f;
    }
    j0 -= in_off;
    j1 -= in_off;
    auto at = [&](int64_t j) -> T {
        if (j < 0 || j >= Hin) return (T)0;
        return z_order == 0 ? in[((int64_t)a * PX + p) * Hin + j]
                            : in[(j * A + a) * PX + p];
    };
    const T v = w0 * at(j0) + w1 * at(j1);
    T* o = z_order == 0 ? &out[((int64_t)a * PX + p) * Hout + h] : &out[(h * A + a) * PX + p];
    *o = accumulate ? *o + v : v;
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, int64_t n, T v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

template <typename T>
hipError_t launch_fill(T* p, int64_t n, T v, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256;
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fill_kernel<T>, dim3((unsigned)nb), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

template hipError_t launch_fill<float>(float*, int64_t, float, hipStream_t);
template hipError_t launch_fill<double>(double*, int64_t, double, hipStream_t);

template <typename T>
hipError_t launch_vshift(int z_order, int transpose, const T* in, T* out, int A, int64_t PX,
                         int64_t Hin, int64_t in_off, int64_t Hout, int64_t out_off,
                         const float* dv, int accumulate, hipStream_t s)
{
    const int64_t n = (int64_t)A * PX * Hout;
    if (n == 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256;
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(vshift_kernel<T>, dim3((unsigned)nb), dim3(256), 0, s, z_order, transpose,
                       in, out, A, PX, Hin, in_off, Hout, out_off, dv, accumulate);
    return hipGetLastError();
}

template hipError_t launch_vshift<float>(int, int, const float*, float*, int, int64_t, int64_t,
                                         int64_t, int64_t, int64_t, const float*, int, hipStream_t);
template hipError_t launch_vshift<double>(int, int, const double*, double*, int, int64_t, int64_t,
                                          int64_t, int64_t, int64_t, const float*, int,
                                          hipStream_t);

}  // namespace vp
