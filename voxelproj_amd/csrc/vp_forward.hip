// Forward projection y (+)= A x in z_order 0 (reference: sp_forward_z0.cuh:107-185).
//
// The reference walks each ray through a texture and evaluates the footprint
// per visit.  Here A is evaluated as the exact transpose of the voxel-driven
// footprint (the same per-voxel r and weights as the backprojector), gathered
// into output-stationary accumulators:
//   wave  = one angle x P consecutive detector pixels x 64 H slices (lane = slice);
//   rows  = lines of the volume along the axis whose direction cosine is
//           larger (|kappa| >= 1/sqrt 2), so at most two voxels of a line
//           share a centre pixel r;
//   per line, lane m evaluates the footprint of the m-th voxel of the line
//   segment that touches the wave's pixels (in increasing-r order), the
//   segment's slices are staged in LDS, and a statically unrolled sweep over
//   r = p0-1 .. p0+P adds each voxel to accumulators r-1, r, r+1 -- so every
//   accumulator index is a compile-time constant.
// No atomics; the summation order is fixed, so results are deterministic.
#include "vp_device.hpp"
#include "vp_launch.hpp"

namespace vp {
namespace {

__device__ __forceinline__ float rdlane(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ int rdlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

__device__ __forceinline__ int popc64(uint64_t m) { return __builtin_popcountll(m); }

constexpr int kSeg = 64;       // max voxels of a line segment (one per lane)
constexpr int kRowsPerPart = 16;

template <int P, int NW>
__global__ __launch_bounds__(64 * NW) void forward_z0_kernel(
    const float* __restrict__ x, float* __restrict__ y, const float* __restrict__ angles,
    int A, int SY, int SX, int H, int PX, double scale, double inv_scale, int accumulate)
{
    // Segment length bound: the r range [p0-1, p0+P] plus one pixel of margin
    // on both sides spans P+5 detector units, i.e. at most (P+5)*sqrt 2 + 4
    // voxels along a line.
    static_assert((P + 5) * 1.4143 + 4 <= kSeg, "segment may exceed one wave");
    __shared__ float s_x[NW][kSeg][64];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int a = blockIdx.y;
    const int h = blockIdx.z * 64 + lane;
    const int p0 = (blockIdx.x * NW + wave) * P;
    if (p0 >= PX) return;  // whole wave; no block-level barriers follow

    const AngleK k = make_angle(angles[a], scale, inv_scale);
    const float cc = k.chi + k.clo;
    const float ss = k.shi + k.slo;
    const bool modeX = fabsf(cc) >= fabsf(ss);
    const float xc = 0.5f * (float)(SX - 1);
    const float yc = 0.5f * (float)(SY - 1);
    const float pc = 0.5f * (float)(PX - 1);
    const int nrows = modeX ? SY : SX;
    const int nalong = modeX ? SX : SY;
    const float kap = modeX ? cc : ss;
    const float oth = modeX ? ss : cc;
    const float ca = modeX ? xc : yc;  // centre along the line
    const float cr = modeX ? yc : xc;  // centre across lines
    const int64_t st_along = modeX ? (int64_t)H : (int64_t)SX * H;
    const int64_t st_row = modeX ? (int64_t)SX * H : (int64_t)H;
    const int pl = p0 - 1;   // smallest centre pixel touching [p0, p0+P)
    const int ph = p0 + P;   // largest
    const float inv_kap = 1.0f / kap;
    const int hc = min(h, H - 1);  // clamped slice for loads of idle lanes

    float acc[P], part[P];
#pragma unroll
    for (int i = 0; i < P; ++i) acc[i] = part[i] = 0.0f;

    for (int row = 0; row < nrows; ++row) {
        // Along-line coordinate j with u(j) ~ ub + j*kap.
        const float R = (float)row - cr;
        const float ub = fmaf(-ca, kap, fmaf(R, oth, pc));
        const float ja = ((float)pl - 1.5f - ub) * inv_kap;
        const float jb = ((float)ph + 1.5f - ub) * inv_kap;
        const int jlo = max(0, (int)floorf(fminf(ja, jb)) - 1);
        const int jhi = min(nalong - 1, (int)ceilf(fmaxf(ja, jb)) + 1);
        const int n = jhi - jlo + 1;
        if (n <= 0) continue;  // uniform

        // Lane m: m-th voxel in increasing-u order.
        const int m = lane;
        const int jm = kap >= 0.0f ? jlo + min(m, n - 1) : jhi - min(m, n - 1);
        const float Xv = modeX ? (float)jm - xc : R;
        const float Yv = modeX ? R : (float)jm - yc;
        const Taps tp = voxel_taps(Xv, Yv, pc, k);
        const int r = (m < n) ? tp.r : 0x3fffffff;

        // Stage the segment (all 64 slices of each voxel) in this wave's LDS.
        const float* xrow = x + (int64_t)row * st_row + hc;
        for (int b = 0; b < n; b += 8) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int jj = rdlane(jm, min(b + i, n - 1));
                s_x[wave][b + i][lane] = xrow[(int64_t)jj * st_along];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        int mm = popc64(__ballot(r < pl));
        const int mend = popc64(__ballot(r <= ph));
#pragma unroll
        for (int q = 0; q < P + 2; ++q) {
#pragma unroll
            for (int rep = 0; rep < 2; ++rep) {
                if (mm < mend && rdlane(r, mm) == pl + q) {
                    const float xv = s_x[wave][mm][lane];
                    const float wm = rdlane(tp.wm, mm);
                    const float w0 = rdlane(tp.w0, mm);
                    const float wp = rdlane(tp.wp, mm);
                    if (q - 2 >= 0 && q - 2 < P) part[q - 2] = fmaf(wm, xv, part[q - 2]);
                    if (q - 1 >= 0 && q - 1 < P) part[q - 1] = fmaf(w0, xv, part[q - 1]);
                    if (q < P) part[q] = fmaf(wp, xv, part[q]);
                    ++mm;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        if ((row % kRowsPerPart) == kRowsPerPart - 1) {
#pragma unroll
            for (int i = 0; i < P; ++i) {
                acc[i] += part[i];
                part[i] = 0.0f;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < P; ++i) acc[i] += part[i];

    if (h < H) {
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const int p = p0 + i;
            if (p < PX) {
                float* py = y + ((int64_t)a * PX + p) * H + h;
                *py = accumulate ? (*py + acc[i]) : acc[i];
            }
        }
    }
}

constexpr int kTile = 64;

__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ in,
                                                        float* __restrict__ out, int64_t R,
                                                        int64_t C, int accumulate)
{
    __shared__ float t[kTile][kTile + 1];
    const int64_t c0 = (int64_t)blockIdx.x * kTile;
    const int64_t r0 = (int64_t)blockIdx.y * kTile;
    const int tx = threadIdx.x & 63;
    const int ty = threadIdx.x >> 6;
    for (int i = ty; i < kTile; i += 4) {
        const int64_t r = r0 + i, c = c0 + tx;
        t[i][tx] = (r < R && c < C) ? in[r * C + c] : 0.0f;
    }
    __syncthreads();
    for (int i = ty; i < kTile; i += 4) {
        const int64_t c = c0 + i, r = r0 + tx;
        if (r < R && c < C) {
            float* po = out + c * R + r;
            *po = accumulate ? (*po + t[tx][i]) : t[tx][i];
        }
    }
}

}  // namespace

hipError_t launch_forward_z0(const float* x, float* y, const float* angles, const Geom& g,
                             int accumulate, hipStream_t s)
{
    constexpr int P = 32, NW = 4;
    const double scale = ldexp(1.0, g.bits);
    const double inv_scale = ldexp(1.0, -g.bits);
    dim3 grid((unsigned)((g.PX + P * NW - 1) / (P * NW)), (unsigned)g.A,
              (unsigned)((g.H + 63) / 64));
    hipLaunchKernelGGL((forward_z0_kernel<P, NW>), grid, dim3(64 * NW), 0, s, x, y, angles, g.A,
                       (int)g.SY, (int)g.SX, (int)g.H, (int)g.PX, scale, inv_scale, accumulate);
    return hipGetLastError();
}

hipError_t launch_transpose(const float* in, float* out, int64_t R, int64_t C, int accumulate,
                            hipStream_t s)
{
    dim3 grid((unsigned)((C + kTile - 1) / kTile), (unsigned)((R + kTile - 1) / kTile));
    hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, s, in, out, R, C, accumulate);
    return hipGetLastError();
}

}  // namespace vp
