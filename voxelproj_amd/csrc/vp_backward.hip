// Voxel-driven backprojection x += AT y (reference: sp_backward_z0.cuh:8-100,
// sp_backward_z2.cuh:8-101), gather form, no atomics, deterministic.
//
// Work decomposition
//   block  = TX x TY voxel tile x one chunk of 64 consecutive H slices;
//   wave   = 64 voxels of the tile; lane = H slice (so every weight is
//            wave-uniform and each sinogram read is one 256-byte row);
//   angles = processed in chunks of AC.  Per chunk the sinogram window that
//            the tile touches (WIN pixels x 64 slices per angle) is staged in
//            LDS, so any number of angles works (the reference's 512-angle
//            shared-memory table limit does not exist here).
// Weights: lane v computes the footprint of voxel v for the current angle
// (one evaluation per voxel-angle, shared by all 64 slices), and the 64
// voxel iterations pick them up with v_readlane into SGPRs.
#include "vp_device.hpp"
#include "vp_launch.hpp"

namespace vp {
namespace {

__device__ __forceinline__ float rdlane(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ int rdlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

constexpr int kRow = 65;  // LDS row: 64 slices + 1 pad word (conflict-free transposed stores)

template <int Z, int TX, int TY, int AC, int WIN>
__global__ __launch_bounds__(TX * TY) void backproject_kernel(
    const float* __restrict__ y, float* __restrict__ x, const float* __restrict__ angles,
    int A, int SY, int SX, int H, int PX, double scale, double inv_scale)
{
    constexpr int NT = TX * TY;
    constexpr int NW = NT / 64;
    static_assert(NT % 64 == 0, "tile must hold whole waves");
    constexpr int STAGE = AC * WIN * kRow;
    constexpr int OUTT = 64 * kRow;
    constexpr int LDSN = STAGE > OUTT ? STAGE : OUTT;
    __shared__ float s_y[LDSN];
    __shared__ AngleK s_k[AC];
    __shared__ int s_rlo[AC];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int tx0 = blockIdx.x * TX;
    const int ty0 = blockIdx.y * TY;
    const int h0 = blockIdx.z * 64;
    const float xc = 0.5f * (float)(SX - 1);
    const float yc = 0.5f * (float)(SY - 1);
    const float pc = 0.5f * (float)(PX - 1);

    // Voxel whose footprint this lane evaluates (wave-local index = lane).
    const int vi = wave * 64 + lane;
    const int vx = tx0 + vi % TX;
    const int vy = ty0 + vi / TX;
    const float X = (float)vx - xc;
    const float Y = (float)vy - yc;

    float acc[64];
#pragma unroll
    for (int v = 0; v < 64; ++v) acc[v] = 0.0f;

    for (int c0 = 0; c0 < A; c0 += AC) {
        __syncthreads();
        if (tid < AC) {
            const int a = c0 + tid;
            const AngleK k = make_angle(a < A ? angles[a] : 0.0f, scale, inv_scale);
            s_k[tid] = k;
            const float xa = (float)tx0 - xc;
            const float xb = (float)(min(tx0 + TX, SX) - 1) - xc;
            const float ya = (float)ty0 - yc;
            const float yb = (float)(min(ty0 + TY, SY) - 1) - yc;
            int r = voxel_r(xa, ya, pc, k);
            r = min(r, voxel_r(xb, ya, pc, k));
            r = min(r, voxel_r(xa, yb, pc, k));
            r = min(r, voxel_r(xb, yb, pc, k));
            s_rlo[tid] = r - 2;  // one spare pixel below r-1 for rounding
        }
        __syncthreads();
        for (int e = tid; e < AC * WIN * 64; e += NT) {
            int a, k, hl;
            if (Z == 0) {
                hl = e & 63;
                k = (e >> 6) % WIN;
                a = e / (64 * WIN);
            } else {
                k = e % WIN;
                hl = (e / WIN) & 63;
                a = e / (64 * WIN);
            }
            const int ag = c0 + a;
            const int p = s_rlo[a] + k;
            const int h = h0 + hl;
            float v = 0.0f;
            if (ag < A && p >= 0 && p < PX && h < H) {
                v = (Z == 0) ? y[((int64_t)ag * PX + p) * H + h]
                             : y[((int64_t)h * A + ag) * PX + p];
            }
            s_y[(a * WIN + k) * kRow + hl] = v;
        }
        __syncthreads();

        float part[64];
#pragma unroll
        for (int v = 0; v < 64; ++v) part[v] = 0.0f;
        const int na = min(AC, A - c0);
        for (int a = 0; a < na; ++a) {
            const AngleK k = s_k[a];
            const Taps t = voxel_taps(X, Y, pc, k);
            int off = t.r - 1 - s_rlo[a];
            off = min(max(off, 0), WIN - 3);
            const float* base = s_y + a * WIN * kRow + lane;
#pragma unroll
            for (int v = 0; v < 64; ++v) {
                const int o = rdlane(off, v);
                const float wm = rdlane(t.wm, v);
                const float w0 = rdlane(t.w0, v);
                const float wp = rdlane(t.wp, v);
                const float* pr = base + o * kRow;
                float s = part[v];
                s = fmaf(wm, pr[0], s);
                s = fmaf(w0, pr[kRow], s);
                s = fmaf(wp, pr[2 * kRow], s);
                part[v] = s;
            }
        }
#pragma unroll
        for (int v = 0; v < 64; ++v) acc[v] += part[v];
    }

    const int h = h0 + lane;
    if (Z == 0) {
        if (h < H) {
#pragma unroll
            for (int v = 0; v < 64; ++v) {
                const int wv = wave * 64 + v;
                const int ix = tx0 + wv % TX;
                const int iy = ty0 + wv / TX;
                if (ix < SX && iy < SY) {
                    float* px = x + ((int64_t)iy * SX + ix) * H + h;
                    *px += acc[v];
                }
            }
        }
    } else {
        // Transpose through LDS so the stores run along ix.
        for (int w = 0; w < NW; ++w) {
            __syncthreads();
            if (wave == w) {
#pragma unroll
                for (int v = 0; v < 64; ++v) s_y[v * kRow + lane] = acc[v];
            }
            __syncthreads();
            for (int e = tid; e < 64 * 64; e += NT) {
                const int v = e & 63;
                const int hl = e >> 6;
                const int wv = w * 64 + v;
                const int ix = tx0 + wv % TX;
                const int iy = ty0 + wv / TX;
                const int hh = h0 + hl;
                if (ix < SX && iy < SY && hh < H) {
                    float* px = x + ((int64_t)hh * SY + iy) * SX + ix;
                    *px += s_y[v * kRow + hl];
                }
            }
        }
    }
}

template <int Z>
hipError_t launch_backward_t(const float* y, float* x, const float* angles, const Geom& g,
                             hipStream_t s)
{
    constexpr int TX = 16, TY = 16, AC = 4, WIN = 28;
    const double scale = ldexp(1.0, g.bits);
    const double inv_scale = ldexp(1.0, -g.bits);
    dim3 grid((unsigned)((g.SX + TX - 1) / TX), (unsigned)((g.SY + TY - 1) / TY),
              (unsigned)((g.H + 63) / 64));
    hipLaunchKernelGGL((backproject_kernel<Z, TX, TY, AC, WIN>), grid, dim3(TX * TY), 0, s, y, x,
                       angles, g.A, (int)g.SY, (int)g.SX, (int)g.H, (int)g.PX, scale, inv_scale);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_backward(int z_order, const float* y, float* x, const float* angles,
                           const Geom& g, hipStream_t s)
{
    return z_order == 0 ? launch_backward_t<0>(y, x, angles, g, s)
                        : launch_backward_t<2>(y, x, angles, g, s);
}

}  // namespace vp
