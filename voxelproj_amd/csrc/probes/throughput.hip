// Throughput probe for the instruction mixes the projector kernels are built
// from (measurement program; not part of the library).  Every LDS access
// stays inside its workgroup's buffer and every loop has a fixed trip count.
//
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops \
//       -o /tmp/vp_probe voxelproj_amd/csrc/probes/throughput.hip && /tmp/vp_probe
//
// It reports, per CU and clock at the device's peak clock:
//   * LDS read bandwidth of ds_read_b32 / b64 / b128 at 1 and 2 workgroups
//     of 256 threads per CU (conflict-free rows);
//   * FP32 FMA rate with the multiplier in a VGPR, in an SGPR, as a DPP
//     row broadcast (v_fmac_f32_dpp row_newbcast), and the FMA-with-literal
//     form, 16 independent accumulators per lane, 4 and 8 waves per SIMD;
//   * v_readlane + FMA pairs;
//   * the backprojection tap loop per voxel-angle and wave (4 slices per
//     lane, 3 taps): weights by v_readlane, copied to VGPRs, or by DPP.
#include <hip/hip_runtime.h>
#include <stdio.h>

#define CHECK(x)                                                                           \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);          \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

template <int BYTES>
__global__ __launch_bounds__(256) void lds_read(float* out, int iters)
{
    __shared__ float4 buf[4096];  // 16 rows of 1024 floats
    const int t = threadIdx.x;
    for (int i = t; i < 4096; i += 256) buf[i] = make_float4(i, i + 1, i + 2, i + 3);
    __syncthreads();
    float acc = 0.f;
    const float* fb = reinterpret_cast<const float*>(buf);
    const int lane = t & 63;
    for (int it = 0; it < iters; ++it) {
        const int rot = (it * 7) & 15;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int row = (rot + k) & 15;
            if constexpr (BYTES == 4) {
                acc += fb[row * 1024 + lane];
            } else if constexpr (BYTES == 8) {
                const float2 v = *reinterpret_cast<const float2*>(fb + row * 1024 + 2 * lane);
                acc += v.x + v.y;
            } else {
                const float4 v = *reinterpret_cast<const float4*>(fb + row * 1024 + 4 * lane);
                acc += (v.x + v.y) + (v.z + v.w);
            }
        }
    }
    out[blockIdx.x * 256 + t] = acc;
}

// MODE 0: VGPR multiplier, 1: SGPR multiplier, 2: DPP row broadcast, 3: literal form
template <int MODE>
__global__ __launch_bounds__(256) void fma_rate(float* out, float a, float b, int iters)
{
    float acc[16];
    const float w = a + threadIdx.x * 1e-7f;
    float y = b + threadIdx.x;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = i;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if constexpr (MODE == 0)
                acc[i] = fmaf(w, y, acc[i]);
            else if constexpr (MODE == 1)
                acc[i] = fmaf(a, y, acc[i]);
            else if constexpr (MODE == 2)
                asm("v_fmac_f32_dpp %0, %1, %2 row_newbcast:5 row_mask:0xf bank_mask:0xf"
                    : "+v"(acc[i])
                    : "v"(w), "v"(y));
            else
                acc[i] = fmaf(y, acc[i], 0.999f);
        }
        y = fmaf(y, 0.5f, 1.0f);
    }
    float s = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) s += acc[i];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void readlane_fma(float* out, float a, float b, int iters)
{
    float acc[16];
    float w = a + threadIdx.x * 1e-7f;
    float y = b + threadIdx.x;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = i;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const float s = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, w), i));
            acc[i] = fmaf(s, y, acc[i]);
        }
        y = fmaf(y, 0.5f, 1.0f);
        w = fmaf(w, 0.5f, 0.25f);
    }
    float s = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) s += acc[i];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

__device__ __forceinline__ float rdl(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// Backprojection tap loop: per item (one voxel-angle of a wave, 4 slices per
// lane) one offset readlane, 3 ds_read_b128 taps, 3 weights, 12 FMAs.
// MODE 0: weights by v_readlane (SGPR FMAs); 1: copied to VGPRs; 2: DPP row
// broadcast (per-row weights); 3: the LDS reads only (taps summed).
template <int MODE>
__global__ __launch_bounds__(1024) void tap_loop(float* out, int iters)
{
    __shared__ float4 buf[21 * 2 * 64];  // 2 angles x 21 pixel rows x 256 floats
    const int t = threadIdx.x, lane = t & 63;
    for (int i = t; i < 21 * 2 * 64; i += 1024) buf[i] = make_float4(i, i * 0.5f, i * 0.25f, 1.0f);
    __syncthreads();
    const int foff = (((lane & 1) * 21) + (lane * 7) % 19) * 256;  // rows 0 .. 18 of angle lane & 1
    const float wm = 0.1f + lane * 1e-3f, w0 = 0.5f - lane * 1e-3f, wp = 0.05f;
    float acc[8][4];
    for (int v = 0; v < 8; ++v)
        for (int j = 0; j < 4; ++j) acc[v][j] = 0.f;
    const float* sw = reinterpret_cast<const float*>(buf) + lane * 4;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int L = (i + it) & 63;
            const int v = i & 7;
            const float* pr = sw + __builtin_amdgcn_readlane(foff, L);
            const float4 a = *reinterpret_cast<const float4*>(pr);
            const float4 b = *reinterpret_cast<const float4*>(pr + 256);
            const float4 c = *reinterpret_cast<const float4*>(pr + 512);
            if constexpr (MODE == 3) {
                acc[v][0] += a.x + b.x + c.x;
                acc[v][1] += a.y + b.y + c.y;
                acc[v][2] += a.z + b.z + c.z;
                acc[v][3] += a.w + b.w + c.w;
                continue;
            }
            float c0, c1, c2;
            if constexpr (MODE == 2) {
                asm volatile("v_mov_b32_dpp %0, %1 row_newbcast:5 row_mask:0xf bank_mask:0xf" : "=v"(c0) : "v"(wm));
                asm volatile("v_mov_b32_dpp %0, %1 row_newbcast:6 row_mask:0xf bank_mask:0xf" : "=v"(c1) : "v"(w0));
                asm volatile("v_mov_b32_dpp %0, %1 row_newbcast:7 row_mask:0xf bank_mask:0xf" : "=v"(c2) : "v"(wp));
            } else {
                c0 = rdl(wm, L);
                c1 = rdl(w0, L);
                c2 = rdl(wp, L);
                if constexpr (MODE == 1) {
                    asm volatile("v_mov_b32 %0, %1" : "=v"(c0) : "s"(c0));
                    asm volatile("v_mov_b32 %0, %1" : "=v"(c1) : "s"(c1));
                    asm volatile("v_mov_b32 %0, %1" : "=v"(c2) : "s"(c2));
                }
            }
            acc[v][0] = fmaf(c0, a.x, acc[v][0]);
            acc[v][1] = fmaf(c0, a.y, acc[v][1]);
            acc[v][2] = fmaf(c0, a.z, acc[v][2]);
            acc[v][3] = fmaf(c0, a.w, acc[v][3]);
            acc[v][0] = fmaf(c1, b.x, acc[v][0]);
            acc[v][1] = fmaf(c1, b.y, acc[v][1]);
            acc[v][2] = fmaf(c1, b.z, acc[v][2]);
            acc[v][3] = fmaf(c1, b.w, acc[v][3]);
            acc[v][0] = fmaf(c2, c.x, acc[v][0]);
            acc[v][1] = fmaf(c2, c.y, acc[v][1]);
            acc[v][2] = fmaf(c2, c.z, acc[v][2]);
            acc[v][3] = fmaf(c2, c.w, acc[v][3]);
        }
    }
    float s = 0;
    for (int v = 0; v < 8; ++v)
        for (int j = 0; j < 4; ++j) s += acc[v][j];
    out[blockIdx.x * 1024 + t] = s;
}

int main()
{
    hipDeviceProp_t p;
    CHECK(hipGetDeviceProperties(&p, 0));
    const int cus = p.multiProcessorCount;
    const double clk = p.clockRate * 1e3;
    printf("%d CUs, peak clock %.0f MHz\n", cus, clk / 1e6);
    float* out;
    CHECK(hipMalloc(&out, (size_t)cus * 16 * 256 * 4));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    // best of 3 timed launches after a short warm-up; returns per-CU-per-clock rate of `work`
    auto timed = [&](auto launch, int nb, int iters, double work_per_iter_thread, int threads) -> double {
        float best = 1e30f;
        for (int r = 0; r < 3; ++r) {
            launch(nb, 50);
            if (hipDeviceSynchronize() != hipSuccess || hipEventRecord(e0) != hipSuccess) return -1.0;
            launch(nb, iters);
            float ms = 0;
            if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                return -1.0;
            if (ms < best) best = ms;
        }
        const double total = work_per_iter_thread * iters * (double)nb * threads;
        return total / (best * 1e-3) / cus / clk;
    };
    for (int bpc : {1, 2}) {
        printf("LDS read, %d workgroup(s) of 256 per CU: b32 %.1f, b64 %.1f, b128 %.1f B/clk/CU\n", bpc,
               timed([&](int nb, int it) { lds_read<4><<<nb, 256>>>(out, it); }, cus * bpc, 20000, 32, 256),
               timed([&](int nb, int it) { lds_read<8><<<nb, 256>>>(out, it); }, cus * bpc, 20000, 64, 256),
               timed([&](int nb, int it) { lds_read<16><<<nb, 256>>>(out, it); }, cus * bpc, 20000, 128, 256));
    }
    for (int wps : {4, 8}) {
        const int nb = cus * wps;  // 256-thread workgroups: wps waves per SIMD
        printf("FP32 FMA lanes/clk/CU, %d waves per SIMD: VGPR %.1f, SGPR %.1f, DPP %.1f, literal %.1f, "
               "readlane+FMA %.1f\n", wps,
               timed([&](int n, int it) { fma_rate<0><<<n, 256>>>(out, 0.5f, 1.f, it); }, nb, 40000, 16, 256),
               timed([&](int n, int it) { fma_rate<1><<<n, 256>>>(out, 0.5f, 1.f, it); }, nb, 40000, 16, 256),
               timed([&](int n, int it) { fma_rate<2><<<n, 256>>>(out, 0.5f, 1.f, it); }, nb, 40000, 16, 256),
               timed([&](int n, int it) { fma_rate<3><<<n, 256>>>(out, 0.5f, 1.f, it); }, nb, 40000, 16, 256),
               timed([&](int n, int it) { readlane_fma<<<n, 256>>>(out, 0.5f, 1.f, it); }, nb, 40000, 16, 256));
    }
    // tap loop: items per clk per CU -> clk per item
    const char* names[] = {"weights by v_readlane (SGPR FMAs)", "weights copied to VGPRs",
                           "weights by DPP row broadcast", "LDS taps only"};
    for (int m = 0; m < 4; ++m) {
        auto go = [&](int n, int it) {
            switch (m) {
                case 0: tap_loop<0><<<n, 1024>>>(out, it); break;
                case 1: tap_loop<1><<<n, 1024>>>(out, it); break;
                case 2: tap_loop<2><<<n, 1024>>>(out, it); break;
                default: tap_loop<3><<<n, 1024>>>(out, it); break;
            }
        };
        const double rate = timed(go, cus * 2, 4000, 16.0 / 64.0, 1024);  // items per thread-iter = 16 / 64 lanes
        printf("tap loop, %s: %.2f clk per voxel-angle and wave per CU\n", names[m], 1.0 / rate);
    }
    return 0;
}
