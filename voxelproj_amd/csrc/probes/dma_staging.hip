// Staging-throughput probe: how fast a workgroup can move 1 KB rows (64
// lanes x 16 bytes) from global memory into LDS, the operation the
// projectors stage volume lines and sinogram windows with (measurement
// program; not part of the library).  Every source offset is a 1 KB-aligned
// offset below the span, inside the allocation; every LDS write stays inside
// the workgroup's buffer; every loop has a fixed trip count.
//
//   hipcc --offload-arch=gfx950 -O3 -o /tmp/vp_dma voxelproj_amd/csrc/probes/dma_staging.hip && /tmp/vp_dma
//
// Two forms, one workgroup per CU, W waves each with D 1 KB rows in flight:
//   dma:  buffer_load_dwordx4 ... lds (global -> LDS without registers);
//   vgpr: buffer_load_dwordx4 into registers, then ds_write_b128;
// over sources that stay in L2 (1 MB), fit the Infinity Cache (128 MB) or
// stream from HBM (1 GB).  Rows are picked by a stride walk, so consecutive
// rows of a wave are far apart.  Reported: bytes per clock and CU at the
// 2.4 GHz peak clock, and per wave.
#include <hip/hip_runtime.h>
#include <stdio.h>

#include <algorithm>
#include <vector>

#define CHECK(x)                                                                           \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);          \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)0xffffffff, 0x00020000);
}

// MODE 0: global -> LDS DMA; MODE 1: load to registers, then ds_write_b128.
template <int WAVES, int DEPTH, int MODE>
__global__ __launch_bounds__(64 * WAVES) void stage_rows(const float* src, uint32_t span_kb, int iters,
                                                         float* sink)
{
    __shared__ float4 buf[WAVES][DEPTH][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const __amdgpu_buffer_rsrc_t r = raw_rsrc(src);
    uint32_t kb = (blockIdx.x * WAVES + wave) * 7919u % span_kb;
    float acc = 0.f;
    for (int it = 0; it < iters; ++it) {
        if constexpr (MODE == 0) {
#pragma unroll
            for (int d = 0; d < DEPTH; ++d) {
                kb = (kb + 12345u) % span_kb;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)&buf[wave][d][0],
                                                         16, lane * 16, kb * 1024u, 0, 0);
            }
            __builtin_amdgcn_s_waitcnt(0);
        } else {
            float4 v[DEPTH];
#pragma unroll
            for (int d = 0; d < DEPTH; ++d) {
                kb = (kb + 12345u) % span_kb;
                v[d] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, lane * 16, kb * 1024u, 0));
            }
#pragma unroll
            for (int d = 0; d < DEPTH; ++d) {
                buf[wave][(d + it) % DEPTH][lane] = v[d];
                acc += v[d].x + v[d].w;  // keeps every load alive
            }
        }
    }
    __syncthreads();
    sink[blockIdx.x * 64 * WAVES + threadIdx.x] = buf[0][0][1].x + buf[WAVES - 1][DEPTH - 1][63].y + acc;
}

template <int WAVES, int DEPTH, int MODE>
int run(const float* src, uint32_t span_kb, float* sink, int cus)
{
    const int iters = 64;
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL((stage_rows<WAVES, DEPTH, MODE>), dim3(cus), dim3(64 * WAVES), 0, 0, src, span_kb, iters, sink);
    CHECK(hipDeviceSynchronize());
    std::vector<float> ts;
    for (int rep = 0; rep < 5; ++rep) {
        CHECK(hipEventRecord(e0));
        hipLaunchKernelGGL((stage_rows<WAVES, DEPTH, MODE>), dim3(cus), dim3(64 * WAVES), 0, 0, src, span_kb, iters,
                           sink);
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        ts.push_back(ms);
    }
    std::sort(ts.begin(), ts.end());
    const double bytes = (double)cus * WAVES * DEPTH * iters * 1024.0;
    const double bpc = bytes / (ts[2] * 1e-3) / 2.4e9 / cus;
    printf("%s  span %7u KB  %2d waves x %2d rows: %6.1f B/clk/CU  (%5.2f per wave)\n", MODE ? "vgpr" : "dma ",
           span_kb, WAVES, DEPTH, bpc, bpc / WAVES);
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    return 0;
}

int main()
{
    int cus = 0;
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    const size_t bytes = (size_t)1 << 30;
    float *src = nullptr, *sink = nullptr;
    CHECK(hipMalloc(&src, bytes));
    CHECK(hipMemset(src, 0, bytes));
    CHECK(hipMalloc(&sink, (size_t)cus * 1024 * sizeof(float)));
    printf("%d CUs\n", cus);
    for (uint32_t span : {1024u, 131072u, 1048575u}) {  // KB: L2, Infinity Cache, HBM
        int rc = 0;
        rc |= run<1, 8, 0>(src, span, sink, cus);
        rc |= run<1, 8, 1>(src, span, sink, cus);
        rc |= run<1, 16, 0>(src, span, sink, cus);
        rc |= run<4, 8, 0>(src, span, sink, cus);
        rc |= run<4, 8, 1>(src, span, sink, cus);
        rc |= run<8, 8, 0>(src, span, sink, cus);
        rc |= run<8, 8, 1>(src, span, sink, cus);
        rc |= run<16, 4, 0>(src, span, sink, cus);
        rc |= run<16, 4, 1>(src, span, sink, cus);
        rc |= run<16, 8, 0>(src, span, sink, cus);
        if (rc) return 1;
    }
    CHECK(hipFree(src));
    CHECK(hipFree(sink));
    return 0;
}
