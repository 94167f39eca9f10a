// Voxel-driven backprojection x (+)= AT y (reference: sp_backward_z0.cuh:8-100,
// sp_backward_z2.cuh:8-101), gather form, no atomics, deterministic.
//
// Work decomposition (z_order 0: H fastest)
//   block  = TX x TY voxel tile x 64*HL consecutive slices; NW waves
//            (configurations in BwdCfg below);
//   wave   = VPW voxels of the tile; lane = HL consecutive slices, so every
//            footprint weight is wave-uniform and each sinogram read is one
//            contiguous 256*HL-byte row;
//   angles = chunks of AC.  The sinogram window that the tile touches
//            (WIN pixels x 64*HL slices per angle) is staged in LDS by
//            global->LDS DMA, double-buffered: chunk i+1 streams in while
//            chunk i is consumed, one barrier per chunk.  Any number of
//            angles works (no 512-angle shared-memory table).
// Footprints: lanes evaluate (voxel, angle) pairs in parallel -- VPW voxels x
// 64/VPW angles per instruction -- and the voxel loop picks the weights up
// with v_readlane into SGPRs.  Taps that fall off the detector get weight 0
// (the reference reads 0 there through border addressing), so the DMA may
// fetch a clamped, valid pixel row instead.
// Summation: per chunk in registers, then into a second register set, so the
// FP32 sum over angles is two-level.
#include "vp_device.hpp"
#include "vp_launch.hpp"

#include <algorithm>
#include <atomic>

namespace vp {
namespace {


__device__ __forceinline__ float rdlane(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ int rdlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Copy of a wave-uniform value (an SGPR) into a VGPR.  FP32 FMAs with all
// operands in VGPRs issue about 1.5x faster than with an SGPR multiplier, so a
// weight that feeds four FMAs pays for its copy.  (A register move; the
// compiler would otherwise keep the SGPR operand.)
__device__ __forceinline__ float to_vgpr(float s)
{
    float r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(s));
    return r;
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <int TX, int TY, int HL>
__host__ __device__ inline void backward_grid(int64_t SY, int64_t SX, int64_t H, int& nbx, int& nby,
                                              int& nbh)
{
    nbx = (int)((SX + TX - 1) / TX);
    nby = (int)((SY + TY - 1) / TY);
    nbh = (int)((H + 64 * HL - 1) / (64 * HL));
}

// floor(sqrt(n)) for compile-time checks.
constexpr int isqrt(int n)
{
    int r = 0;
    while ((r + 1) * (r + 1) <= n) ++r;
    return r;
}

template <int TX, int TY, int NW, int AC, int WIN, int HL>
__device__ __forceinline__ void backproject_block(const OpArgs& g, const int logical)
{
    const float* __restrict__ y = g.in;
    float* __restrict__ x = g.out;
    const float* __restrict__ angles = g.angles;
    const int A = g.A, SY = g.SY, SX = g.SX, H = g.H, PX = g.PX;
    const double scale = g.scale, inv_scale = g.inv_scale;
    const int accumulate = g.accumulate;
    int nbx, nby, nbh_;
    backward_grid<TX, TY, HL>(SY, SX, H, nbx, nby, nbh_);
    constexpr int VPW = TX * TY / NW;        // voxels per wave
    constexpr int APL = 64 / VPW;            // angles evaluated per lane sweep
    constexpr int VOX = 64 * HL;             // floats per staged pixel row
    constexpr int STAGE = AC * WIN * VOX;    // floats per buffer
    static_assert(VPW * APL == 64 && AC % APL == 0, "lane layout");
    // Window coverage: over the tile, r - min r <= floor(L) + 1 with
    // L = sqrt((TX-1)^2 + (TY-1)^2) (the largest |c|(TX-1) + |s|(TY-1)); the
    // window offset r - 1 - rlo (rlo = min r - 2) is then <= floor(L) + 2
    // and the rows read, offset + {0, 1, 2}, <= floor(L) + 4 <= WIN - 1, so
    // the clamp of the offset never acts.
    static_assert(isqrt((TX - 1) * (TX - 1) + (TY - 1) * (TY - 1)) + 5 <= WIN,
                  "sinogram window too small for the tile");
    static_assert(1024 % (VOX * 4) == 0 && (AC * WIN) % (1024 / (VOX * 4)) == 0,
                  "one 16-byte-per-lane DMA instruction must cover whole pixel rows");
    __shared__ float4 s_y4[2][STAGE / 4];
    // Angle constants and window bases for a range of kTabAngles angles,
    // double-buffered (the next range is built while the current one is in
    // use); every thread builds at most one angle, so the FP64 angle math of
    // a range costs one latency and the chunk loop reads only LDS.
    constexpr int kTabAngles = 256;
    static_assert(kTabAngles % AC == 0 && kTabAngles <= 64 * NW, "angle table range");
    __shared__ AngleK s_k[2][kTabAngles];
    __shared__ int s_rlo[2][kTabAngles];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = uni(tid >> 6);

    const int bx = logical % nbx;
    const int by = (logical / nbx) % nby;
    const int bhc = logical / (nbx * nby);
    const int tx0 = bx * TX, ty0 = by * TY;
    const int hb = bhc * VOX;
    const float xc = 0.5f * (float)(SX - 1);
    const float yc = 0.5f * (float)(SY - 1);
    const bool vec4 = (H & 3) == 0;

    // Footprint lane layout: voxel (lane % VPW) of this wave, angle lane / VPW.
    const int lv = lane % VPW;
    const int la = lane / VPW;
    const int vi = wave * VPW + lv;
    const float X = (float)(tx0 + vi % TX) - xc;
    const float Y = (float)(ty0 + vi / TX) - yc;

    const int nchunks = (A + AC - 1) / AC;

    // Angle constants and window base of angles [a0, a0 + kTabAngles) into
    // table buffer tb (one angle per thread).
    auto tables = [&](int a0, int tb) {
        if (tid < kTabAngles && a0 + tid < A) {
            const int a = a0 + tid;
            const AngleK k = make_angle(angles[a], g.du ? g.du[a] : 0.0f, 0.5 * (double)(PX - 1), scale, inv_scale);
            s_k[tb][tid] = k;
            const float xa = (float)tx0 - xc;
            const float xb = (float)(min(tx0 + TX, SX) - 1) - xc;
            const float ya = (float)ty0 - yc;
            const float yb = (float)(min(ty0 + TY, SY) - 1) - yc;
            int r = voxel_r(xa, ya, k);
            r = min(r, voxel_r(xb, ya, k));
            r = min(r, voxel_r(xa, yb, k));
            r = min(r, voxel_r(xb, yb, k));
            s_rlo[tb][tid] = r - 2;  // one spare pixel below r-1 for rounding
        }
    };
    // Table buffer and entry of angle a (a < A).
    auto tab_of = [&](int a) { return (a / kTabAngles) & 1; };
    auto ent_of = [&](int a) { return a % kTabAngles; };
    // DMA of a chunk's windows: pixel rows [rlo, rlo+WIN) of each angle,
    // clamped to the detector (out-of-detector taps carry zero weight).
    auto fetch = [&](int chunk, int b) {
        float* dst0 = reinterpret_cast<float*>(s_y4[b]);
        if (vec4) {
            constexpr int RPI = 1024 / (VOX * 4);  // pixel rows per DMA instruction
            constexpr int LPR = 64 / RPI;                                  // lanes per row
            const int sub = lane / LPR;
            const int hq = (lane % LPR) * 4;
            const int hoff = min(hb + hq, H - 4);
            for (int ins = wave; ins < AC * WIN / RPI; ins += NW) {
                const int row = ins * RPI + sub;  // = a * WIN + k
                const int a = row / WIN;
                const int kk = row % WIN;
                const int ag = min(chunk * AC + a, A - 1);
                const int p = min(max(s_rlo[tab_of(ag)][ent_of(ag)] + kk, 0), PX - 1);
                __builtin_amdgcn_global_load_lds(
                    (const void*)(y + ((int64_t)ag * PX + p) * H + hoff),
                    (__attribute__((address_space(3))) void*)(dst0 + ins * RPI * VOX), 16, 0, 0);
            }
        } else {
            for (int ins = wave; ins < AC * WIN * HL; ins += NW) {
                const int row = ins / HL;
                const int c = ins % HL;
                const int a = row / WIN;
                const int kk = row % WIN;
                const int ag = min(chunk * AC + a, A - 1);
                const int p = min(max(s_rlo[tab_of(ag)][ent_of(ag)] + kk, 0), PX - 1);
                const int hh = min(hb + 64 * c + lane, H - 1);
                __builtin_amdgcn_global_load_lds(
                    (const void*)(y + ((int64_t)ag * PX + p) * H + hh),
                    (__attribute__((address_space(3))) void*)(dst0 + row * VOX + 64 * c), 4, 0, 0);
            }
        }
    };

    float acc[VPW][HL];
#pragma unroll
    for (int v = 0; v < VPW; ++v)
#pragma unroll
        for (int j = 0; j < HL; ++j) acc[v][j] = 0.0f;

    tables(0, 0);
    __syncthreads();
    fetch(0, 0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    for (int c = 0; c < nchunks; ++c) {
        const int b = c & 1;
        const int a0 = c * AC;  // first angle of the chunk
        const int tb = tab_of(a0);
        if (c + 1 < nchunks) {
            const int an = a0 + AC;
            if (an % kTabAngles == 0) {
                // the next chunk starts a new angle range: build its table
                // (into the other buffer, idle since the previous range)
                tables(an, tab_of(an));
                __syncthreads();
            }
            fetch(c + 1, b ^ 1);
        }

        float part[VPW][HL];
#pragma unroll
        for (int v = 0; v < VPW; ++v)
#pragma unroll
            for (int j = 0; j < HL; ++j) part[v][j] = 0.0f;
        const float* sy = reinterpret_cast<const float*>(s_y4[b]) + lane * HL;

#pragma unroll
        for (int ag = 0; ag < AC; ag += APL) {
            // Footprint of (voxel lv, angle ag + la) in this lane.
            const int al = ag + la;
            const int ae = ent_of(min(a0 + al, A - 1));  // same range: a0 <= A - 1
            const AngleK k = s_k[tb][ae];
            const Taps t = voxel_taps(X, Y, k);
            const bool aval = c * AC + al < A;
            const int p0 = t.r - 1;
            const float wm = (aval && p0 >= 0 && p0 < PX) ? t.wm : 0.0f;
            const float w0 = (aval && p0 + 1 >= 0 && p0 + 1 < PX) ? t.w0 : 0.0f;
            const float wp = (aval && p0 + 2 >= 0 && p0 + 2 < PX) ? t.wp : 0.0f;
            int off = p0 - s_rlo[tb][ae];
            off = min(max(off, 0), WIN - 3);
            const int foff = (al * WIN + off) * VOX;  // LDS float offset of pixel r-1
            // Taps (pixels r-1, r, r+1; HL slices each) of item L of this
            // lane sweep, read from the staged windows.  Software-pipelined:
            // the taps of item L+1 are loaded before the FMAs of item L.
            struct Y3 {
                float t[3][HL];
            };
            auto load_taps = [&](int L) {
                const float* pr = sy + rdlane(foff, L);
                Y3 o;
#pragma unroll
                for (int k3 = 0; k3 < 3; ++k3) {
                    if constexpr (HL == 2) {
                        const float2 v2 = *reinterpret_cast<const float2*>(pr + k3 * VOX);
                        o.t[k3][0] = v2.x;
                        o.t[k3][1] = v2.y;
                    } else {
#pragma unroll
                        for (int j = 0; j < HL; ++j) o.t[k3][j] = pr[k3 * VOX + j];
                    }
                }
                return o;
            };
            Y3 nxt = load_taps(0);
#pragma unroll
            for (int s2 = 0; s2 < APL; ++s2) {
#pragma unroll
                for (int v = 0; v < VPW; ++v) {
                    const int L = s2 * VPW + v;
                    const Y3 cur = nxt;
                    if (L + 1 < APL * VPW) nxt = load_taps(L + 1);
                    const float a0 = rdlane(wm, L);
                    const float a1 = rdlane(w0, L);
                    const float a2 = rdlane(wp, L);
#pragma unroll
                    for (int j = 0; j < HL; ++j) {
                        float s = part[v][j];
                        s = fmaf(a0, cur.t[0][j], s);
                        s = fmaf(a1, cur.t[1][j], s);
                        s = fmaf(a2, cur.t[2][j], s);
                        part[v][j] = s;
                    }
                }
            }
        }
#pragma unroll
        for (int v = 0; v < VPW; ++v)
#pragma unroll
            for (int j = 0; j < HL; ++j) acc[v][j] += part[v][j];

        __builtin_amdgcn_s_waitcnt(0);  // this wave's DMA into the other buffer landed
        __syncthreads();
    }

    // Write-out.  The old values of a batch of voxels are loaded first, all
    // unconditionally (clamped to valid addresses), so their latencies overlap
    // and no store waits on a load of another branch; the stores of a wave
    // whose slices all exist are unmasked.
    constexpr int WB = 8;  // voxels per batch
    static_assert(VPW % WB == 0, "batch");
    const bool hfull = hb + VOX <= H;
#pragma unroll
    for (int v0 = 0; v0 < VPW; v0 += WB) {
        float old[WB][HL];
        int64_t base[WB];
        bool vin[WB];
#pragma unroll
        for (int v = 0; v < WB; ++v) {
            const int wv = wave * VPW + v0 + v;
            const int ix = tx0 + wv % TX;
            const int iy = ty0 + wv / TX;
            vin[v] = ix < SX && iy < SY;
            base[v] = vin[v] ? ((int64_t)iy * SX + ix) * H : 0;
        }
        if (accumulate) {
#pragma unroll
            for (int v = 0; v < WB; ++v)
#pragma unroll
                for (int j = 0; j < HL; ++j) old[v][j] = x[base[v] + min(hb + lane * HL + j, H - 1)];
        } else {
#pragma unroll
            for (int v = 0; v < WB; ++v)
#pragma unroll
                for (int j = 0; j < HL; ++j) old[v][j] = 0.0f;
        }
        if (hfull) {
#pragma unroll
            for (int v = 0; v < WB; ++v) {
                if (vin[v]) {
#pragma unroll
                    for (int j = 0; j < HL; ++j) x[base[v] + hb + lane * HL + j] = old[v][j] + acc[v0 + v][j];
                }
            }
        } else {
#pragma unroll
            for (int v = 0; v < WB; ++v) {
                if (vin[v]) {
#pragma unroll
                    for (int j = 0; j < HL; ++j) {
                        const int hh = hb + lane * HL + j;
                        if (hh < H) x[base[v] + hh] = old[v][j] + acc[v0 + v][j];
                    }
                }
            }
        }
    }
}

// 4 slices per lane.  Block = TX x TY voxel tile x 256 slices in NW waves of
// VPW voxels; lane = 4 consecutive slices, so each tap is one 16-byte LDS
// read per lane (ds_read_b128, per byte the fastest LDS read) and each
// wave-uniform weight feeds 4 FMAs.  Footprints are evaluated once per group
// of APL = 64 / VPW angles (lane = (voxel, angle)), anchored at their first
// nonzero pixel: two taps always, the third (about one voxel-angle in eight
// at 512^3 x 192) behind a wave-uniform bit test of a ballot, which leaves
// every sum unchanged (a skipped tap has weight 0).  Each group is consumed
// over NSUB window chunks of AC angles; the window double buffer is two
// separately declared LDS arrays and the chunk loop is unrolled over the
// group, so every tap read names its buffer statically and the DMA into the
// other buffer does not hold it up.  The DMA rows of a group's chunks are
// located once per group, lane-parallel, so issuing a row costs one readlane
// and scalar address arithmetic.  Sums are two-level: per group of APL angles
// in registers, then across groups.
template <int TX, int TY, int NW, int AC, int WIN>
__host__ __device__ inline void backward4_grid(int64_t SY, int64_t SX, int64_t H, int& nbx, int& nby,
                                               int& nbh)
{
    nbx = (int)((SX + TX - 1) / TX);
    nby = (int)((SY + TY - 1) / TY);
    nbh = (int)((H + 255) / 256);
}

template <int TX, int TY, int NW, int AC, int WIN, int TAB>
__device__ __forceinline__ void backproject4_block(const OpArgs& g, const int logical)
{
    constexpr int HL = 4;
    constexpr int VPW = TX * TY / NW;      // voxels per wave
    constexpr int APL = 64 / VPW;          // angles per footprint group
    constexpr int NSUB = APL / AC;         // window chunks per group
    constexpr int ROW = 64 * HL;           // floats per staged pixel row
    constexpr int STAGE = AC * WIN * ROW;  // floats per window buffer
    constexpr int RPW = (AC * WIN + NW - 1) / NW;  // DMA rows per wave and chunk (at most)
    static_assert(VPW * APL == 64 && APL % AC == 0 && NSUB % 2 == 0, "lane layout");
    static_assert(NSUB * RPW <= 64, "DMA rows of a group must fit one lane each");
    // Window coverage: over the tile, r - min r <= floor(L) + 1 with
    // L = sqrt((TX-1)^2 + (TY-1)^2); the first tap l is r - 1 or r, and the
    // third tap is read only when l = r - 1, so the rows read are at most
    // r + 1 - rlo <= floor(L) + 4 <= WIN - 1 (rlo = min r - 2).
    static_assert(isqrt((TX - 1) * (TX - 1) + (TY - 1) * (TY - 1)) + 5 <= WIN,
                  "sinogram window too small for the tile");
    __shared__ float4 s_wa[STAGE / 4];
    __shared__ float4 s_wb[STAGE / 4];
    constexpr int kTabAngles = TAB;  // angles per table range (double-buffered)
    static_assert(kTabAngles % APL == 0 && kTabAngles <= 64 * NW, "angle table range");
    __shared__ AngleK s_k[2][kTabAngles];
    __shared__ int s_rlo[2][kTabAngles];
    // per wave, lane = voxel-angle of the current group: its two anchored
    // weights, read back wave-uniformly by the tap loop (VGPR operands)
    __shared__ float2 s_wt[NW][64];

    const float* __restrict__ y = g.in;
    float* __restrict__ x = g.out;
    const float* __restrict__ angles = g.angles;
    const int A = g.A, SY = g.SY, SX = g.SX, H = g.H, PX = g.PX;
    const double scale = g.scale, inv_scale = g.inv_scale;
    int nbx, nby, nbh_;
    backward4_grid<TX, TY, NW, AC, WIN>(SY, SX, g.hn > 0 ? g.hn : H - g.h0, nbx, nby, nbh_);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = uni(tid >> 6);
    const int bx = logical % nbx;
    const int by = (logical / nbx) % nby;
    const int bhc = logical / (nbx * nby);
    const int tx0 = bx * TX, ty0 = by * TY;
    const int hb = g.h0 + bhc * ROW;
    const float xc = 0.5f * (float)(SX - 1);
    const float yc = 0.5f * (float)(SY - 1);
    const bool vec4 = (H & 3) == 0;

    // footprint lane layout: voxel lv of this wave, angle la of the group
    const int lv = lane % VPW;
    const int la = lane / VPW;
    const int vi = wave * VPW + lv;
    const float X = (float)(tx0 + vi % TX) - xc;
    const float Y = (float)(ty0 + vi / TX) - yc;

    auto tables = [&](int a0, int tb) {
        if (tid < kTabAngles && a0 + tid < A) {
            const int a = a0 + tid;
            const AngleK k = make_angle(angles[a], g.du ? g.du[a] : 0.0f, 0.5 * (double)(PX - 1), scale, inv_scale);
            s_k[tb][tid] = k;
            const float xa = (float)tx0 - xc;
            const float xb = (float)(min(tx0 + TX, SX) - 1) - xc;
            const float ya = (float)ty0 - yc;
            const float yb = (float)(min(ty0 + TY, SY) - 1) - yc;
            int r = voxel_r(xa, ya, k);
            r = min(r, voxel_r(xb, ya, k));
            r = min(r, voxel_r(xa, yb, k));
            r = min(r, voxel_r(xb, yb, k));
            s_rlo[tb][tid] = r - 2;
        }
    };
    auto tab_of = [&](int a) { return (a / kTabAngles) & 1; };
    auto ent_of = [&](int a) { return a % kTabAngles; };
    // Sinogram row (angle * PX + pixel) of DMA row `row` (= a * WIN + k) of
    // chunk `chunk`: pixel rlo + k of the chunk's angle a, clamped to the
    // detector and to the last angle (taps there carry zero weight).
    auto dma_row = [&](int chunk, int row) {
        const int ag = min(chunk * AC + row / WIN, A - 1);
        const int p = min(max(s_rlo[tab_of(ag)][ent_of(ag)] + row % WIN, 0), PX - 1);
        return (uint32_t)ag * (uint32_t)PX + (uint32_t)p;
    };
    // DMA of one staged row (sinogram row srow) into row `row` of dst; float q
    // of a staged row holds slice hb + q.
    const uint32_t voff4 = (uint32_t)min(hb + lane * 4, H - 4) * 4u;
    auto issue_row = [&](uint32_t srow, int row, float4* dst) __attribute__((always_inline)) {
        const float* base = y + (int64_t)srow * H;
        if (vec4) {
            dma_to_lds<16>(base, (uint32_t)H * 4u, voff4,
                           (__attribute__((address_space(3))) void*)(dst + row * (ROW / 4)));
        } else {
#pragma unroll
            for (int c4 = 0; c4 < HL; ++c4) {
                const uint32_t hh = (uint32_t)min(hb + 64 * c4 + lane, H - 1);
                dma_to_lds<4>(base, (uint32_t)H * 4u, hh * 4u,
                              (__attribute__((address_space(3))) void*)(reinterpret_cast<float*>(dst + row * (ROW / 4)) + 64 * c4));
            }
        }
    };
    // DMA of a whole chunk (the first one; later chunks go through the rows
    // located per group)
    auto fetch = [&](int chunk, float4* dst) {
        for (int row = wave; row < AC * WIN; row += NW) issue_row((uint32_t)uni((int)dma_row(chunk, row)), row, dst);
    };

    float acc[VPW][HL];
#pragma unroll
    for (int v = 0; v < VPW; ++v)
#pragma unroll
        for (int j = 0; j < HL; ++j) acc[v][j] = 0.0f;

    const int nch = (A + AC - 1) / AC;  // chunks holding at least one angle
    const int ngroups = (A + APL - 1) / APL;
    tables(0, 0);
    __syncthreads();
    fetch(0, s_wa);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    for (int gi = 0; gi < ngroups; ++gi) {
        const int a0 = gi * APL;
        const int tb = tab_of(a0);
        // The last group of a table range builds the next range's table first
        // (into the other buffer, idle since the previous range), so that the
        // DMA rows of the next group's first chunk can be located below.
        if ((a0 + APL) % kTabAngles == 0 && a0 + APL < A) {
            tables(a0 + APL, tab_of(a0 + APL));
            __syncthreads();
        }
        // footprint of (voxel lv, angle a0 + la); the group lies in one table
        // range.  This per-group part (a dependent chain ending in the DMA
        // rows of the group's chunks) runs at a raised issue priority, the
        // tap loop at the base one, so a wave entering its group is not held
        // up by the other waves' FMAs (256^3 x 192: 0.45 -> 0.40 ms, 512^3:
        // 3.04 -> 3.01 ms; same results).
        __builtin_amdgcn_s_setprio(1);
        const int al = a0 + la;
        const bool aval = al < A;
        const int ae = ent_of(min(al, A - 1));
        const AngleK k = s_k[tb][ae];
        const Taps2 t = anchor_left(voxel_taps(X, Y, k));
        const float wl = (aval && t.l >= 0 && t.l < PX) ? t.wl : 0.0f;
        const float wr = (aval && t.l + 1 >= 0 && t.l + 1 < PX) ? t.wr : 0.0f;
        const float w3 = (aval && t.l + 2 >= 0 && t.l + 2 < PX) ? t.w3 : 0.0f;
        const uint64_t has3 = __ballot(w3 != 0.0f);
        // (after the previous group's reads of s_wt: one wave's LDS accesses
        // are performed in order)
        s_wt[wave][lane] = make_float2(wl, wr);
        // Window row of pixel l.  For the tile's voxels the clamp never acts:
        // l - rlo <= floor(L) + 3 <= WIN - 2, and <= WIN - 3 when the third tap
        // is used (l = r - 1); it keeps the reads of voxels past the volume's
        // edge (never stored) inside their angle's window.
        int off = t.l - s_rlo[tb][ae];
        off = min(max(off, 0), w3 != 0.0f ? WIN - 3 : WIN - 2);
        const int foff = ((la % AC) * WIN + off) * ROW;  // float offset of pixel l in its buffer
        // DMA rows of the chunks that this group's sub-chunks prefetch (c + 1
        // for c = gi * NSUB + sub): lane sub * RPW + j holds row wave + j * NW.
        uint32_t rows_l = 0;
        {
            const int sub = lane / RPW, j = lane % RPW;
            const int row = wave + j * NW;
            if (lane < NSUB * RPW && row < AC * WIN) rows_l = dma_row(gi * NSUB + sub + 1, row);
        }

        __builtin_amdgcn_s_setprio(0);
        // The group's first item of voxel v (sub 0, i = v; chunk gi * NSUB
        // always holds an angle) sets part[v] to its first product instead of
        // adding it to zeros: the sums differ at most in the sign of a zero,
        // which acc += part removes, so every result is unchanged.
        float part[VPW][HL];

#pragma unroll
        for (int sub = 0; sub < NSUB; ++sub) {
            const int c = gi * NSUB + sub;
            if (c >= nch) break;  // block-uniform
            if (c + 1 < nch) {
                float4* dst = (sub & 1) ? s_wa : s_wb;
#pragma unroll
                for (int j = 0; j < RPW; ++j) {
                    const int row = wave + j * NW;
                    if (j * NW < AC * WIN - NW + 1 || row < AC * WIN)
                        issue_row((uint32_t)__builtin_amdgcn_readlane((int)rows_l, sub * RPW + j), row, dst);
                }
            }
            const float* sw = reinterpret_cast<const float*>((sub & 1) ? s_wb : s_wa) + lane * HL;
            struct T2 {
                float4 t[2];
            };
            auto tap_ptr = [&](int L) { return sw + __builtin_amdgcn_readlane(foff, L); };
            auto load2 = [&](const float* pr) {
                T2 o;
                o.t[0] = *reinterpret_cast<const float4*>(pr);
                o.t[1] = *reinterpret_cast<const float4*>(pr + ROW);
                return o;
            };
            constexpr int NI = AC * VPW;  // voxel-angles of this chunk per wave
            const int L0 = sub * NI;      // lanes L0 .. L0 + NI - 1 hold them
            const float* pn = tap_ptr(L0);
            T2 nxt = load2(pn);
            // the two weights of item L, one item ahead: a wave-uniform LDS
            // read gives them as VGPR operands (faster FMAs, no v_readlane)
            const float2* wt = &s_wt[wave][0];
            float2 cw_n = wt[L0];
#pragma unroll
            for (int i = 0; i < NI; ++i) {
                const int v = i % VPW;
                const int L = L0 + i;
                const float* pc = pn;
                const T2 cur = nxt;
                const float2 cw = cw_n;
                if (i + 1 < NI) {
                    pn = tap_ptr(L + 1);
                    nxt = load2(pn);
                    cw_n = wt[L + 1];
                }
                const float c0 = cw.x;
                const float c1 = cw.y;
                if (sub == 0 && i < VPW) {  // compile-time
                    part[v][0] = c0 * cur.t[0].x;
                    part[v][1] = c0 * cur.t[0].y;
                    part[v][2] = c0 * cur.t[0].z;
                    part[v][3] = c0 * cur.t[0].w;
                } else {
                    part[v][0] = fmaf(c0, cur.t[0].x, part[v][0]);
                    part[v][1] = fmaf(c0, cur.t[0].y, part[v][1]);
                    part[v][2] = fmaf(c0, cur.t[0].z, part[v][2]);
                    part[v][3] = fmaf(c0, cur.t[0].w, part[v][3]);
                }
                part[v][0] = fmaf(c1, cur.t[1].x, part[v][0]);
                part[v][1] = fmaf(c1, cur.t[1].y, part[v][1]);
                part[v][2] = fmaf(c1, cur.t[1].z, part[v][2]);
                part[v][3] = fmaf(c1, cur.t[1].w, part[v][3]);
                if ((has3 >> L) & 1ull) {
                    const float c2 = to_vgpr(rdlane(w3, L));
                    const float4 t3 = *reinterpret_cast<const float4*>(pc + 2 * ROW);
                    part[v][0] = fmaf(c2, t3.x, part[v][0]);
                    part[v][1] = fmaf(c2, t3.y, part[v][1]);
                    part[v][2] = fmaf(c2, t3.z, part[v][2]);
                    part[v][3] = fmaf(c2, t3.w, part[v][3]);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);  // this wave's DMA into the other buffer landed
            __syncthreads();
        }
#pragma unroll
        for (int v = 0; v < VPW; ++v)
#pragma unroll
            for (int j = 0; j < HL; ++j) acc[v][j] += part[v][j];
    }

    const int hh0 = hb + lane * HL;
    if (g.out_z2) {
        // x in (H, SY, SX): the wave's voxels are one tile row of TX
        // consecutive voxels, so per slice a lane adds its VPW sums to one
        // run of consecutive floats: two 16-byte accesses when SX is a
        // multiple of 4 and the row lies inside the volume (tx0 is a multiple
        // of 8).  Old values first, then the stores.
        static_assert(VPW == TX && VPW == 8, "one tile row of 8 voxels per wave");
        const int iy = ty0 + wave;
        if (iy < SY) {
            const int64_t plane = (int64_t)SY * SX;
            float* row = x + (int64_t)iy * SX + tx0;
            if ((SX & 3) == 0 && tx0 + TX <= SX) {
                float4 o[HL][2];
#pragma unroll
                for (int j = 0; j < HL; ++j) {
                    const int64_t off = (int64_t)min(hh0 + j, H - 1) * plane;
#pragma unroll
                    for (int k = 0; k < 2; ++k)
                        o[j][k] = g.accumulate ? *reinterpret_cast<const float4*>(row + off + 4 * k)
                                               : make_float4(0.f, 0.f, 0.f, 0.f);
                }
#pragma unroll
                for (int j = 0; j < HL; ++j) {
                    if (hh0 + j < H) {
                        const int64_t off = (int64_t)(hh0 + j) * plane;
#pragma unroll
                        for (int k = 0; k < 2; ++k)
                            *reinterpret_cast<float4*>(row + off + 4 * k) =
                                make_float4(o[j][k].x + acc[4 * k][j], o[j][k].y + acc[4 * k + 1][j],
                                            o[j][k].z + acc[4 * k + 2][j], o[j][k].w + acc[4 * k + 3][j]);
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < HL; ++j) {
                    if (hh0 + j < H) {
                        float* p = row + (int64_t)(hh0 + j) * plane;
#pragma unroll
                        for (int v = 0; v < VPW; ++v)
                            if (tx0 + v < SX) p[v] = (g.accumulate ? p[v] : 0.0f) + acc[v][j];
                    }
                }
            }
        }
        return;
    }

    // Write-out: the old values of the wave's voxels are loaded first (clamped
    // addresses), then all stores; 16-byte accesses where the lane's 4 slices
    // exist and H is a multiple of 4.
    const bool full = vec4 && hh0 + HL <= H;
    float4 old[VPW];
    int64_t base[VPW];
    bool vin[VPW];
#pragma unroll
    for (int v = 0; v < VPW; ++v) {
        const int wv = wave * VPW + v;
        const int ix = tx0 + wv % TX;
        const int iy = ty0 + wv / TX;
        vin[v] = ix < SX && iy < SY;
        base[v] = vin[v] ? ((int64_t)iy * SX + ix) * H : 0;
    }
    if (full) {
#pragma unroll
        for (int v = 0; v < VPW; ++v)
            old[v] = g.accumulate ? *reinterpret_cast<const float4*>(x + base[v] + hh0)
                                  : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int v = 0; v < VPW; ++v) {
            if (vin[v])
                *reinterpret_cast<float4*>(x + base[v] + hh0) =
                    make_float4(old[v].x + acc[v][0], old[v].y + acc[v][1], old[v].z + acc[v][2],
                                old[v].w + acc[v][3]);
        }
    } else {
#pragma unroll
        for (int v = 0; v < VPW; ++v) {
            if (vin[v]) {
#pragma unroll
                for (int j = 0; j < HL; ++j) {
                    const int hh = hh0 + j;
                    if (hh < H) x[base[v] + hh] = (g.accumulate ? x[base[v] + hh] : 0.0f) + acc[v][j];
                }
            }
        }
    }
}

// Small H: 4 slices per lane, RW lanes per voxel.  A wave's 64 lanes form
// R = 64 / RW rows; row r holds voxel r of the wave's current R voxels, lane
// q of a row slices hb + 4q .. hb + 4q + 3, so a block covers 4 * RW slices
// (32, 64 or 128) with no idle lanes where H fills it.  Everything else is
// the 4-slice kernel's: 8 x 8 tiles, footprints per group of 8 angles
// anchored at their first nonzero pixel, two taps always and the third when
// any row of the item needs it, windows staged by global -> LDS DMA into a
// double buffer, sums in registers per group of 8 angles, then across
// groups -- the same terms in the same order per voxel, so the results equal
// the 4-slice kernel's bit for bit.  The tap weights and window offset of
// each row's voxel-angle come from a per-wave LDS record (one ds_read_b128
// per lane at a constant offset from the lane's row), so the FMAs take VGPR
// operands and no v_readlane is needed.
template <int TX, int TY, int NW, int AC, int WIN, int RW>
__host__ __device__ inline void backward_rows_grid(int64_t SY, int64_t SX, int64_t H, int& nbx, int& nby,
                                                   int& nbh)
{
    nbx = (int)((SX + TX - 1) / TX);
    nby = (int)((SY + TY - 1) / TY);
    nbh = (int)((H + 4 * RW - 1) / (4 * RW));
}

template <int TX, int TY, int NW, int AC, int WIN, int TAB, int RW>
__device__ __forceinline__ void backproject_rows_block(const OpArgs& g, const int logical)
{
    constexpr int R = 64 / RW;             // voxel rows per wave
    constexpr int ROW = 4 * RW;            // slices per block = floats per staged pixel row
    constexpr int VPW = TX * TY / NW;      // voxels per wave
    constexpr int VPL = VPW / R;           // voxels per lane
    constexpr int APL = 64 / VPW;          // angles per footprint group
    constexpr int NSUB = APL / AC;         // window chunks per group
    constexpr int STAGE = AC * WIN * ROW;  // floats per window buffer
    constexpr int NI = AC * VPL;           // tap items per chunk (R voxel-angles each)
    static_assert(RW == 8 || RW == 16 || RW == 32, "rows of 8, 16 or 32 lanes");
    static_assert(VPW % R == 0 && VPW * APL == 64 && APL % AC == 0 && NSUB % 2 == 0, "lane layout");
    static_assert(WIN % R == 0, "a 16-byte DMA instruction stages R whole pixel rows of one angle");
    static_assert((WIN * ROW) % 64 == 0, "a dword DMA instruction stays within one angle's rows");
    // Window coverage as in the 4-slice kernel: the rows read are at most
    // r + 1 - rlo <= floor(L) + 4 <= WIN - 1 (rlo = min r - 2 over the tile).
    static_assert(isqrt((TX - 1) * (TX - 1) + (TY - 1) * (TY - 1)) + 5 <= WIN,
                  "sinogram window too small for the tile");
    __shared__ float4 s_wa[STAGE / 4];
    __shared__ float4 s_wb[STAGE / 4];
    constexpr int kTabAngles = TAB;
    static_assert(kTabAngles % APL == 0 && kTabAngles <= 64 * NW, "angle table range");
    __shared__ AngleK s_k[2][kTabAngles];
    __shared__ int s_rlo[2][kTabAngles];
    // per wave, lane = voxel-angle of the current group: (wl, wr, w3, window offset)
    __shared__ float4 s_wt[NW][64];

    const float* __restrict__ y = g.in;
    float* __restrict__ x = g.out;
    const float* __restrict__ angles = g.angles;
    const int A = g.A, SY = g.SY, SX = g.SX, H = g.H, PX = g.PX;
    const double scale = g.scale, inv_scale = g.inv_scale;
    int nbx, nby, nbh_;
    backward_rows_grid<TX, TY, NW, AC, WIN, RW>(SY, SX, g.hn > 0 ? g.hn : H - g.h0, nbx, nby, nbh_);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = uni(tid >> 6);
    const int bx = logical % nbx;
    const int by = (logical / nbx) % nby;
    const int bhc = logical / (nbx * nby);
    const int tx0 = bx * TX, ty0 = by * TY;
    const int hb = g.h0 + bhc * ROW;
    const float xc = 0.5f * (float)(SX - 1);
    const float yc = 0.5f * (float)(SY - 1);
    const bool vec4 = (H & 3) == 0;
    const int vr = lane / RW;  // this lane's voxel row
    const int q4 = (lane % RW) * 4;  // its first slice, relative to hb

    // footprint lane layout: voxel lv of this wave, angle la of the group
    const int lv = lane % VPW;
    const int la = lane / VPW;
    const int vi = wave * VPW + lv;
    const float X = (float)(tx0 + vi % TX) - xc;
    const float Y = (float)(ty0 + vi / TX) - yc;

    auto tables = [&](int a0, int tb) {
        if (tid < kTabAngles && a0 + tid < A) {
            const int a = a0 + tid;
            const AngleK k = make_angle(angles[a], g.du ? g.du[a] : 0.0f, 0.5 * (double)(PX - 1), scale, inv_scale);
            s_k[tb][tid] = k;
            const float xa = (float)tx0 - xc;
            const float xb = (float)(min(tx0 + TX, SX) - 1) - xc;
            const float ya = (float)ty0 - yc;
            const float yb = (float)(min(ty0 + TY, SY) - 1) - yc;
            int r = voxel_r(xa, ya, k);
            r = min(r, voxel_r(xb, ya, k));
            r = min(r, voxel_r(xa, yb, k));
            r = min(r, voxel_r(xb, yb, k));
            s_rlo[tb][tid] = r - 2;
        }
    };
    auto tab_of = [&](int a) { return (a / kTabAngles) & 1; };
    auto ent_of = [&](int a) { return a % kTabAngles; };
    // DMA of chunk `chunk`'s windows (pixel rows rlo .. rlo + WIN - 1 of each
    // of its AC angles, slices hb .. hb + ROW - 1) into dst.  Each angle's
    // rows are read through a buffer descriptor spanning exactly that angle's
    // sinogram rows, so rows off the detector (and past the last angle) read
    // zeros; their taps carry zero weight.  Lanes past H read the next pixel
    // row's slices (or zeros) and are never stored.
    auto fetch = [&](int chunk, float4* dst) __attribute__((always_inline)) {
        const uint32_t range = (uint32_t)PX * (uint32_t)H * 4u;
        if (vec4) {
            constexpr int NINS = AC * WIN / R;  // 16-byte instructions: R pixel rows each
            for (int ins = wave; ins < NINS; ins += NW) {
                const int row0 = ins * R;          // first staged row: a * WIN + k0
                const int a = row0 / WIN, k0 = row0 % WIN;
                const int ag = min(chunk * AC + a, A - 1);
                const int rlo = s_rlo[tab_of(ag)][ent_of(ag)];
                const int pix = rlo + k0 + vr;
                const uint32_t voff = (uint32_t)((pix * H + hb + q4) * 4);
                dma_to_lds<16>(y + (int64_t)ag * PX * H, range, voff,
                               (__attribute__((address_space(3))) void*)(dst + row0 * (ROW / 4)));
            }
        } else {
            constexpr int NINS = AC * WIN * ROW / 64;  // dword instructions: 64 floats each
            for (int ins = wave; ins < NINS; ins += NW) {
                const int f = ins * 64 + lane;      // float of the chunk's windows
                const int row = f / ROW, col = f % ROW;
                const int a = (ins * 64) / (WIN * ROW), k = row % WIN;
                const int ag = min(chunk * AC + a, A - 1);
                const int rlo = s_rlo[tab_of(ag)][ent_of(ag)];
                const uint32_t voff = (uint32_t)(((rlo + k) * H + hb + col) * 4);
                dma_to_lds<4>(y + (int64_t)ag * PX * H, range, voff,
                              (__attribute__((address_space(3))) void*)(reinterpret_cast<float*>(dst) + ins * 64));
            }
        }
    };

    float acc[VPL][4];
#pragma unroll
    for (int v = 0; v < VPL; ++v)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[v][j] = 0.0f;

    const int nch = (A + AC - 1) / AC;
    const int ngroups = (A + APL - 1) / APL;
    tables(0, 0);
    __syncthreads();
    fetch(0, s_wa);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    const float4* wt_row = &s_wt[wave][vr];  // this lane's row: entry L of the table at wt_row[L - vr]
    for (int gi = 0; gi < ngroups; ++gi) {
        const int a0 = gi * APL;
        const int tb = tab_of(a0);
        if ((a0 + APL) % kTabAngles == 0 && a0 + APL < A) {
            tables(a0 + APL, tab_of(a0 + APL));
            __syncthreads();
        }
        // footprint of (voxel lv, angle a0 + la); the group lies in one table
        // range (at a raised issue priority, as in the 4-slice kernel:
        // 512^2 x 64 x 360: 1.01 -> 0.90 ms, 128^3: 0.116 -> 0.114 ms)
        __builtin_amdgcn_s_setprio(1);
        uint64_t has3;  // bit L: entry L has a third tap
        {
            const int al = a0 + la;
            const bool aval = al < A;
            const int ae = ent_of(min(al, A - 1));
            const AngleK k = s_k[tb][ae];
            const Taps2 t = anchor_left(voxel_taps(X, Y, k));
            const float wl = (aval && t.l >= 0 && t.l < PX) ? t.wl : 0.0f;
            const float wr = (aval && t.l + 1 >= 0 && t.l + 1 < PX) ? t.wr : 0.0f;
            const float w3 = (aval && t.l + 2 >= 0 && t.l + 2 < PX) ? t.w3 : 0.0f;
            int off = t.l - s_rlo[tb][ae];
            off = min(max(off, 0), w3 != 0.0f ? WIN - 3 : WIN - 2);
            const int foff = ((la % AC) * WIN + off) * ROW;  // float offset of pixel l in its buffer
            // (after the previous group's reads of s_wt: one wave's LDS accesses are performed in order)
            s_wt[wave][lane] = make_float4(wl, wr, w3, __builtin_bit_cast(float, foff));
            has3 = __ballot(w3 != 0.0f);
        }

        __builtin_amdgcn_s_setprio(0);
        float part[VPL][4];  // set by the group's first item of each voxel, as in the 4-slice kernel

#pragma unroll
        for (int sub = 0; sub < NSUB; ++sub) {
            const int c = gi * NSUB + sub;
            if (c >= nch) break;  // block-uniform
            if (c + 1 < nch) fetch(c + 1, (sub & 1) ? s_wa : s_wb);
            const float* sw = reinterpret_cast<const float*>((sub & 1) ? s_wb : s_wa) + q4;
            // item i: angle sub * AC + i / VPL of the group, voxels (i % VPL) * R + row
            auto entry = [&](int i) { return wt_row[(sub * AC + i / VPL) * VPW + (i % VPL) * R]; };
            auto taps = [&](const float4& e, float4& t0, float4& t1) {
                const float* pr = sw + __builtin_bit_cast(int, e.w);
                t0 = *reinterpret_cast<const float4*>(pr);
                t1 = *reinterpret_cast<const float4*>(pr + ROW);
            };
            float4 e_n = entry(0), t0_n, t1_n;
            taps(e_n, t0_n, t1_n);
            float4 e_n2 = NI > 1 ? entry(1) : e_n;
#pragma unroll
            for (int i = 0; i < NI; ++i) {
                const int v = i % VPL;
                const float4 e = e_n;
                const float4 t0 = t0_n, t1 = t1_n;
                if (i + 1 < NI) {
                    e_n = e_n2;
                    taps(e_n, t0_n, t1_n);
                    if (i + 2 < NI) e_n2 = entry(i + 2);
                }
                if (sub == 0 && i < VPL) {  // compile-time
                    part[v][0] = e.x * t0.x;
                    part[v][1] = e.x * t0.y;
                    part[v][2] = e.x * t0.z;
                    part[v][3] = e.x * t0.w;
                } else {
                    part[v][0] = fmaf(e.x, t0.x, part[v][0]);
                    part[v][1] = fmaf(e.x, t0.y, part[v][1]);
                    part[v][2] = fmaf(e.x, t0.z, part[v][2]);
                    part[v][3] = fmaf(e.x, t0.w, part[v][3]);
                }
                part[v][0] = fmaf(e.y, t1.x, part[v][0]);
                part[v][1] = fmaf(e.y, t1.y, part[v][1]);
                part[v][2] = fmaf(e.y, t1.z, part[v][2]);
                part[v][3] = fmaf(e.y, t1.w, part[v][3]);
                // the third tap when any row's voxel-angle of the item has one
                const int L0 = (sub * AC + i / VPL) * VPW + v * R;
                if ((has3 >> L0) & ((1ull << R) - 1)) {
                    const float4 t2 = *reinterpret_cast<const float4*>(sw + __builtin_bit_cast(int, e.w) + 2 * ROW);
                    part[v][0] = fmaf(e.z, t2.x, part[v][0]);
                    part[v][1] = fmaf(e.z, t2.y, part[v][1]);
                    part[v][2] = fmaf(e.z, t2.z, part[v][2]);
                    part[v][3] = fmaf(e.z, t2.w, part[v][3]);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);  // this wave's DMA into the other buffer landed
            __syncthreads();
        }
#pragma unroll
        for (int v = 0; v < VPL; ++v)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[v][j] += part[v][j];
    }

    // Write-out: old values first (clamped addresses), then the stores;
    // 16-byte accesses where the lane's 4 slices exist and H is a multiple of 4.
    const int hh0 = hb + q4;
    const bool full = vec4 && hh0 + 4 <= H;
    float4 old[VPL];
    int64_t base[VPL];
    bool vin[VPL];
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
        const int wv = wave * VPW + v * R + vr;
        const int ix = tx0 + wv % TX;
        const int iy = ty0 + wv / TX;
        vin[v] = ix < SX && iy < SY;
        base[v] = vin[v] ? ((int64_t)iy * SX + ix) * H : 0;
    }
    if (full) {
#pragma unroll
        for (int v = 0; v < VPL; ++v)
            old[v] = g.accumulate ? *reinterpret_cast<const float4*>(x + base[v] + hh0)
                                  : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            if (vin[v])
                *reinterpret_cast<float4*>(x + base[v] + hh0) =
                    make_float4(old[v].x + acc[v][0], old[v].y + acc[v][1], old[v].z + acc[v][2],
                                old[v].w + acc[v][3]);
        }
    } else {
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            if (vin[v]) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int hh = hh0 + j;
                    if (hh < H) x[base[v] + hh] = (g.accumulate ? x[base[v] + hh] : 0.0f) + acc[v][j];
                }
            }
        }
    }
}

// Block order: consecutive workgroups (round-robin over the 8 XCDs) are
// neighbouring tiles of one slice chunk, so concurrently running tiles read
// overlapping sinogram rows, served from the shared Infinity Cache.
// (Giving each XCD a contiguous tile range instead was measured slower.)
__device__ __forceinline__ int xcd_logical(int bid, int /*nb*/) { return bid; }

__device__ __forceinline__ OpArgs item_args(const BatchItem& it)
{
    OpArgs g;
    g.in = it.in;
    g.out = it.out;
    g.angles = it.angles;
    g.A = it.A;
    g.SY = it.SY;
    g.SX = it.SX;
    g.H = it.H;
    g.PX = it.PX;
    g.accumulate = it.accumulate;
    g.out_z2 = 0;
    g.h0 = 0;
    g.hn = 0;
    const int bits = hi_bits(it.SY, it.SX, it.PX);
    g.du = nullptr;
    g.scale = ldexp(1.0, bits);
    g.inv_scale = ldexp(1.0, -bits);
    return g;
}

template <int TX, int TY, int NW, int AC, int WIN, int HL>
__global__ __launch_bounds__(64 * NW) void backproject_z0_kernel(OpArgs g)
{
    backproject_block<TX, TY, NW, AC, WIN, HL>(g, xcd_logical(blockIdx.x, gridDim.x));
}

// Persistent launch over a list of instances (fixed work order, see the
// forward kernel of the same name).
template <int TX, int TY, int NW, int AC, int WIN, int HL>
__global__ __launch_bounds__(64 * NW) void backproject_batched_kernel(
    const BatchItem* __restrict__ items, int n, int64_t total)
{
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        int i = 0;
        int64_t base = 0;
        for (; i < n - 1; ++i) {
            int nbx, nby, nbh;
            backward_grid<TX, TY, HL>(items[i].SY, items[i].SX, items[i].H, nbx, nby, nbh);
            const int64_t nbl = (int64_t)nbx * nby * nbh;
            if (w < base + nbl) break;
            base += nbl;
        }
        backproject_block<TX, TY, NW, AC, WIN, HL>(item_args(items[i]), (int)(w - base));
        __syncthreads();
    }
}

// 4 slices per lane: 8 x 8 voxels in 8 waves of 8, windows of 14 pixels in
// chunks of 2 angles, footprints in groups of 8 angles, angle tables per 64
// angles: 62 KB of LDS and at most 128 VGPRs, so two workgroups share a CU
// (4 waves per SIMD) and one computes while the other waits at its chunk
// barrier.
struct Bwd4Cfg {
    static constexpr int TX = 8, TY = 8, NW = 8, AC = 2, WIN = 14, TAB = 64;
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// This order gives XCD j the j-th contiguous range of tiles (a band of whole
// tile rows), so neighbouring tiles, which read overlapping sinogram windows,
// share an L2: 512^3 x 192: 3.40 -> 3.34 ms, 1024^2 x 512 x 1800: 145 -> 129
// ms.  It is a bijection for any grid size (XCDs 0 .. r-1 take one tile more
// when the grid is not a multiple of 8), so it changes no result.
__device__ __forceinline__ int xcd_contiguous(int bid, int nb)
{
    constexpr int kXcds = 8;
    const int xcd = bid % kXcds, k = bid / kXcds;
    const int q = nb / kXcds, r = nb % kXcds;
    return xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

// Work item of persistent workgroup w's current round in a batched launch:
// the items of each round (gridDim.x of them, fewer in the last) are handed
// out in the XCD-contiguous order above, so neighbouring tiles of an
// instance share an L2 as in a single launch.  A bijection on each round.
__device__ __forceinline__ int64_t xcd_round_item(int64_t w, int64_t total)
{
    const int64_t round0 = w - w % gridDim.x;
    const int nb = (int)min<int64_t>(gridDim.x, total - round0);
    return round0 + xcd_contiguous((int)(w - round0), nb);
}

// Register budget for 4 waves per SIMD (128 VGPRs): two 8-wave workgroups
// per CU, which hide each other's chunk barriers.
template <int TX, int TY, int NW, int AC, int WIN, int TAB>
__global__ __launch_bounds__((64 * NW), 4) void backproject4_z0_kernel(OpArgs g)
{
    backproject4_block<TX, TY, NW, AC, WIN, TAB>(g, xcd_contiguous(blockIdx.x, gridDim.x));
}

template <int TX, int TY, int NW, int AC, int WIN, int TAB>
__global__ __launch_bounds__((64 * NW), 4) void backproject4_batched_kernel(
    const BatchItem* __restrict__ items, int n, int64_t total)
{
    for (int64_t w0 = blockIdx.x; w0 < total; w0 += gridDim.x) {
        const int64_t w = xcd_round_item(w0, total);
        int i = 0;
        int64_t base = 0;
        for (; i < n - 1; ++i) {
            int nbx, nby, nbh;
            backward4_grid<TX, TY, NW, AC, WIN>(items[i].SY, items[i].SX, items[i].H, nbx, nby, nbh);
            const int64_t nbl = (int64_t)nbx * nby * nbh;
            if (w < base + nbl) break;
            base += nbl;
        }
        backproject4_block<TX, TY, NW, AC, WIN, TAB>(item_args(items[i]), (int)(w - base));
        __syncthreads();
    }
}

template <int TX, int TY, int NW, int AC, int WIN, int TAB, int RW>
__global__ __launch_bounds__((64 * NW), 4) void backproject_rows_z0_kernel(OpArgs g)
{
    backproject_rows_block<TX, TY, NW, AC, WIN, TAB, RW>(g, xcd_contiguous(blockIdx.x, gridDim.x));
}

template <int TX, int TY, int NW, int AC, int WIN, int TAB, int RW>
__global__ __launch_bounds__((64 * NW), 4) void backproject_rows_batched_kernel(
    const BatchItem* __restrict__ items, int n, int64_t total)
{
    for (int64_t w0 = blockIdx.x; w0 < total; w0 += gridDim.x) {
        const int64_t w = xcd_round_item(w0, total);
        int i = 0;
        int64_t base = 0;
        for (; i < n - 1; ++i) {
            int nbx, nby, nbh;
            backward_rows_grid<TX, TY, NW, AC, WIN, RW>(items[i].SY, items[i].SX, items[i].H, nbx, nby, nbh);
            const int64_t nbl = (int64_t)nbx * nby * nbh;
            if (w < base + nbl) break;
            base += nbl;
        }
        backproject_rows_block<TX, TY, NW, AC, WIN, TAB, RW>(item_args(items[i]), (int)(w - base));
        __syncthreads();
    }
}

// Small-H configurations (block of 4 * RW slices): the 4-slice kernel's
// 8 x 8 tiles, 8 waves, chunks of 2 angles, footprint groups of 8 angles;
// windows of 16 pixels (a multiple of the R rows one DMA instruction stages)
// for rows of 8 and 16 lanes, 14 for rows of 32.
template <int RW>
struct BwdRowsCfg {
    static constexpr int TX = 8, TY = 8, NW = 8, AC = 2, WIN = (RW == 32 ? 14 : 16), TAB = 64;
};

template <int RW>
hipError_t backward_rows_single(const OpArgs& g, hipStream_t s)
{
    using C = BwdRowsCfg<RW>;
    int nbx, nby, nbh;
    backward_rows_grid<C::TX, C::TY, C::NW, C::AC, C::WIN, RW>(g.SY, g.SX, g.hn > 0 ? g.hn : g.H - g.h0, nbx, nby,
                                                               nbh);
    const int64_t nb = (int64_t)nbx * nby * nbh;
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((backproject_rows_z0_kernel<C::TX, C::TY, C::NW, C::AC, C::WIN, C::TAB, RW>),
                       dim3((unsigned)nb), dim3(64 * C::NW), 0, s, g);
    return hipGetLastError();
}

template <int RW>
hipError_t backward_rows_batched(const BatchItem* host, const BatchItem* dev, int n, int grid, hipStream_t s)
{
    using C = BwdRowsCfg<RW>;
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        int nbx, nby, nbh;
        backward_rows_grid<C::TX, C::TY, C::NW, C::AC, C::WIN, RW>(host[i].SY, host[i].SX, host[i].H, nbx, nby, nbh);
        total += (int64_t)nbx * nby * nbh;
    }
    if (total == 0) return hipSuccess;
    const int64_t gsz = grid > 0 ? std::min<int64_t>(grid, total) : total;
    hipLaunchKernelGGL((backproject_rows_batched_kernel<C::TX, C::TY, C::NW, C::AC, C::WIN, C::TAB, RW>),
                       dim3((unsigned)gsz), dim3(64 * C::NW), 0, s, dev, n, total);
    return hipGetLastError();
}

hipError_t backward4_single(const OpArgs& g, hipStream_t s)
{
    using C = Bwd4Cfg;
    int nbx, nby, nbh;
    backward4_grid<C::TX, C::TY, C::NW, C::AC, C::WIN>(g.SY, g.SX, g.hn > 0 ? g.hn : g.H - g.h0, nbx, nby, nbh);
    const int64_t nb = (int64_t)nbx * nby * nbh;
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((backproject4_z0_kernel<C::TX, C::TY, C::NW, C::AC, C::WIN, C::TAB>),
                       dim3((unsigned)nb), dim3(64 * C::NW), 0, s, g);
    return hipGetLastError();
}

hipError_t backward4_batched(const BatchItem* host, const BatchItem* dev, int n, int grid,
                             hipStream_t s)
{
    using C = Bwd4Cfg;
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        int nbx, nby, nbh;
        backward4_grid<C::TX, C::TY, C::NW, C::AC, C::WIN>(host[i].SY, host[i].SX, host[i].H, nbx, nby, nbh);
        total += (int64_t)nbx * nby * nbh;
    }
    if (total == 0) return hipSuccess;
    const int64_t gsz = grid > 0 ? std::min<int64_t>(grid, total) : total;
    hipLaunchKernelGGL((backproject4_batched_kernel<C::TX, C::TY, C::NW, C::AC, C::WIN, C::TAB>),
                       dim3((unsigned)gsz), dim3(64 * C::NW), 0, s, dev, n, total);
    return hipGetLastError();
}

// Tile configurations.  2 slices per lane: 16 x 16 voxels and 16 waves of 16
// voxels each, so 4 waves share a SIMD (104 VGPRs, 137 KB LDS); measured 5 %
// faster than 16 x 24 voxels with 12 waves of 32 (168 VGPRs) at 512^3 x 192
// and 2 % faster at 1024^2 x 512 x 1800, bitwise-equal results.  1 slice per
// lane keeps 16 x 16 with 8 waves (bigger tiles measured slower).
template <int HL>
struct BwdCfg;
template <> struct BwdCfg<1> { static constexpr int TX = 16, TY = 16, NW = 8, AC = 4, WIN = 28; };
template <> struct BwdCfg<2> { static constexpr int TX = 16, TY = 16, NW = 16, AC = 4, WIN = 28; };

// Smaller tiles for 1 slice per lane, used by single launches when 16 x 16
// tiles would give fewer workgroups than the device has CUs: 16 x 8 voxels in
// 8 waves of 16 (the tile shape does not change any voxel's sum; 128^3 x 192:
// 0.28 ms with 16 x 16 tiles, 0.23 ms with 4 waves of 32, 0.16 ms).
struct BwdSmall {
    static constexpr int TX = 16, TY = 8, NW = 8, AC = 4, WIN = 24;
};

template <typename C, int HL>
hipError_t backward_launch(const OpArgs& g, hipStream_t s)
{
    int nbx, nby, nbh;
    backward_grid<C::TX, C::TY, HL>(g.SY, g.SX, g.H, nbx, nby, nbh);
    const int64_t nb = (int64_t)nbx * nby * nbh;
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((backproject_z0_kernel<C::TX, C::TY, C::NW, C::AC, C::WIN, HL>),
                       dim3((unsigned)nb), dim3(64 * C::NW), 0, s, g);
    return hipGetLastError();
}

template <int HL>
hipError_t backward_single(const OpArgs& g, hipStream_t s)
{
    return backward_launch<BwdCfg<HL>, HL>(g, s);
}

template <int HL>
hipError_t backward_batched(const BatchItem* host, const BatchItem* dev, int n, int grid,
                            hipStream_t s)
{
    using C = BwdCfg<HL>;
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        int nbx, nby, nbh;
        backward_grid<C::TX, C::TY, HL>(host[i].SY, host[i].SX, host[i].H, nbx, nby, nbh);
        total += (int64_t)nbx * nby * nbh;
    }
    if (total == 0) return hipSuccess;
    const int64_t gsz = grid > 0 ? std::min<int64_t>(grid, total) : total;
    hipLaunchKernelGGL((backproject_batched_kernel<C::TX, C::TY, C::NW, C::AC, C::WIN, HL>),
                       dim3((unsigned)gsz), dim3(64 * C::NW), 0, s, dev, n, total);
    return hipGetLastError();
}

}  // namespace

int default_variant(int op, int64_t H)
{
    if (op == 1 && H <= 128) return H <= 32 ? 32 : (H <= 64 ? 64 : 128);  // AT: rows of 8, 16, 32 lanes
    return H > 128 ? 4 : (H > 64 ? 2 : 1);
}

int device_cus()
{
    constexpr int kMaxDevices = 64;
    static std::atomic<int> cached[kMaxDevices];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 256;
    int n = cached[dev].load(std::memory_order_relaxed);
    if (n <= 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cached[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

int single_variant(int op, int A, int64_t SY, int64_t SX, int64_t H, int64_t PX)
{
    const int v = default_variant(op, H);
    // AT's configurations (4 slices per lane, 256 slices per block or rows
    // of 8, 16, 32 lanes per voxel below 129 slices) all add their terms in
    // the same order, and the row configurations fill any volume's lanes
    if (op == 0 && v == 4) return forward_strip_variant(A, H, PX);
    if (op == 1 || v != 2) return v;
    const int cus = device_cus();
    if (op == 0) return forward_blocks(2, A, H, PX) < cus ? 1 : 2;
    // AT: with fewer 2-slice workgroups than CUs, 1 slice per lane.  Else
    // estimated time = rounds of resident workgroups x time per round:
    // 2 slices per lane, one workgroup per CU (LDS); 1 slice per lane, two
    // per CU, each about 22 % less efficient per voxel-slice (measured at
    // 512^3 x 192).  Picks 1 at 128^3 and 192^3 and 2 from 256^3 up, as
    // measured best (at 320^3 the two are within 2 %).
    const int64_t n2 = backward_blocks(2, SY, SX, H), n1 = backward_blocks(1, SY, SX, H);
    if (n2 < cus) return 1;
    const double w2 = (double)BwdCfg<2>::TX * BwdCfg<2>::TY * 128;
    const double w1 = (double)BwdCfg<1>::TX * BwdCfg<1>::TY * 64 / 0.78;
    const double t2 = (double)((n2 + cus - 1) / cus) * w2;
    const double t1 = (double)((n1 + 2 * cus - 1) / (2 * cus)) * 2.0 * w1;
    return t1 < t2 ? 1 : 2;
}

namespace {
OpArgs backward_args(const float* y, float* x, const float* angles, const Geom& g, int accumulate,
                     int out_z2)
{
    OpArgs a;
    a.in = y;
    a.out = x;
    a.angles = angles;
    a.A = g.A;
    a.SY = (int)g.SY;
    a.SX = (int)g.SX;
    a.H = (int)g.H;
    a.PX = (int)g.PX;
    a.accumulate = accumulate;
    a.out_z2 = out_z2;
    a.h0 = 0;
    a.hn = 0;
    a.scale = ldexp(1.0, g.bits);
    a.inv_scale = ldexp(1.0, -g.bits);
    a.du = g.du;
    return a;
}
}  // namespace

bool backward_writes_z2(const Geom& g) { return single_variant(1, g.A, g.SY, g.SX, g.H, g.PX) == 4; }

// Always the 4-slice kernel (also for a slab of 128 slices or fewer, where
// single launches use the row kernels: their sums are bitwise the same).
hipError_t launch_backward_to_z2(const float* y, float* x, const float* angles, const Geom& g,
                                 hipStream_t s)
{
    return backward4_single(backward_args(y, x, angles, g, 1, 1), s);
}

hipError_t launch_backward_z0(const float* y, float* x, const float* angles, const Geom& g,
                              int accumulate, hipStream_t s)
{
    const OpArgs a = backward_args(y, x, angles, g, accumulate, 0);
    const int v = single_variant(1, g.A, g.SY, g.SX, g.H, g.PX);
    if (v == 4) {
        // A ragged last block of at most 128 slices goes to a row kernel of
        // its size instead of a 256-slice block with most lanes idle (the
        // same terms in the same order, so the same bits).  The row kernel
        // addresses an angle's sinogram rows through 32-bit byte offsets:
        // PX * H * 4 must stay below 2^31.
        const int tail = (int)(g.H % 256);
        if (g.H > 256 && tail > 0 && tail <= 128 && g.PX * g.H < ((int64_t)1 << 29)) {
            OpArgs main = a, rest = a;
            main.hn = (int)g.H - tail;
            rest.h0 = (int)g.H - tail;
            rest.hn = tail;
            const hipError_t e = backward4_single(main, s);
            if (e != hipSuccess) return e;
            if (tail <= 32) return backward_rows_single<8>(rest, s);
            if (tail <= 64) return backward_rows_single<16>(rest, s);
            return backward_rows_single<32>(rest, s);
        }
        return backward4_single(a, s);
    }
    if (v == 32) return backward_rows_single<8>(a, s);
    if (v == 64) return backward_rows_single<16>(a, s);
    if (v == 128) return backward_rows_single<32>(a, s);
    if (v == 2) return backward_single<2>(a, s);
    if (backward_blocks(1, g.SY, g.SX, g.H) < device_cus()) return backward_launch<BwdSmall, 1>(a, s);
    return backward_single<1>(a, s);
}

hipError_t launch_backward_batched(int variant, const BatchItem* items_host,
                                   const BatchItem* items_dev, int n, int grid, hipStream_t s)
{
    switch (variant) {
        case 1: return backward_batched<1>(items_host, items_dev, n, grid, s);
        case 2: return backward_batched<2>(items_host, items_dev, n, grid, s);
        case 4: return backward4_batched(items_host, items_dev, n, grid, s);
        case 32: return backward_rows_batched<8>(items_host, items_dev, n, grid, s);
        case 64: return backward_rows_batched<16>(items_host, items_dev, n, grid, s);
        case 128: return backward_rows_batched<32>(items_host, items_dev, n, grid, s);
        default: return hipErrorInvalidValue;
    }
}

int64_t backward_blocks(int variant, int64_t SY, int64_t SX, int64_t H)
{
    int nbx = 0, nby = 0, nbh = 0;
    switch (variant) {
        case 1: backward_grid<BwdCfg<1>::TX, BwdCfg<1>::TY, 1>(SY, SX, H, nbx, nby, nbh); break;
        case 2: backward_grid<BwdCfg<2>::TX, BwdCfg<2>::TY, 2>(SY, SX, H, nbx, nby, nbh); break;
        case 4: backward4_grid<Bwd4Cfg::TX, Bwd4Cfg::TY, Bwd4Cfg::NW, Bwd4Cfg::AC, Bwd4Cfg::WIN>(SY, SX, H, nbx, nby, nbh); break;
        case 32: backward_rows_grid<8, 8, 8, 2, 16, 8>(SY, SX, H, nbx, nby, nbh); break;
        case 64: backward_rows_grid<8, 8, 8, 2, 16, 16>(SY, SX, H, nbx, nby, nbh); break;
        case 128: backward_rows_grid<8, 8, 8, 2, 14, 32>(SY, SX, H, nbx, nby, nbh); break;
        default: return -1;
    }
    return (int64_t)nbx * nby * nbh;
}

}  // namespace vp
