// Forward projection y (+)= A x in z_order 0 (reference: sp_forward_z0.cuh:107-185).
//
// The reference walks each ray through a texture and evaluates the footprint
// per visit.  Here A is evaluated as the exact transpose of the voxel-driven
// footprint (the same per-voxel r and weights as the backprojector), gathered
// into output-stationary accumulators:
//   sweep wave = one angle x P consecutive detector pixels x 64*HL H slices
//                (lane = HL consecutive slices; variant 32: two angles x 32
//                slices, one per half-wave);
//   rows  = lines of the volume along the axis whose direction cosine is
//           larger (|kappa| >= 1/sqrt 2), so consecutive voxels of a line
//           are at least 1/sqrt 2 pixel apart on the detector and at most
//           two of them share the first pixel l of their footprint;
//   per line, lane q of a staging wave finds the (at most two) voxels whose
//   footprint starts at pixel l = p0-2+q and files them in slot q of the
//   angle's slot table (two angles per staging instruction stream);
//   the angle's sweep wave then runs a statically unrolled sweep over
//   l = p0-2 .. p0+P-1 that adds each voxel to accumulators l, l+1 and, for
//   the one voxel-angle in eight whose footprint covers three pixels, l+2 --
//   so every accumulator index is a compile-time constant and only the voxel
//   index is dynamic.
// No atomics; the summation order is fixed, so results are deterministic.
#include "vp_device.hpp"
#include "vp_launch.hpp"

#include <algorithm>

namespace vp {
namespace {

__device__ __forceinline__ float rdlane(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

constexpr int kSeg = 64;       // max voxels of one wave's line segment (one per lane)
constexpr int kRowsPerFold = 128;

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// One staged piece of work: voxels [j0, j1] of line `row` (-1 = none).
struct Task {
    int row, j0, j1;
};

// Block = NW sweep waves (one angle each) x P detector pixels x 64*HL slices,
// plus NS staging waves.  For every line of the volume along the angles'
// dominant axis, the union of the NW line segments that touch the block's
// pixels is staged in LDS once (HL slices per lane, 256*HL bytes per voxel)
// and read by all NW sweep waves, so each volume load feeds up to NW angles.
// Lines are split into pieces of at most UMAX voxels.  The work of a piece is
// split by role, one barrier per piece:
//   staging waves: issue the piece's global->LDS DMA into the other buffer,
//     build the footprint slot tables of their angles for it in LDS, and
//     find the piece after it (at 4 slices per lane one staging wave finds
//     the pieces for all and builds no tables);
//   sweep waves: read their angle's slot table for the current piece and
//     sweep it (the FMAs).
// So the latency-bound part of a piece (piece iteration, footprints, slot
// tables, the DMA in flight) overlaps the sweep of the previous piece instead
// of preceding it in every wave.  Angles whose dominant axis is x and those
// whose dominant axis is y are handled in two passes over the volume.
//
// With AW = 2 (1 slice per lane only) each sweep wave serves two angles, one
// per half-wave of 32 lanes, so a block covers 2 NW angles x 32 slices: for
// H <= 32, where a wave of 64 one-slice lanes would leave half its lanes
// idle.  Every output still gets the same terms in the same order.
template <int P, int NW, int HL, int AW = 1>
__host__ __device__ inline void forward_grid(int A, int64_t H, int64_t PX, int& nbp, int& nba, int& nbh)
{
    constexpr int SB = 64 * HL / AW;  // slices per block
    nbp = (int)((PX + P - 1) / P);
    nba = (A + NW * AW - 1) / (NW * AW);
    nbh = (int)((H + SB - 1) / SB);
}

// A block of NW sweep waves (AW angles each) and NS staging waves.
template <int NW, int NS, int AW = 1>
struct FwdRoles {
    static_assert(NW + NS <= 16, "at most 1024 threads per block");
    static constexpr int AS = (NW * AW + NS - 1) / NS;  // angles per staging wave
    static constexpr int THREADS = 64 * (NW + NS);
};

template <int P, int NW, int NS, int HL, int UMAX, int AW = 1>
__device__ __forceinline__ void forward_block(const OpArgs& g, const int logical)
{
    const float* __restrict__ x = g.in;
    float* __restrict__ y = g.out;
    const float* __restrict__ angles = g.angles;
    const int A = g.A, SY = g.SY, SX = g.SX, H = g.H, PX = g.PX;
    const double scale = g.scale, inv_scale = g.inv_scale;
    const int accumulate = g.accumulate;
    int nbp, nba, nbh;
    forward_grid<P, NW, HL, AW>(A, g.hn > 0 ? g.hn : H - g.h0, PX, nbp, nba, nbh);
    static_assert((P + 5) * 1.4143 + 4 <= kSeg, "segment may exceed one wave");
    static_assert(AW == 1 || (AW == 2 && HL == 1), "two angles per wave only at 1 slice per lane");
    constexpr int NA = NW * AW;                   // angles per block
    constexpr int AS = FwdRoles<NW, NS, AW>::AS;
    constexpr int SLOTS = P + 2;                 // first pixels p0-2 .. p0+P-1
    static_assert(SLOTS <= 32, "slot masks are 32-bit");
    constexpr int VOX = 64 * HL / AW;            // floats per staged voxel (slices per block)
    // floats per buffer: UMAX voxels rounded up to whole DMA groups (the last
    // DMA instruction of a piece fills a whole group), plus one, so that the
    // sweep's read of a slot's second voxel (one voxel above the first) stays
    // inside the array even where the slot has none
    constexpr int GRP = 1024 / (VOX * 4);  // voxels per 16-byte DMA (a multiple of the dword DMA's)
    constexpr int UCAP = (UMAX + GRP - 1) / GRP * GRP;
    constexpr int STAGE = (UCAP + 1) * VOX;
    // Staging double buffer as two separate arrays: every access names its
    // buffer statically, so the compiler can tell that the sweep's LDS reads
    // of one buffer do not alias the DMA into the other and does not wait for
    // the DMA before them.
    __shared__ float4 s_xa[STAGE / 4];
    __shared__ float4 s_xb[STAGE / 4];
    // Slot tables [buffer][angle][rank][slot]: slot q of rank r holds the r-th
    // voxel whose footprint starts at pixel p0 - 2 + q as (wl, wr, w3, LDS
    // offset of its staged slices); empty slots hold zeros.
    __shared__ float4 s_tab[2][NA][2][SLOTS];
    // piece k in s_rec[k % 2]: (line, first voxel, last voxel, pass); line -1 after the last
    __shared__ int4 s_rec[2];
    // the piece after the one in buffer b (the staging waves' piece finder
    // passes it on): (line, first voxel, last voxel, pass)
    __shared__ int4 s_next[2];
    __shared__ AngleK s_k[NA];
    __shared__ int s_mode[NA];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = uni(tid >> 6);

    const int ba = logical % nba;
    const int bp = (logical / nba) % nbp;
    const int bh = logical / (nba * nbp);

    const int hb = g.h0 + bh * VOX;              // first slice of the block
    const int p0 = bp * P;
    const int pl = p0 - 1;
    const int ph = p0 + P;
    const float xc = 0.5f * (float)(SX - 1);
    const float yc = 0.5f * (float)(SY - 1);
    const bool vec4 = (H & 3) == 0;

    if (tid < NA) {
        const int at = ba * NA + tid;
        const int ac = min(at, A - 1);
        const AngleK kk = make_angle(angles[ac], g.du ? g.du[ac] : 0.0f, 0.5 * (double)(PX - 1), scale, inv_scale);
        s_k[tid] = kk;
        const float cc = kk.chi + kk.clo, ss = kk.shi + kk.slo;
        s_mode[tid] = (at < A) ? (fabsf(cc) >= fabsf(ss) ? 0 : 1) : -1;
    }
    __syncthreads();

    // Staging DMA of a piece, issued by the staging waves: staging wave s
    // issues the instructions s, s + NS, ...; one instruction moves RPI
    // voxels of VOX slices (16 B per lane when H is a multiple of 4; else
    // dword instructions: HL per voxel, or one per 64 / VOX voxels).  Idle
    // lanes past H load a clamped, valid address; their results are never
    // stored.  Staged slot v holds voxel j0 + v (the same order for every
    // angle of the block, whatever its direction along the line).
    constexpr int RPI = 1024 / (VOX * 4);    // voxels per 16-byte-per-lane DMA
    constexpr int LPR = 64 / RPI;            // lanes per voxel
    constexpr int DPI = VOX >= 64 ? 1 : 64 / VOX;  // voxels per dword DMA
    static_assert(UCAP % RPI == 0 && UCAP % DPI == 0, "staging buffer must hold whole DMA groups");
    struct Dma {
        const float* base;
        int64_t st_along;
        int nv, nins;
        bool fast;
        __amdgpu_buffer_rsrc_t r;
    };
    // (line, first voxel, last voxel, pass) of a piece
    auto plan = [&](const int row, const int j0, const int j1, const int pass) {
        Dma d;
        const bool px = pass == 0;
        d.st_along = px ? (int64_t)H : (int64_t)SX * H;
        const int64_t st_row = px ? (int64_t)SX * H : (int64_t)H;
        d.nv = j1 - j0 + 1;
        d.base = x + (int64_t)row * st_row + (int64_t)j0 * d.st_along;
        d.nins = vec4 ? (d.nv + RPI - 1) / RPI : (d.nv + DPI - 1) / DPI;
        // one descriptor for the piece when its extent fits 31 bits
        const int64_t extent = ((int64_t)(d.nv - 1) * d.st_along + H) * 4;
        d.fast = extent < (1ll << 31);
        d.r = dma_rsrc(d.base, (uint32_t)(d.fast ? extent : 0));
        return d;
    };
    auto issue = [&](const Dma& d, float4* dst, const int ins) __attribute__((always_inline)) {
        if (ins >= d.nins) return;
        if (vec4) {
            const int v0 = ins * RPI;
            // lanes past the piece (last instruction, RPI > 1) load its last voxel again
            const int v = min(v0 + lane / LPR, d.nv - 1);
            const int hoff = min(hb + (lane % LPR) * 4, H - 4);
            __attribute__((address_space(3))) void* lds =
                (__attribute__((address_space(3))) void*)&dst[v0 * (VOX / 4)];
            if (d.fast) {
                dma16_to_lds(d.r, (uint32_t)(((int64_t)(v - v0) * d.st_along + hoff) * 4),
                             (uint32_t)((int64_t)v0 * d.st_along * 4), lds);
            } else {
                // lane offsets below (RPI-1) * st_along + H floats, under 2^29
                // (RPI > 1 only for up to 128 slices in single launches;
                // batched launches check it)
                dma_to_lds<16>(d.base + (int64_t)v0 * d.st_along,
                               (uint32_t)(((int64_t)(RPI - 1) * d.st_along + H) * 4),
                               (uint32_t)(((int64_t)(v - v0) * d.st_along + hoff) * 4), lds);
            }
        } else if constexpr (DPI > 1) {
            // lane = (voxel v0 + lane / VOX, slice hb + lane % VOX); offsets
            // below (DPI-1) * st_along + H floats, under 2^29 (DPI > 1 only
            // at up to 32 slices)
            const int v0 = ins * DPI;
            const int v = min(v0 + lane / VOX, d.nv - 1);
            const int hh = min(hb + lane % VOX, H - 1);
            dma_to_lds<4>(d.base + (int64_t)v0 * d.st_along,
                          (uint32_t)(((int64_t)(DPI - 1) * d.st_along + H) * 4),
                          (uint32_t)(((int64_t)(v - v0) * d.st_along + hh) * 4),
                          (__attribute__((address_space(3))) void*)&dst[v0 * (VOX / 4)]);
        } else {
#pragma unroll
            for (int c = 0; c < HL; ++c) {
                // dword instruction c fills floats [64c, 64c+64) of the
                // voxel's row, i.e. slices hb + 64c + lane.
                const int hh = min(hb + 64 * c + lane, H - 1);
                dma_to_lds<4>(d.base + (int64_t)ins * d.st_along, (uint32_t)H * 4, (uint32_t)hh * 4,
                              (__attribute__((address_space(3))) void*)(reinterpret_cast<float*>(&dst[ins * (VOX / 4)]) + 64 * c));
            }
        }
    };

    if (wave < NW) {
        // ---------------------------------------------------------------- sweep
        // lane = (half-wave sub, slice group sl); half-wave sub serves angle
        // wave * AW + sub of the block (AW = 1: the whole wave, sl = lane)
        const int sub = AW > 1 ? lane / VOX : 0;
        const int sl = AW > 1 ? lane % VOX : lane;
        const int a = ba * NA + wave * AW + sub;
        // Register partial sums over at most kRowsPerFold lines of one pass;
        // the owning wave folds them into y (it is the only writer of its
        // outputs), which keeps the FP32 summation two-level without a second
        // register array.
        float part[P][HL];
#pragma unroll
        for (int i = 0; i < P; ++i)
#pragma unroll
            for (int j = 0; j < HL; ++j) part[i][j] = 0.0f;
        bool folded = false;
        auto fold = [&]() {
            if (a < A) {
                const int hh0 = hb + sl * HL;
                const bool full = vec4 && hh0 + 3 < H;
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    const int p = p0 + i;
                    if (p < PX) {
                        float* py = y + ((int64_t)a * PX + p) * H + hh0;
                        bool done = false;
                        if constexpr (HL == 4) {
                            if (full) {
                                // one 16-byte read-modify-write per lane: whole cache lines
                                float4 v = make_float4(part[i][0], part[i][1], part[i][2], part[i][3]);
                                if (accumulate || folded) {
                                    const float4 o = *reinterpret_cast<const float4*>(py);
                                    v.x += o.x; v.y += o.y; v.z += o.z; v.w += o.w;
                                }
                                *reinterpret_cast<float4*>(py) = v;
                                done = true;
                            }
                        }
                        if (!done) {
#pragma unroll
                            for (int j = 0; j < HL; ++j) {
                                if (hh0 + j < H) py[j] = (accumulate || folded) ? (py[j] + part[i][j]) : part[i][j];
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < P; ++i)
#pragma unroll
                for (int j = 0; j < HL; ++j) part[i][j] = 0.0f;
            folded = true;
        };
        int fpass = 0, frow = 0;  // fold group (pass, line group) of the partial sums
        // One piece: sweep the piece in buffer b (xcur) with its slot table.
        // Returns false at the end marker.  Inlined at two call sites so that
        // both buffers are compile-time known.
        auto step = [&](const float4* xcur, const int b) __attribute__((always_inline)) -> bool {
            // every LDS read of the piece's header first, so that their
            // latencies overlap: the piece's record and its slot table
            // (unmasked reads; lanes past the slots drop theirs)
            const int4 rc = s_rec[b];
            constexpr int TL = AW > 1 ? 32 : 64;  // lanes per angle of the table reads
            const int tq = lane % TL;            // slot read by the lane
            const int ql = tq < SLOTS ? tq : SLOTS - 1;
            const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
            float4 t0 = s_tab[b][wave * AW + lane / TL][0][ql];
            float4 t1 = s_tab[b][wave * AW + lane / TL][1][ql];
            if (tq >= SLOTS) {
                t0 = zero;
                t1 = zero;
            }
            // the angle adds something from the piece iff a slot holds a voxel
            // (a voxel's first weight wl is positive; empty tables are zeros)
            const bool swept = __ballot(t0.x != 0.f) != 0ull;
            const int row = uni(rc.x);
            if (row < 0) return false;
            const int pass = uni(rc.w);
            if (pass != fpass || row / kRowsPerFold != frow / kRowsPerFold) {
                fold();
                fpass = pass;
                frow = row;
            }
            if (swept) {
                using Mask = uint32_t;
                // per slot of either angle of the wave (AW = 2: the two halves'
                // bits merged; a lane whose angle lacks the term applies a
                // zero weight)
                auto slot_bits = [](const uint64_t m) -> Mask {
                    if constexpr (AW > 1) return (Mask)m | (Mask)(m >> 32);
                    else return (Mask)m;
                };
                // bit q: slot q holds a second voxel (about 1 in 10 at 512^3 x
                // 192); a voxel's first weight wl is always positive
                const Mask has2 = slot_bits(__ballot(t1.x != 0.f));
                // bit q: the rank-0 / rank-1 voxel of slot q covers three pixels
                const Mask has3a = slot_bits(__ballot(t0.z != 0.f));
                const Mask has3b = slot_bits(__ballot(t1.z != 0.f));
                const Mask extra = has3a | has2;  // has3b implies has2
                const float* sx = reinterpret_cast<const float*>(xcur) + sl * HL;
                auto ld = [&](const float* src) {
                    struct V { float a[HL]; } out;
                    if constexpr (HL == 4) {
                        const float4 t4 = *reinterpret_cast<const float4*>(src);
                        out.a[0] = t4.x;
                        out.a[1] = t4.y;
                        out.a[2] = t4.z;
                        out.a[3] = t4.w;
                    } else if constexpr (HL == 2) {
                        const float2 t2 = *reinterpret_cast<const float2*>(src);
                        out.a[0] = t2.x;
                        out.a[1] = t2.y;
                    } else {
#pragma unroll
                        for (int j = 0; j < HL; ++j) out.a[j] = src[j];
                    }
                    return out;
                };
                // Software-pipelined sweep: the two slot rows of pixel q+PF are
                // fetched (unconditionally) before the FMAs of pixel q (PF = 2
                // at 1 slice per lane, whose slots are too short to cover the
                // LDS latency one slot ahead: 2-4 % faster; 1 otherwise, within
                // the register budget).  The second voxel is the next staged
                // slot, so both reads share one address and the second takes a
                // constant instruction offset.
                constexpr int PF = HL == 1 ? 2 : 1;  // slots read ahead
                // Weights and row offsets: the rank-0 weights (wl, wr) of
                // pixel q+PF are read PF pixels ahead, as a wave-uniform LDS
                // read: VGPR operands, for the faster all-VGPR form of the FMAs
                // and no v_readlane.  The table's base stays in a VGPR
                // (512^3 x 192: 4.28 -> 4.23 ms; 128^3 0.235 -> 0.220 ms).
                // At 1 slice per lane the whole record (wl, wr, w3, offset) is
                // read that way, PF + 1 pixels ahead, and gives the address of
                // the slot's rows without a v_readlane (256^2 x 32 x 360:
                // 0.936 -> 0.869 ms); at 4 slices per lane that form was 2.5 %
                // slower (512^3 x 192), and the offset comes by v_readlane from
                // the lane's own table entry.
                constexpr bool kRecords = HL == 1;
                static_assert(kRecords || AW == 1, "per-lane records for two angles per wave");
                constexpr int NR = kRecords ? PF + 2 : PF;  // table entries in flight
                const auto wtab = lds_vgpr_base<pair_f32>(&s_tab[b][wave][0][0]);
                // the lane's angle's table (AW = 2: one per half-wave)
                const auto rtab = lds_vgpr_base<quad_f32>(&s_tab[b][wave * AW][0][0]) + sub * 2 * SLOTS;
                using V = decltype(ld(sx));
                V xar[PF], xbr[PF];
                pair_f32 wv[NR];                 // (wl, wr) of the entries in flight
                float offv[kRecords ? NR : 1];   // their row offsets (records only)
                float w3v[kRecords ? NR : 1];    // and third weights
                auto fetch_entry = [&](const int q) {
                    if constexpr (kRecords) {
                        const quad_f32 r = rtab[q];
                        wv[q % NR] = r.xy;
                        w3v[q % NR] = r.z;
                        offv[q % NR] = r.w;
                    } else {
                        wv[q % NR] = wtab[2 * q];  // float4 entries: (wl, wr) at float2 index 2q
                    }
                };
                auto slot_src = [&](const int q) {
                    if constexpr (kRecords) return sx + __float_as_int(offv[q % NR]);
                    else return sx + __float_as_int(rdlane(t0.w, q));
                };
                auto fetch_rows = [&](const int q) {
                    const float* s0 = slot_src(q);
                    xar[q % PF] = ld(s0);
                    xbr[q % PF] = ld(s0 + VOX);
                };
                // (a record before the rows it addresses; the rows before the
                // weights otherwise)
                if constexpr (kRecords) {
#pragma unroll
                    for (int dd = 0; dd <= PF; ++dd)
                        if (dd < SLOTS) fetch_entry(dd);
                }
#pragma unroll
                for (int dd = 0; dd < PF; ++dd) {
                    if (dd < SLOTS) {
                        fetch_rows(dd);
                        if constexpr (!kRecords) fetch_entry(dd);
                    }
                }
#pragma unroll
                for (int q = 0; q < SLOTS; ++q) {
                    const auto xa = xar[q % PF];
                    const auto xb = xbr[q % PF];
                    const float wl = wv[q % NR].x, wr = wv[q % NR].y;
                    const float w3a = kRecords ? w3v[q % NR] : 0.f;
                    if constexpr (kRecords) {
                        if (q + PF + 1 < SLOTS) fetch_entry(q + PF + 1);
                        if (q + PF < SLOTS) fetch_rows(q + PF);
                    } else if (q + PF < SLOTS) {
                        fetch_rows(q + PF);
                        fetch_entry(q + PF);
                    }
#pragma unroll
                    for (int j = 0; j < HL; ++j) {
                        if (q - 2 >= 0 && q - 2 < P) part[q - 2][j] = fmaf(wl, xa.a[j], part[q - 2][j]);
                        if (q - 1 >= 0 && q - 1 < P) part[q - 1][j] = fmaf(wr, xa.a[j], part[q - 1][j]);
                    }
                    // one test for the common case of a slot with neither a
                    // three-pixel rank-0 footprint nor a second voxel
                    if ((extra >> q) & 1u) {
                        if (q < P && ((has3a >> q) & 1u)) {
                            const float w3 = kRecords ? w3a : rdlane(t0.z, q);
#pragma unroll
                            for (int j = 0; j < HL; ++j) part[q][j] = fmaf(w3, xa.a[j], part[q][j]);
                        }
                        if ((has2 >> q) & 1u) {
                            // rank 1 of slot q: lane q's table entry, or with
                            // two angles per wave the lane's own angle's entry
                            quad_f32 r1 = {0.f, 0.f, 0.f, 0.f};
                            if constexpr (AW > 1) r1 = rtab[SLOTS + q];
                            const float wl2 = AW > 1 ? (float)r1.x : rdlane(t1.x, q);
                            const float wr2 = AW > 1 ? (float)r1.y : rdlane(t1.y, q);
                            // With two angles per wave a lane whose own angle has
                            // no second voxel here runs this branch for the other
                            // half: its staged row one voxel up may lie past the
                            // piece (stale LDS), so it adds zeros instead of
                            // zero-weighted products of that row.
                            auto xb2 = xb;
                            if constexpr (AW > 1) {
#pragma unroll
                                for (int j = 0; j < HL; ++j) xb2.a[j] = wl2 != 0.f ? xb.a[j] : 0.f;
                            }
#pragma unroll
                            for (int j = 0; j < HL; ++j) {
                                if (q - 2 >= 0 && q - 2 < P) part[q - 2][j] = fmaf(wl2, xb2.a[j], part[q - 2][j]);
                                if (q - 1 >= 0 && q - 1 < P) part[q - 1][j] = fmaf(wr2, xb2.a[j], part[q - 1][j]);
                            }
                            if (q < P && ((has3b >> q) & 1u)) {
                                const float w3 = AW > 1 ? (float)r1.z : rdlane(t1.z, q);
#pragma unroll
                                for (int j = 0; j < HL; ++j) part[q][j] = fmaf(w3, xb2.a[j], part[q][j]);
                            }
                        }
                    }
                }
            }
            __syncthreads();  // the piece's buffer, table and record may be reused
            return true;
        };
        __syncthreads();  // the first piece is staged
        while (step(s_xa, 0) && step(s_xb, 1)) {
        }
        fold();
        return;
    }

    // ---------------------------------------------------------------- staging
    // Staging waves issue first: their work is short latency-bound chains
    // that the sweep waves' FMAs would otherwise delay, holding up the barrier.
    __builtin_amdgcn_s_setprio(1);
    const int s = wave - NW;  // staging wave s
    // Piece finding: at 4 slices per lane, where the staging stream is the
    // critical path of a piece and the other NS - 1 staging waves can build
    // all NA slot tables in as many table rounds as NS waves would (NA = 12,
    // NS = 4: two rounds either way), staging wave 0 alone finds and plans
    // the next piece and passes it on through LDS, and builds no tables;
    // otherwise every staging wave finds the pieces (the same sequence) and
    // builds tables.  (At 1 and 2 slices per lane the sweep is the critical
    // path, and the LDS hand-over only adds latency: 2-3 % slower there.)
    constexpr int AS1 = NS > 1 ? (NA + NS - 2) / (NS - 1) : NA;
    constexpr bool kFinder = HL == 4 && NS > 1 && (AS1 + 1) / 2 == (AS + 1) / 2;
    constexpr int TAS = kFinder ? AS1 : AS;  // angles per table-building wave
    const int tw = kFinder ? s - 1 : s;      // table wave tw builds angles tw*TAS .. tw*TAS+TAS-1
    // Lane w < NA describes angle w of the block (segment union of a line).
    const AngleK kl = s_k[lane < NA ? lane : 0];
    const int model = (lane < NA) ? s_mode[lane] : -1;
    // Slot-table lanes: lane = 32 * half + q builds slot q of angle 2 * pr + half
    // of this wave's TAS angles in table round pr, so one instruction stream
    // serves two angles.
    constexpr int ROUNDS = (TAS + 1) / 2;
    static_assert(SLOTS <= 32, "two angles' slots per wave");
    const int half = lane >> 5;
    const int q = lane & 31;
    AngleK kv[ROUNDS];
    int wv[ROUNDS], modev[ROUNDS];  // block angle of the lane per round (NA: none) and its pass
#pragma unroll
    for (int pr = 0; pr < ROUNDS; ++pr) {
        const int i = 2 * pr + half;
        const int w = tw * TAS + i;
        const bool ok = tw >= 0 && i < TAS && w < NA && q < SLOTS;
        kv[pr] = s_k[ok ? w : 0];
        wv[pr] = ok ? w : NA;
        modev[pr] = ok ? s_mode[w] : -1;
    }

    // Pass state: lines run along x in pass 0 and along y in pass 1.
    int pass = 0;
    bool px_ = true;
    int nrows = SY, nalong = SX;
    float ca = xc, cr = yc;
    float kp_l = 0.f, ik_l = 0.f, ot_l = 0.f, pc_l = 0.f;  // lane w < NA: angle w
    float kapv[ROUNDS], ikapv[ROUNDS], othv[ROUNDS], pcv[ROUNDS];  // the lane's angle per round
    bool minev[ROUNDS];
    auto enter_pass = [&](int p) {
        pass = p;
        px_ = (p == 0);
        nrows = px_ ? SY : SX;
        nalong = px_ ? SX : SY;
        ca = px_ ? xc : yc;
        cr = px_ ? yc : xc;
        const float kcl = kl.chi + kl.clo, ksl = kl.shi + kl.slo;
        kp_l = px_ ? kcl : ksl;
        ot_l = px_ ? ksl : kcl;
        ik_l = 1.0f / kp_l;
        pc_l = kl.pch + kl.plo;  // pc - du of angle w (segment bounds only)
#pragma unroll
        for (int pr = 0; pr < ROUNDS; ++pr) {
            const float kc = kv[pr].chi + kv[pr].clo, ks = kv[pr].shi + kv[pr].slo;
            kapv[pr] = px_ ? kc : ks;
            othv[pr] = px_ ? ks : kc;
            ikapv[pr] = 1.0f / kapv[pr];
            pcv[pr] = kv[pr].pch + kv[pr].plo;
            minev[pr] = modev[pr] == p;
        }
    };

    // Segment [jlo, jhi] of line R (centred row coordinate) for an angle
    // (kap, oth): voxels whose centre pixel can lie in [pl, ph], with one
    // pixel of margin each side.
    auto segment = [&](float R, float kp, float ikp, float ot, float pcw, int& jlo, int& jhi) {
        const float ub = fmaf(-ca, kp, fmaf(R, ot, pcw));
        const float ja = ((float)pl - 1.5f - ub) * ikp;
        const float jb = ((float)ph + 1.5f - ub) * ikp;
        jlo = max(0, (int)floorf(fminf(ja, jb)) - 1);
        jhi = min(nalong - 1, (int)ceilf(fmaxf(ja, jb)) + 1);
    };
    // Line table for 64 consecutive lines c0 .. c0+63, lane = line: the union
    // of the block's segments (what is staged); `nonempty` marks lines with
    // anything to stage.
    int c0 = -64;
    int u_lo = 0, u_hi = -1;
    uint64_t nonempty = 0;
    auto load_lines = [&](int base) {
        c0 = base;
        const int row = base + lane;
        const float R = (float)row - cr;
        int lo = 0x7fffffff, hi = -1;
        // not unrolled: keeps the per-angle reads inside the loop
#pragma unroll 1
        for (int w = 0; w < NA; ++w) {
            if (__builtin_amdgcn_readlane(model, w) == pass) {
                int l, h;
                segment(R, rdlane(kp_l, w), rdlane(ik_l, w), rdlane(ot_l, w), rdlane(pc_l, w), l, h);
                if (h >= l) {
                    lo = min(lo, l);
                    hi = max(hi, h);
                }
            }
        }
        const bool valid = row < nrows;
        u_lo = lo;
        u_hi = valid ? hi : -1;
        nonempty = __ballot(valid && hi >= lo);
    };
    // First piece at or after line `row` of the current pass (pieces hold at
    // most UMAX voxels).
    auto first_task = [&](int row) {
        Task t{-1, 0, -1};
        while (row < nrows) {
            if (row >= c0 + 64) load_lines(row);
            const uint64_t m = nonempty & (~0ull << (row - c0));
            if (m) {
                row = c0 + (int)__builtin_ctzll(m);
                const int ulo = uni(__builtin_amdgcn_readlane(u_lo, row - c0));
                const int uhi = uni(__builtin_amdgcn_readlane(u_hi, row - c0));
                t = Task{row, ulo, min(uhi, ulo + UMAX - 1)};
                break;
            }
            row = c0 + 64;
        }
        return t;
    };
    // First piece of pass p or of a later pass (row -1 when there is none).
    auto first_from_pass = [&](int p) {
        for (; p < 2; ++p) {
            if (__ballot(model == p) == 0ull) continue;  // no angle of the block in pass p
            enter_pass(p);
            c0 = -64;
            const Task t = first_task(0);
            if (t.row >= 0) return t;
        }
        pass = 2;
        return Task{-1, 0, -1};
    };
    // The line of `t` is always in the current table (lines only advance).
    auto next_piece = [&](const Task& t) {
        const int uhi = uni(__builtin_amdgcn_readlane(u_hi, t.row - c0));
        if (t.j1 < uhi) return Task{t.row, t.j1 + 1, min(uhi, t.j1 + UMAX)};
        const Task n = first_task(t.row + 1);
        if (n.row >= 0) return n;
        return first_from_pass(pass + 1);
    };

    // Slot tables of this wave's angles for piece t, in buffer b.  Slot q of
    // an angle lists the voxels of the piece whose footprint starts at pixel
    // l = p0 - 2 + q: (wl, wr, w3, LDS offset of its staged slices) for the
    // first staged (rank 0), (wl, wr, w3, 0) for the one after it (rank 1:
    // its slices follow rank 0's), zeros where there is none.  The first
    // nonzero pixel of a footprint is floor(u - b + 1/2) (u its centre, b its
    // half-width), so those voxels have u in [l + b - 1/2, l + b + 1/2):
    // an interval of at most 1/|kappa| <= sqrt 2 voxels along the line, which
    // holds at most two of them.  The lane solves for that interval (widened
    // by 0.01 voxel against rounding), evaluates the footprints of its two
    // candidate voxels and keeps those whose first pixel is l -- the same
    // footprint function for every lane, so each voxel lands in exactly one
    // slot.  Every slot of every angle of the block is written, so an angle
    // with nothing to add from the piece gets an all-zero table.
    auto tables = [&](const Task& t, const int b) __attribute__((always_inline)) {
        const float R = (float)t.row - cr;
        const int lq = pl - 1 + q;  // pixel of slot q
#pragma unroll
        for (int pr = 0; pr < ROUNDS; ++pr) {
            const AngleK& k = kv[pr];
            const float ub = fmaf(-ca, kapv[pr], fmaf(R, othv[pr], pcv[pr]));  // u at j = 0
            const float ulo = ((float)lq + k.b) - 0.5f;
            const float ja = (ulo - ub) * ikapv[pr];
            const float jb = ((ulo + 1.0f) - ub) * ikapv[pr];
            // clamped before the conversion (lanes of angles outside this pass)
            const float jf = fminf(fmaxf(ceilf(fminf(ja, jb) - 0.01f), -2.0f), (float)nalong + 2.0f);
            const int j0 = (int)jf;
            Taps2 c[2];
            bool in[2];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int j = j0 + e;
                const float jc = (float)j;
                c[e] = anchor_left(voxel_taps(px_ ? jc - xc : R, px_ ? R : jc - yc, k));
                in[e] = minev[pr] && j >= t.j0 && j <= t.j1 && c[e].l == lq;
            }
            if (wv[pr] < NA) {
                const bool any = in[0] || in[1];
                const Taps2 r0 = in[0] ? c[0] : c[1];
                const int off = (j0 + (in[0] ? 0 : 1) - t.j0) * VOX;
                const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
                s_tab[b][wv[pr]][0][q] = any ? make_float4(r0.wl, r0.wr, r0.w3, __builtin_bit_cast(float, off)) : zero;
                s_tab[b][wv[pr]][1][q] = (in[0] && in[1]) ? make_float4(c[1].wl, c[1].wr, c[1].w3, 0.f) : zero;
            }
        }
    };
    // A piece with its DMA plan (planned when the piece is found).
    struct Planned {
        Task t;
        Dma d;
    };
    auto with_plan = [&](const Task& t) {
        Planned r;
        r.t = t;
        if (t.row >= 0) r.d = plan(t.row, t.j0, t.j1, pass);
        return r;
    };
    // Publish piece pt.t (or the end marker) in buffer b: issue its DMA,
    // then, at the base issue priority so that the sweep waves' FMAs go
    // first, build its slot tables and find and plan the next piece (work on
    // the wave's own registers) while the DMA lands; wait for the DMA and the
    // table writes, barrier, and return the next piece.  Measured at
    // 512^3 x 192 (A): DMA split in three parts around the table rounds and
    // the next piece found after the barrier 4.73 ms; tables, DMA, then the
    // next piece, all at the staging priority, 4.46 ms; this order 4.28 ms
    // (and 3-8 % faster at 128^3 - 512^2 x 128 x 720).  Results are the same
    // in every order.
    auto publish = [&](const Planned& pt, float4* xdst, const int b) __attribute__((always_inline)) {
        const Task& t = pt.t;
        if (t.row >= 0) {
            const Dma& d = pt.d;
            const int nmine = d.nins > s ? (d.nins - s + NS - 1) / NS : 0;  // this wave's instructions
            for (int k = 0; k < nmine; ++k) issue(d, xdst, s + k * NS);
            __builtin_amdgcn_s_setprio(0);  // the tables and the next piece: after the sweep's issue
            if (tw >= 0) tables(t, b);
        }
        if (s == 0 && lane == 0) s_rec[b] = make_int4(t.row, t.j0, t.j1, pass);
        Planned nx;
        if (!kFinder || s == 0) {
            // (next_piece may enter the next pass: after the record of t)
            nx = with_plan(t.row >= 0 ? next_piece(t) : t);
            if (kFinder && lane == 0) s_next[b] = make_int4(nx.t.row, nx.t.j0, nx.t.j1, pass);
        }
        __builtin_amdgcn_s_waitcnt(0);  // this wave's DMA and table writes have landed
        __syncthreads();
        __builtin_amdgcn_s_setprio(1);
        if (kFinder && s != 0) {
            // the piece wave 0 found, and its pass
            const int4 nr = s_next[b];
            const Task nt{uni(nr.x), uni(nr.y), uni(nr.z)};
            if (nt.row >= 0 && uni(nr.w) != pass) enter_pass(uni(nr.w));
            nx = with_plan(nt);
        }
        return nx;
    };
    Planned pt = with_plan(first_from_pass(0));
    Planned nx = publish(pt, s_xa, 0);
    while (pt.t.row >= 0) {
        pt = nx;
        nx = publish(pt, s_xb, 1);
        if (pt.t.row < 0) break;
        pt = nx;
        nx = publish(pt, s_xa, 0);
    }
}

// Block order: consecutive workgroups (dispatched round-robin over the 8
// XCDs) are consecutive angle groups of one detector strip and slice chunk,
// so all XCDs stream the same volume strip at the same time and it is
// served from the shared Infinity Cache.  (Giving each XCD a contiguous
// range of logical blocks instead was measured 25 % slower.)
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

template <int P, int NW, int NS, int HL, int UMAX, int WPE, int AW>
__global__ __launch_bounds__((FwdRoles<NW, NS>::THREADS), WPE) void forward_z0_kernel(OpArgs g)
{
    forward_block<P, NW, NS, HL, UMAX, AW>(g, xcd_logical(blockIdx.x, gridDim.x));
}

// Persistent launch over a list of instances: block b handles work items
// b, b + gridDim.x, ... of the concatenated (instance, block) list, in a
// fixed order, so results do not depend on the grid or on the other
// instances of the batch.
template <int P, int NW, int NS, int HL, int UMAX, int WPE, int AW>
__global__ __launch_bounds__((FwdRoles<NW, NS>::THREADS), WPE) void forward_batched_kernel(const BatchItem* __restrict__ items,
                                                                  int n, int64_t total)
{
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        int i = 0;
        int64_t base = 0;
        for (; i < n - 1; ++i) {
            int nbp, nba, nbh;
            forward_grid<P, NW, HL, AW>(items[i].A, items[i].H, items[i].PX, nbp, nba, nbh);
            const int64_t nbl = (int64_t)nbp * nba * nbh;
            if (w < base + nbl) break;
            base += nbl;
        }
        forward_block<P, NW, NS, HL, UMAX, AW>(item_args(items[i]), (int)(w - base));
        __syncthreads();
    }
}

constexpr int kTile = 64;

// One 64 x 64 tile of out[c][r] (+)= in[r][c] for an R x C row-major
// matrix, by a 256-thread workgroup through an LDS tile: 16-byte loads and
// stores on full tiles when both extents are multiples of 4 and both
// pointers 16-byte aligned (an accumulating store loads all its old values
// first), element-wise on ragged edges.  Ends with a barrier, so the tile
// buffer can be reused at once.
__device__ __forceinline__ void transpose_tile(const float* __restrict__ in, float* __restrict__ out,
                                               int64_t R, int64_t C, int accumulate, int64_t tile,
                                               float (&t)[kTile][kTile + 1])
{
    const int64_t tiles_c = (C + kTile - 1) / kTile;
    const bool vec = ((R | C) & 3) == 0 && ((reinterpret_cast<uintptr_t>(in) |
                                             reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // scalar path
    const int lr = threadIdx.x >> 4, lc = (threadIdx.x & 15) * 4;  // vector path
    const int64_t r0 = (tile / tiles_c) * kTile, c0 = (tile % tiles_c) * kTile;
    if (vec && r0 + kTile <= R && c0 + kTile <= C) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = lr + 16 * k;
            const float4 v = *reinterpret_cast<const float4*>(in + (r0 + i) * C + c0 + lc);
            t[i][lc] = v.x;
            t[i][lc + 1] = v.y;
            t[i][lc + 2] = v.z;
            t[i][lc + 3] = v.w;
        }
        __syncthreads();
        float4 o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int j = lr + 16 * k;
            o[k] = make_float4(t[lc][j], t[lc + 1][j], t[lc + 2][j], t[lc + 3][j]);
        }
        if (accumulate) {
            float4 old[4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                old[k] = *reinterpret_cast<const float4*>(out + (c0 + lr + 16 * k) * R + r0 + lc);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                o[k].x += old[k].x;
                o[k].y += old[k].y;
                o[k].z += old[k].z;
                o[k].w += old[k].w;
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            *reinterpret_cast<float4*>(out + (c0 + lr + 16 * k) * R + r0 + lc) = o[k];
    } else {
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
    __syncthreads();  // the tile buffer is reused by the next tile
}

__host__ __device__ inline int64_t transpose_tiles(int64_t R, int64_t C)
{
    return ((R + kTile - 1) / kTile) * ((C + kTile - 1) / kTile);
}

__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ in,
                                                        float* __restrict__ out, int64_t R,
                                                        int64_t C, int accumulate)
{
    __shared__ float t[kTile][kTile + 1];
    const int64_t n = transpose_tiles(R, C);
    for (int64_t tile = blockIdx.x; tile < n; tile += gridDim.x) transpose_tile(in, out, R, C, accumulate, tile, t);
}

// Persistent launch over the concatenated tile lists of n matrices.
__global__ __launch_bounds__(256) void transpose_batched_kernel(const TransposeItem* __restrict__ items,
                                                                int n, int64_t total)
{
    __shared__ float t[kTile][kTile + 1];
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        int i = 0;
        int64_t base = 0;
        for (; i < n - 1; ++i) {
            const int64_t nt = transpose_tiles(items[i].R, items[i].C);
            if (w < base + nt) break;
            base += nt;
        }
        transpose_tile(items[i].in, items[i].out, items[i].R, items[i].C, items[i].accumulate, w - base, t);
    }
}

// Configuration of variant V: HL slices per lane, AW angles per sweep wave,
// P pixels and NW sweep waves per block, NS staging waves, pieces of up to
// UMAX voxels; WPE is the occupancy (waves per SIMD) the register budget
// targets.
template <int V>
struct FwdCfg;
// 1 and 2 slices per lane: 12 angles per block with 28-pixel strips.  Both
// variants share P, NW, NS and UMAX, so they add every output's terms in the
// same order (bitwise-equal results).
template <> struct FwdCfg<1> { static constexpr int HL = 1, AW = 1, P = 28, NW = 12, NS = 4, UMAX = 128, WPE = 4; };
template <> struct FwdCfg<2> { static constexpr int HL = 2, AW = 1, P = 28, NW = 12, NS = 4, UMAX = 128, WPE = 4; };
// 4 slices per lane: 12 angles with 22-pixel strips (88 accumulators within
// the 128 VGPRs of 16 waves per CU; the spills are fold addresses, outside
// the sweep).  Pieces of up to 69 voxels: 140 KB of staging, 18 KB of slot
// tables.
template <> struct FwdCfg<4> { static constexpr int HL = 4, AW = 1, P = 22, NW = 12, NS = 4, UMAX = 69, WPE = 4; };
// The same with 16-, 18- and 20-pixel strips (variants 416, 418, 420): more,
// shorter workgroups, for grids that 22-pixel strips would leave a CU round
// partly empty (forward_strip_variant).  Another strip width stages another
// union of segments per line, so where a line is split into pieces the
// pieces end elsewhere: float32 rounding against variant 4.
template <> struct FwdCfg<416> { static constexpr int HL = 4, AW = 1, P = 16, NW = 12, NS = 4, UMAX = 69, WPE = 4; };
template <> struct FwdCfg<418> { static constexpr int HL = 4, AW = 1, P = 18, NW = 12, NS = 4, UMAX = 69, WPE = 4; };
template <> struct FwdCfg<420> { static constexpr int HL = 4, AW = 1, P = 20, NW = 12, NS = 4, UMAX = 69, WPE = 4; };
// 32 slices per block (variant 32, for H <= 32): 1 slice per lane and two
// angles per sweep wave, 24 angles per block, otherwise as variant 1.  Where
// no line's staged union exceeds UMAX (volumes of at most 128 voxels per
// side) the pieces are whole lines and the results equal variant 1's bit for
// bit; on larger volumes the 24-angle union splits some lines elsewhere, and
// an output whose voxels straddle a split in reverse order adds two of its
// terms in the other order (float32 rounding).  It doubles the useful lanes
// at H <= 32 but halves the grid and doubles each block's slot-table work,
// so it pays on large grids only: batched launches offer it to the
// autotuner; single launches keep variant 1.
template <> struct FwdCfg<32> { static constexpr int HL = 1, AW = 2, P = 28, NW = 12, NS = 4, UMAX = 128, WPE = 4; };
// The last block of a ragged H (at most 128 slices past whole 256-slice
// blocks) in a single launch of a 4-slice variant: 1 or 2 slices per lane
// with that variant's strips, angles and pieces (P, NW, UMAX = 69), so the
// same terms in the same order as the 4-slice kernel over all of H.
template <int TL, int TP>
struct FwdTail { static constexpr int HL = TL, AW = 1, P = TP, NW = 12, NS = 4, UMAX = 69, WPE = 4; };

// Narrower strips for 1 slice per lane, used by single launches of volumes
// of at most UMAX = 128 voxels per side when the default grid would have fewer
// workgroups than the device has CUs: twice the workgroups, each with half the
// pixels and 8 angles, two angles per staging wave, and a register budget for
// six waves per SIMD (one block and part of the next per CU).  Within that size no line is split
// into pieces and all lines fall in one fold group, so every output adds the
// same terms in the same order as with the default strips: results are equal
// bit for bit.  (For larger volumes the strip width changes where a line's staged
// union is split and which pieces pair up around a fold, hence the last bits.)
struct FwdSmall {
    static constexpr int HL = 1, AW = 1, P = 16, NW = 8, NS = 4, UMAX = 128, WPE = 6;
};

template <typename C>
hipError_t forward_single(const OpArgs& g, hipStream_t s)
{
    int nbp, nba, nbh;
    forward_grid<C::P, C::NW, C::HL, C::AW>(g.A, g.hn > 0 ? g.hn : g.H - g.h0, g.PX, nbp, nba, nbh);
    const int64_t nb = (int64_t)nbp * nba * nbh;
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((forward_z0_kernel<C::P, C::NW, C::NS, C::HL, C::UMAX, C::WPE, C::AW>), dim3((unsigned)nb),
                       dim3(FwdRoles<C::NW, C::NS>::THREADS), 0, s, g);
    return hipGetLastError();
}

template <typename C>
hipError_t forward_batched(const BatchItem* host, const BatchItem* dev, int n, int grid,
                           hipStream_t s)
{
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        int nbp, nba, nbh;
        forward_grid<C::P, C::NW, C::HL, C::AW>(host[i].A, host[i].H, host[i].PX, nbp, nba, nbh);
        total += (int64_t)nbp * nba * nbh;
    }
    if (total == 0) return hipSuccess;
    const int64_t gsz = grid > 0 ? std::min<int64_t>(grid, total) : total;
    hipLaunchKernelGGL((forward_batched_kernel<C::P, C::NW, C::NS, C::HL, C::UMAX, C::WPE, C::AW>), dim3((unsigned)gsz),
                       dim3(FwdRoles<C::NW, C::NS>::THREADS), 0, s, dev, n, total);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_forward_z0(const float* x, float* y, const float* angles, const Geom& g,
                             int accumulate, hipStream_t s)
{
    OpArgs a;
    a.in = x;
    a.out = y;
    a.angles = angles;
    a.A = g.A;
    a.SY = (int)g.SY;
    a.SX = (int)g.SX;
    a.H = (int)g.H;
    a.PX = (int)g.PX;
    a.accumulate = accumulate;
    a.out_z2 = 0;
    a.h0 = 0;
    a.hn = 0;
    a.scale = ldexp(1.0, g.bits);
    a.inv_scale = ldexp(1.0, -g.bits);
    a.du = g.du;
    switch (single_variant(0, g.A, g.SY, g.SX, g.H, g.PX)) {
        case 1:
            if (std::max(g.SY, g.SX) <= FwdSmall::UMAX && forward_blocks(1, g.A, g.H, g.PX) < device_cus())
                return forward_single<FwdSmall>(a, s);
            return forward_single<FwdCfg<1>>(a, s);
        case 2: return forward_single<FwdCfg<2>>(a, s);
        default: break;
    }
    // 4 slices per lane.  A ragged last block of at most 128 slices goes to
    // a 1- or 2-slice kernel of the same strips and pieces instead of a
    // 256-slice block with most lanes idle (the same bits).  Its DMA groups
    // several voxels per instruction through 32-bit lane offsets: at most
    // 3 * SX * H + H floats.
    const int v = single_variant(0, g.A, g.SY, g.SX, g.H, g.PX);
    const int tail = (int)(g.H % 256);
    const bool split = g.H > 256 && tail > 0 && tail <= 128 && 3 * g.SX * g.H + g.H < ((int64_t)1 << 29);
    OpArgs main = a, rest = a;
    if (split) {
        main.hn = (int)g.H - tail;
        rest.h0 = (int)g.H - tail;
        rest.hn = tail;
    }
    auto both = [&](auto cfg4, auto tail1, auto tail2) {
        const hipError_t e = forward_single<decltype(cfg4)>(main, s);
        if (e != hipSuccess || !split) return e;
        return tail <= 64 ? forward_single<decltype(tail1)>(rest, s) : forward_single<decltype(tail2)>(rest, s);
    };
    switch (v) {
        case 416: return both(FwdCfg<416>{}, FwdTail<1, 16>{}, FwdTail<2, 16>{});
        case 418: return both(FwdCfg<418>{}, FwdTail<1, 18>{}, FwdTail<2, 18>{});
        case 420: return both(FwdCfg<420>{}, FwdTail<1, 20>{}, FwdTail<2, 20>{});
        default: return both(FwdCfg<4>{}, FwdTail<1, 22>{}, FwdTail<2, 22>{});
    }
}

hipError_t launch_forward_batched(int variant, const BatchItem* items_host,
                                  const BatchItem* items_dev, int n, int grid, hipStream_t s)
{
    switch (variant) {
        case 1: return forward_batched<FwdCfg<1>>(items_host, items_dev, n, grid, s);
        case 2: return forward_batched<FwdCfg<2>>(items_host, items_dev, n, grid, s);
        case 4: return forward_batched<FwdCfg<4>>(items_host, items_dev, n, grid, s);
        case 416: return forward_batched<FwdCfg<416>>(items_host, items_dev, n, grid, s);
        case 418: return forward_batched<FwdCfg<418>>(items_host, items_dev, n, grid, s);
        case 420: return forward_batched<FwdCfg<420>>(items_host, items_dev, n, grid, s);
        case 32: return forward_batched<FwdCfg<32>>(items_host, items_dev, n, grid, s);
        default: return hipErrorInvalidValue;
    }
}

int forward_strip_variant(int A, int64_t H, int64_t PX)
{
    // Measured at 192 angles (one MI355X, 256 CUs): 256^3 22 -> 16 pixels
    // 0.75 -> 0.53 ms, 288^3 22 -> 18 1.40 -> 1.25 ms, 320^3 22 -> 20
    // 1.68 -> 1.64 ms, 384^3 22 -> 16 2.48 -> 2.13 ms, 448^3 22 -> 20
    // 3.30 -> 3.14 ms; 22 pixels kept at 512^3, 512^2 x 256 x 720 and
    // 1024^2 x 512 x 1800, as the rule picks.
    const int64_t cus = device_cus();
    int best = 4;
    int64_t best_cost = -1;
    for (const int v : {4, 420, 418, 416}) {  // ties keep the wider strip
        const int64_t nb = forward_blocks(v, A, H, PX);
        const int P = v == 4 ? 22 : v - 400;
        const int64_t cost = (nb + cus - 1) / cus * (P + 4);
        if (best_cost < 0 || cost < best_cost) {
            best = v;
            best_cost = cost;
        }
    }
    return best;
}

int64_t forward_blocks(int variant, int A, int64_t H, int64_t PX)
{
    int nbp = 0, nba = 0, nbh = 0;
    switch (variant) {
        case 1: forward_grid<FwdCfg<1>::P, FwdCfg<1>::NW, 1>(A, H, PX, nbp, nba, nbh); break;
        case 2: forward_grid<FwdCfg<2>::P, FwdCfg<2>::NW, 2>(A, H, PX, nbp, nba, nbh); break;
        case 4: forward_grid<FwdCfg<4>::P, FwdCfg<4>::NW, 4>(A, H, PX, nbp, nba, nbh); break;
        case 416: forward_grid<FwdCfg<416>::P, FwdCfg<416>::NW, 4>(A, H, PX, nbp, nba, nbh); break;
        case 418: forward_grid<FwdCfg<418>::P, FwdCfg<418>::NW, 4>(A, H, PX, nbp, nba, nbh); break;
        case 420: forward_grid<FwdCfg<420>::P, FwdCfg<420>::NW, 4>(A, H, PX, nbp, nba, nbh); break;
        case 32: forward_grid<FwdCfg<32>::P, FwdCfg<32>::NW, 1, 2>(A, H, PX, nbp, nba, nbh); break;
        default: return -1;
    }
    return (int64_t)nbp * nba * nbh;
}

hipError_t launch_transpose(const float* in, float* out, int64_t R, int64_t C, int accumulate,
                            hipStream_t s)
{
    // workgroups stride over the tiles (one-dimensional grid: no limit on
    // either extent's tile count)
    const int64_t tiles = transpose_tiles(R, C);
    if (tiles == 0) return hipSuccess;
    const unsigned grid = (unsigned)std::min<int64_t>(tiles, 8192);
    hipLaunchKernelGGL(transpose_kernel, dim3(grid), dim3(256), 0, s, in, out, R, C, accumulate);
    return hipGetLastError();
}

hipError_t launch_transpose_batched(const TransposeItem* host, const TransposeItem* dev, int n,
                                    int grid, hipStream_t s)
{
    int64_t total = 0;
    for (int i = 0; i < n; ++i) total += transpose_tiles(host[i].R, host[i].C);
    if (total == 0) return hipSuccess;
    const unsigned g = (unsigned)std::min<int64_t>(grid > 0 ? grid : total, std::min<int64_t>(total, 8192));
    hipLaunchKernelGGL(transpose_batched_kernel, dim3(g), dim3(256), 0, s, dev, n, total);
    return hipGetLastError();
}

}  // namespace vp
