// Forward projection y (+)= A x in z_order 0 (reference: sp_forward_z0.cuh:107-185).
//
// The reference walks each ray through a texture and evaluates the footprint
// per visit.  Here A is evaluated as the exact transpose of the voxel-driven
// footprint (the same per-voxel r and weights as the backprojector), gathered
// into output-stationary accumulators:
//   wave  = one angle x P consecutive detector pixels x 64 H slices (lane = slice);
//   rows  = lines of the volume along the axis whose direction cosine is
//           larger (|kappa| >= 1/sqrt 2), so consecutive voxels of a line
//           are at least 1/sqrt 2 pixel apart on the detector and at most
//           two of them share the first pixel l of their footprint;
//   per line, lane m evaluates the footprint of the m-th voxel of the line
//   segment that touches the wave's pixels (in increasing-u order), and a
//   statically unrolled sweep over l = p0-2 .. p0+P-1 adds each voxel to
//   accumulators l, l+1 and, for the one voxel-angle in eight whose
//   footprint covers three pixels, l+2 -- so every accumulator index is a
//   compile-time constant and only the voxel index is dynamic.
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
__device__ __forceinline__ float uni(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// One staged piece of work: voxels [j0, j1] of line `row` (-1 = none).
struct Task {
    int row, j0, j1;
};

// Footprint slot table of one piece, lane = slot q: t0 a voxel whose
// footprint starts at pixel p0 - 2 + q, as (wl, wr, w3, LDS offset of the
// voxel's staged slices), and t1 the other one (if two do), as (wl, wr, w3,
// 0).  Of two such voxels (neighbours on the line), t0 is the one staged
// first and t1 the next staged voxel, so t1's offset is t0's plus one voxel.
struct Slots {
    float4 t0, t1;
};

// Value of lane - 1 (lane 0: 0), by a DPP wavefront shift.
__device__ __forceinline__ int lane_shift_up(int v)
{
    constexpr int kWaveShr1 = 0x138;
    return __builtin_amdgcn_update_dpp(0, v, kWaveShr1, 0xf, 0xf, false);
}

// Value of lane + 1 (lane 63: 0), by a DPP wavefront shift.
__device__ __forceinline__ int lane_shift_down(int v)
{
    constexpr int kWaveShl1 = 0x130;
    return __builtin_amdgcn_update_dpp(0, v, kWaveShl1, 0xf, 0xf, false);
}

// Cross-lane scatter of (a, b, c, o): lane d/4 of the result receives the
// values of the highest-numbered lane whose byte address d points at it;
// lanes that no lane points at receive zeros (ds_permute_b32 semantics).
__device__ __forceinline__ float4 permute4(int d, float a, float b, float c, int o)
{
    float4 r;
    r.x = __builtin_bit_cast(float, __builtin_amdgcn_ds_permute(d, __builtin_bit_cast(int, a)));
    r.y = __builtin_bit_cast(float, __builtin_amdgcn_ds_permute(d, __builtin_bit_cast(int, b)));
    r.z = __builtin_bit_cast(float, __builtin_amdgcn_ds_permute(d, __builtin_bit_cast(int, c)));
    r.w = __builtin_bit_cast(float, __builtin_amdgcn_ds_permute(d, o));
    return r;
}

// The same for (a, b, c) only; the fourth component is zero.
__device__ __forceinline__ float4 permute3(int d, float a, float b, float c)
{
    float4 r;
    r.x = __builtin_bit_cast(float, __builtin_amdgcn_ds_permute(d, __builtin_bit_cast(int, a)));
    r.y = __builtin_bit_cast(float, __builtin_amdgcn_ds_permute(d, __builtin_bit_cast(int, b)));
    r.z = __builtin_bit_cast(float, __builtin_amdgcn_ds_permute(d, __builtin_bit_cast(int, c)));
    r.w = 0.0f;
    return r;
}

// Block = NW angles (one per wave) x P detector pixels x 64*HL slices.
// For every line of the volume along the angles' dominant axis, the union of
// the NW line segments that touch the block's pixels is staged in LDS once
// (HL slices per lane, 256*HL bytes per voxel) and read by all NW waves, so
// each volume load feeds up to NW angles.  Lines are split into pieces of at
// most UMAX voxels; while one piece is swept, the next streams into the other
// LDS buffer by DMA and each wave builds its footprint slot table for it (one
// barrier per piece).  Angles whose dominant axis is x and those whose
// dominant axis is y are handled in two passes over the volume.
template <int P, int NW, int HL>
__host__ __device__ inline void forward_grid(int A, int64_t H, int64_t PX, int& nbp, int& nba, int& nbh)
{
    nbp = (int)((PX + P - 1) / P);
    nba = (A + NW - 1) / NW;
    nbh = (int)((H + 64 * HL - 1) / (64 * HL));
}

template <int P, int NW, int HL, int UMAX>
__device__ __forceinline__ void forward_block(const OpArgs& g, const int logical)
{
    const float* __restrict__ x = g.in;
    float* __restrict__ y = g.out;
    const float* __restrict__ angles = g.angles;
    const int A = g.A, SY = g.SY, SX = g.SX, H = g.H, PX = g.PX;
    const double scale = g.scale, inv_scale = g.inv_scale;
    const int accumulate = g.accumulate;
    int nbp, nba, nbh;
    forward_grid<P, NW, HL>(A, H, PX, nbp, nba, nbh);
    static_assert((P + 5) * 1.4143 + 4 <= kSeg, "segment may exceed one wave");
    constexpr int VOX = 64 * HL;                 // floats per staged voxel
    // floats per buffer: UMAX voxels plus one, so that the sweep's read of a
    // slot's second voxel (one voxel above the first) stays inside the array
    // even where the slot has none
    constexpr int STAGE = (UMAX + 1) * VOX;
    // Staging double buffer as two separate arrays: every access names its
    // buffer statically, so the compiler can tell that the sweep's LDS reads
    // of one buffer do not alias the DMA into the other and does not wait for
    // the DMA before them.
    __shared__ float4 s_xa[STAGE / 4];
    __shared__ float4 s_xb[STAGE / 4];
    __shared__ AngleK s_k[NW];
    __shared__ int s_mode[NW];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = uni(tid >> 6);

    const int ba = logical % nba;
    const int bp = (logical / nba) % nbp;
    const int bh = logical / (nba * nbp);

    const int a = ba * NW + wave;
    const int hb = bh * VOX;                     // first slice of the block
    const int p0 = bp * P;
    const int pl = p0 - 1;
    const int ph = p0 + P;
    const float xc = 0.5f * (float)(SX - 1);
    const float yc = 0.5f * (float)(SY - 1);
    const bool vec4 = (H & 3) == 0;

    if (tid < NW) {
        const int at = ba * NW + tid;
        const int ac = min(at, A - 1);
        const AngleK kk = make_angle(angles[ac], g.du ? g.du[ac] : 0.0f, 0.5 * (double)(PX - 1), scale, inv_scale);
        s_k[tid] = kk;
        const float cc = kk.chi + kk.clo, ss = kk.shi + kk.slo;
        s_mode[tid] = (at < A) ? (fabsf(cc) >= fabsf(ss) ? 0 : 1) : -1;
    }
    __syncthreads();
    AngleK k = s_k[wave];
    k.chi = uni(k.chi); k.clo = uni(k.clo); k.shi = uni(k.shi); k.slo = uni(k.slo);
    k.a = uni(k.a); k.b = uni(k.b); k.h = uni(k.h); k.cd = uni(k.cd);
    k.pch = uni(k.pch); k.plo = uni(k.plo);
    const int mode = uni(s_mode[wave]);
    // Lane w < NW describes angle w of the block.
    const AngleK kl = s_k[lane < NW ? lane : 0];
    const int model = (lane < NW) ? s_mode[lane] : -1;

    // Register partial sums over at most kRowsPerFold lines; the owning wave
    // folds them into y (it is the only writer of its outputs), which keeps the
    // FP32 summation two-level without a second register array.
    float part[P][HL];
#pragma unroll
    for (int i = 0; i < P; ++i)
#pragma unroll
        for (int j = 0; j < HL; ++j) part[i][j] = 0.0f;
    bool folded = false;
    auto fold = [&]() {
        if (a < A) {
            const int hh0 = hb + lane * HL;
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
        // Retire the fold's loads and stores here.  Left in flight, they make
        // the compiler wait on vmcnt before the next sweep's LDS reads, which
        // also waits for the next piece's DMA issued meanwhile (about 1 % of
        // A at 512^3 x 192).
        __builtin_amdgcn_s_waitcnt(0);
    };

    for (int pass = 0; pass < 2; ++pass) {
        const bool any = __ballot(model == pass) != 0ull;
        if (!any) continue;  // block-uniform: every wave sees the same modes
        const bool mine = (mode == pass);
        const bool px_ = (pass == 0);
        const int nrows = px_ ? SY : SX;
        const int nalong = px_ ? SX : SY;
        const float ca = px_ ? xc : yc;
        const float cr = px_ ? yc : xc;
        const int64_t st_along = px_ ? (int64_t)H : (int64_t)SX * H;
        const int64_t st_row = px_ ? (int64_t)SX * H : (int64_t)H;
        const float kap = px_ ? (k.chi + k.clo) : (k.shi + k.slo);
        const float oth = px_ ? (k.shi + k.slo) : (k.chi + k.clo);
        const float ikap = 1.0f / kap;
        // Direction in which the centre pixel grows along a line.
        const bool fwd = kap >= 0.0f;
        // Lane w < NW: the (kap, 1/kap, oth) of angle w of the block, read
        // with v_readlane where the union of the block's segments is formed.
        const float kcl = kl.chi + kl.clo, ksl = kl.shi + kl.slo;
        const float kp_l = px_ ? kcl : ksl;
        const float ot_l = px_ ? ksl : kcl;
        const float ik_l = 1.0f / kp_l;
        const float pc_l = kl.pch + kl.plo;  // pc - du of angle w (segment bounds only)

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
        // Line table for 64 consecutive lines c0 .. c0+63, lane = line: the
        // union of the block's segments (what is staged) and this wave's own
        // segment; `nonempty` marks lines with anything to stage.
        int c0 = -64;
        int u_lo = 0, u_hi = -1, s_lo = 0, s_hi = -1;
        uint64_t nonempty = 0;
        auto load_lines = [&](int base) {
            c0 = base;
            const int row = base + lane;
            const float R = (float)row - cr;
            int lo = 0x7fffffff, hi = -1;
            // not unrolled: keeps the per-angle reads inside the loop (unrolled,
            // they are hoisted to every caller of the line search)
#pragma unroll 1
            for (int w = 0; w < NW; ++w) {
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
            s_lo = 0;
            s_hi = -1;
            if (mine) segment(R, kap, ikap, oth, k.pch + k.plo, s_lo, s_hi);
        };
        // First piece at or after line `row` (pieces hold at most UMAX voxels).
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
        // The line of `t` is always in the current table (lines only advance).
        auto next_task = [&](const Task& t) {
            const int uhi = uni(__builtin_amdgcn_readlane(u_hi, t.row - c0));
            if (t.j1 < uhi) return Task{t.row, t.j1 + 1, min(uhi, t.j1 + UMAX)};
            return first_task(t.row + 1);
        };
        // Stage a piece straight into LDS with global->LDS DMA: one wave
        // instruction moves one voxel's 64*HL slices (16 B per lane when H is a
        // multiple of 4, else HL dword instructions).  Idle lanes past H load a
        // clamped, valid address; their results are never stored.  Staged
        // slot s holds voxel j0 + s (the order is the same for every wave of
        // the block, whatever the direction of its angle).
        auto fetch = [&](const Task& t, float4* dst) __attribute__((always_inline)) {
            const int nv = t.j1 - t.j0 + 1;
            const float* base = x + (int64_t)t.row * st_row + (int64_t)t.j0 * st_along;
            // Buffer-form DMA: each instruction's base is its first voxel;
            // lane offsets stay below (RPI-1) * st_along + H floats, under
            // 2^29 (RPI > 1 only for up to 128 slices in single launches;
            // batched launches check it).
            if (vec4) {
                constexpr int RPI = 1024 / (VOX * 4);  // voxels per 16-byte-per-lane DMA
                constexpr int LPR = 64 / RPI;          // lanes per voxel
                static_assert(UMAX % RPI == 0, "staging buffer must hold whole DMA groups");
                const int sub = lane / LPR;
                const int hoff = min(hb + (lane % LPR) * 4, H - 4);
                const int nins = (nv + RPI - 1) / RPI;
                const int nfull = nv / RPI;  // instructions whose RPI voxels all exist
                const uint32_t range = (uint32_t)(((int64_t)(RPI - 1) * st_along + H) * 4);
                // full instructions: the same lane offsets every time, the base
                // and the LDS destination advance by constant steps
                const uint32_t voff = (uint32_t)(((int64_t)sub * st_along + hoff) * 4);
                const int64_t step = (int64_t)NW * RPI * st_along;
                const float* src = base + (int64_t)wave * RPI * st_along;
                __attribute__((address_space(3))) float4* d =
                    (__attribute__((address_space(3))) float4*)&dst[wave * RPI * (VOX / 4)];
                int ins = wave;
                for (; ins < nfull; ins += NW) {
                    dma_to_lds<16>(src, range, voff, (__attribute__((address_space(3))) void*)d);
                    src += step;
                    d += NW * RPI * (VOX / 4);
                }
                if (ins < nins) {
                    // the last, partial instruction (RPI > 1 only): lanes past
                    // the piece load its last voxel again
                    const int v0 = ins * RPI;
                    const int v = min(v0 + sub, nv - 1);
                    dma_to_lds<16>(base + (int64_t)v0 * st_along, range,
                                   (uint32_t)(((int64_t)(v - v0) * st_along + hoff) * 4),
                                   (__attribute__((address_space(3))) void*)&dst[ins * RPI * (VOX / 4)]);
                }
            } else {
                for (int v = wave; v < nv; v += NW) {
#pragma unroll
                    for (int c = 0; c < HL; ++c) {
                        // dword instruction c fills floats [64c, 64c+64) of the
                        // voxel's row, i.e. slices hb + 64c + lane.
                        const int e = 64 * c + lane;           // float index in the row
                        const int hh = min(hb + e, H - 1);
                        dma_to_lds<4>(base + (int64_t)v * st_along, (uint32_t)H * 4, (uint32_t)hh * 4,
                                      (__attribute__((address_space(3))) void*)(reinterpret_cast<float*>(&dst[v * (VOX / 4)]) + 64 * c));
                    }
                }
            }
        };

        // Footprints of this wave's voxels in piece t (lane = voxel), moved
        // into the slot table st (lane = slot) by cross-lane permutes: slot
        // (rank, q) holds the rank-th voxel whose footprint starts at pixel
        // p0 - 2 + q (at most two per pixel) as (wl, wr, w3, LDS offset).
        // Empty slots hold zero weights and offset 0, so the sweep is
        // straight-line code.  Returns whether the wave has anything to add
        // from this piece.
        auto setup = [&](const Task& t, Slots& st) -> int {
            if (!mine) return 0;
            const int jlo = uni(__builtin_amdgcn_readlane(s_lo, t.row - c0));
            const int jhi = uni(__builtin_amdgcn_readlane(s_hi, t.row - c0));
            const int n = jhi - jlo + 1;
            if (n <= 0) return 0;
            const float R = (float)t.row - cr;
            const int mcl = min(lane, n - 1);
            const int jm = fwd ? jlo + mcl : jhi - mcl;
            // footprint anchored at its first nonzero pixel l; slot q = l - (pl - 1)
            // covers l in [p0-2, p0+P-1], the voxels that touch the block's pixels
            const Taps2 tp = anchor_left(
                voxel_taps(px_ ? (float)jm - xc : R, px_ ? R : (float)jm - yc, k));
            const int lprev = lane_shift_up(tp.l);    // lane m-1 holds the previous voxel
            const int lnext = lane_shift_down(tp.l);  // lane m+1 the next
            const bool valid = lane < n;
            const int l = valid ? tp.l : 0x3fffffff;
            const int mm0 = uni((int)__popcll(__ballot(valid && l < pl - 1)));
            const int mend = uni((int)__popcll(__ballot(valid && l <= ph - 1)));
            int mA = fwd ? t.j0 - jlo : jhi - t.j1;
            int mB = fwd ? t.j1 - jlo : jhi - t.j0;
            mA = uni(max(mA, mm0));
            mB = uni(min(mB, mend - 1));
            if (mA > mB) return 0;
            const int jb0 = fwd ? (jlo - t.j0) : (jhi - t.j0);  // staged slot of lane 0's voxel
            const int dj = fwd ? 1 : -1;
            // Destination lane of each voxel in the rank-0 and rank-1 tables;
            // voxels outside [mA, mB] go to lane 63, which is never a slot.
            // Of two voxels with the same l (lanes m-1, m), rank 1 is the one
            // staged later: lane m when j grows with the lane (fwd), else
            // lane m-1.
            const bool in = lane >= mA && lane <= mB;
            const bool second = fwd ? (lane > mA && l == lprev) : (lane < mB && l == lnext);
            const int q = l - (pl - 1);
            const int d0 = 4 * ((in && !second) ? q : 63);
            const int d1 = 4 * ((in && second) ? q : 63);
            const int off = (jb0 + dj * lane) * VOX;
            st.t0 = permute4(d0, tp.wl, tp.wr, tp.w3, off);
            st.t1 = permute3(d1, tp.wl, tp.wr, tp.w3);  // its offset is t0's plus one voxel
            return 1;
        };

        // Pipeline over pieces: while one piece is swept from one buffer, the
        // next streams into the other buffer and its slot table is built.
        __syncthreads();  // previous pass's readers of s_xa / s_xb are done
        Task cur = first_task(0);
        int flag = 0;
        Slots sl;
        if (cur.row >= 0) {
            fetch(cur, s_xa);
            flag = setup(cur, sl);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        int rows_done = 0;
        // One piece: sweep `cur` from xcur with slot table sl; stage the next
        // piece into xnxt and build its table.  Inlined at two call sites so
        // that both buffers are compile-time known.
        auto step = [&](const float4* xcur, float4* xnxt) __attribute__((always_inline)) {
            const Task nxt = next_task(cur);
            int flag_n = 0;
            Slots sl_n;
            // Staging DMA of the next piece.  With one voxel per instruction
            // (4 slices per lane), each wave issues its instruction k at pixel
            // slot 2k of its sweep instead of all of them at once right after
            // the barrier, when every wave of the block queues its DMAs at the
            // same time (1-4 % faster; same data, same LDS slots).
            constexpr bool kSpread = (VOX * 4 == 1024);
            constexpr int kMaxIssue = (UMAX + NW - 1) / NW;
            constexpr int kInSweep = kMaxIssue < (P + 3) / 2 ? kMaxIssue : (P + 3) / 2;
            int dcnt = 0;
            const float* dsrc = x;
            float4* dd = xnxt;
            int64_t dstep = 0;
            uint32_t dvoff = 0, drange = 0;
            if (nxt.row >= 0) {
                flag_n = setup(nxt, sl_n);
                if (kSpread && vec4) {
                    const int nv = nxt.j1 - nxt.j0 + 1;
                    dcnt = wave < nv ? (nv - wave + NW - 1) / NW : 0;
                    dsrc = x + (int64_t)nxt.row * st_row + (int64_t)(nxt.j0 + wave) * st_along;
                    dd = xnxt + wave * (VOX / 4);
                    dstep = (int64_t)NW * st_along;
                    dvoff = (uint32_t)min(hb + lane * 4, H - 4) * 4;
                    drange = (uint32_t)H * 4;
                } else {
                    fetch(nxt, xnxt);
                }
            }
            auto issue = [&](int k) __attribute__((always_inline)) {
                if (k < dcnt)
                    dma_to_lds<16>(dsrc + k * dstep, drange, dvoff,
                                   (__attribute__((address_space(3))) void*)(dd + k * NW * (VOX / 4)));
            };
            if (flag) {
                const float4 t0 = sl.t0;
                const float4 t1 = sl.t1;
                // Slot masks, 32-bit (one scalar bit test each).  Only slot
                // lanes 0 .. P+1 and the discard lane 63 receive table entries
                // (the others read zero), and the cast drops lane 63.
                static_assert(P + 2 <= 32, "slot masks are 32-bit");
                using Mask = uint32_t;
                // bit q: slot q holds a second voxel (about 1 in 10 at 512^3 x
                // 192); a voxel's first weight wl is always positive
                const Mask has2 = (Mask)__ballot(t1.x != 0.f);
                // bit q: the rank-0 / rank-1 voxel of slot q covers three pixels
                const Mask has3a = (Mask)__ballot(t0.z != 0.f);
                const Mask has3b = (Mask)__ballot(t1.z != 0.f);
                const Mask extra = has3a | has2;  // has3b implies has2
                const float* sx = reinterpret_cast<const float*>(xcur) + lane * HL;
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
                // Software-pipelined sweep: the two slot rows of pixel q+1 are
                // fetched (unconditionally) before the FMAs of pixel q.  The
                // second voxel is the next staged slot, so both reads share one
                // address and the second takes a constant instruction offset.
                auto slot_src = [&](int q) { return sx + __builtin_bit_cast(int, rdlane(t0.w, q)); };
                const float* src = slot_src(0);
                auto xa_n = ld(src);
                auto xb_n = ld(src + VOX);
                // the rank-0 weights of pixel q+1 are read out one pixel ahead too
                float wl_n = rdlane(t0.x, 0), wr_n = rdlane(t0.y, 0);
#pragma unroll
                for (int q = 0; q < P + 2; ++q) {
                    if (q % 2 == 0 && q / 2 < kInSweep) issue(q / 2);
                    const auto xa = xa_n;
                    const auto xb = xb_n;
                    const float wl = wl_n, wr = wr_n;
                    if (q + 1 < P + 2) {
                        src = slot_src(q + 1);
                        xa_n = ld(src);
                        xb_n = ld(src + VOX);
                        wl_n = rdlane(t0.x, q + 1);
                        wr_n = rdlane(t0.y, q + 1);
                    }
                    {
#pragma unroll
                        for (int j = 0; j < HL; ++j) {
                            if (q - 2 >= 0 && q - 2 < P) part[q - 2][j] = fmaf(wl, xa.a[j], part[q - 2][j]);
                            if (q - 1 >= 0 && q - 1 < P) part[q - 1][j] = fmaf(wr, xa.a[j], part[q - 1][j]);
                        }
                    }
                    // one test for the common case of a slot with neither a
                    // three-pixel rank-0 footprint nor a second voxel
                    if ((extra >> q) & 1u) {
                        if (q < P && ((has3a >> q) & 1u)) {
                            const float w3 = rdlane(t0.z, q);
#pragma unroll
                            for (int j = 0; j < HL; ++j) part[q][j] = fmaf(w3, xa.a[j], part[q][j]);
                        }
                        if ((has2 >> q) & 1u) {
                            const float wl = rdlane(t1.x, q);
                            const float wr = rdlane(t1.y, q);
#pragma unroll
                            for (int j = 0; j < HL; ++j) {
                                if (q - 2 >= 0 && q - 2 < P) part[q - 2][j] = fmaf(wl, xb.a[j], part[q - 2][j]);
                                if (q - 1 >= 0 && q - 1 < P) part[q - 1][j] = fmaf(wr, xb.a[j], part[q - 1][j]);
                            }
                            if (q < P && ((has3b >> q) & 1u)) {
                                const float w3 = rdlane(t1.z, q);
#pragma unroll
                                for (int j = 0; j < HL; ++j) part[q][j] = fmaf(w3, xb.a[j], part[q][j]);
                            }
                        }
                    }
                }
            }
            for (int k = flag ? kInSweep : 0; k < dcnt; ++k) issue(k);  // not issued in a sweep
            __builtin_amdgcn_s_waitcnt(0);  // this wave's DMA into xnxt has landed
            __syncthreads();
            cur = nxt;
            flag = flag_n;
            sl = sl_n;
        };
        while (cur.row >= 0) {
            // Fold the partial sums when a new group of kRowsPerFold lines
            // starts (checked once per pair of pieces, so the fold code exists
            // once; its global loads retire before the next DMA is issued, as
            // vmcnt retires in order).
            if (cur.row / kRowsPerFold != rows_done / kRowsPerFold) {
                fold();
                rows_done = cur.row;
            }
            step(s_xa, s_xb);
            if (cur.row < 0) break;
            step(s_xb, s_xa);
        }
        fold();
    }
    fold();
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
    const int bits = hi_bits(it.SY, it.SX, it.PX);
    g.du = nullptr;
    g.scale = ldexp(1.0, bits);
    g.inv_scale = ldexp(1.0, -bits);
    return g;
}

template <int P, int NW, int HL, int UMAX>
__global__ __launch_bounds__(64 * NW) void forward_z0_kernel(OpArgs g)
{
    forward_block<P, NW, HL, UMAX>(g, xcd_logical(blockIdx.x, gridDim.x));
}

// Persistent launch over a list of instances: block b handles work items
// b, b + gridDim.x, ... of the concatenated (instance, block) list, in a
// fixed order, so results do not depend on the grid or on the other
// instances of the batch.
template <int P, int NW, int HL, int UMAX>
__global__ __launch_bounds__(64 * NW) void forward_batched_kernel(const BatchItem* __restrict__ items,
                                                                  int n, int64_t total)
{
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        int i = 0;
        int64_t base = 0;
        for (; i < n - 1; ++i) {
            int nbp, nba, nbh;
            forward_grid<P, NW, HL>(items[i].A, items[i].H, items[i].PX, nbp, nba, nbh);
            const int64_t nbl = (int64_t)nbp * nba * nbh;
            if (w < base + nbl) break;
            base += nbl;
        }
        forward_block<P, NW, HL, UMAX>(item_args(items[i]), (int)(w - base));
        __syncthreads();
    }
}

constexpr int kTile = 64;

__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ in,
                                                        float* __restrict__ out, int64_t R,
                                                        int64_t C, int accumulate)
{
    __shared__ float t[kTile][kTile + 1];
    const int64_t tiles_r = (R + kTile - 1) / kTile, tiles_c = (C + kTile - 1) / kTile;
    // 16-byte rows: both extents multiples of 4 and both pointers aligned
    const bool vec = ((R | C) & 3) == 0 && ((reinterpret_cast<uintptr_t>(in) |
                                             reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // scalar path
    const int lr = threadIdx.x >> 4, lc = (threadIdx.x & 15) * 4;  // vector path
    for (int64_t tile = blockIdx.x; tile < tiles_r * tiles_c; tile += gridDim.x) {
        const int64_t r0 = (tile / tiles_c) * kTile, c0 = (tile % tiles_c) * kTile;
        if (vec && r0 + kTile <= R && c0 + kTile <= C) {
            // full tile: 16-byte loads and stores, all old values of an
            // accumulating store loaded before any is written
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
}

template <int HL>
struct FwdCfg;
// 1 and 2 slices per lane: 12 angles per block (3 waves per SIMD) with
// 28-pixel strips; against 8 angles with 32-pixel strips, 10-25 % faster on
// most mid-size volumes (256^2 - 448^2, 64 - 128 slices), never more than
// 2 % slower in the shapes measured.  Both variants share P, NW and UMAX, so
// they add every output's terms in the same order (bitwise-equal results).
template <> struct FwdCfg<1> { static constexpr int P = 28, NW = 12, UMAX = 128; };
template <> struct FwdCfg<2> { static constexpr int P = 28, NW = 12, UMAX = 128; };
// 4 slices per lane: 12 angles (waves) per block at 3 waves per SIMD (168
// VGPRs; the few spills are outside the sweep) and 22-pixel strips.  Against
// 8 waves with 28-pixel strips: 6 % faster per unit of work (1024^2 x 512 x
// 1800), and at 512^3 x 192 the grid is exactly 3 x 256 workgroups instead of
// 3.6 x 256 (6.2 vs 7.3 ms); 2-6 % slower at 256^3 and 384^3.  Pieces of up
// to 76 voxels (152 KB of staging) split fewer lines than 72 (1.7 % faster).
template <> struct FwdCfg<4> { static constexpr int P = 22, NW = 12, UMAX = 76; };

// Narrower strips for 1 slice per lane, used by single launches of volumes
// of at most UMAX = 128 voxels per side when the default grid would have fewer
// workgroups than the device has CUs: twice the workgroups, each with half the
// pixels (128^3 x 192: 0.46 -> 0.38 ms).  Within that size no line is split
// into pieces and all lines fall in one fold group, so every output adds the
// same terms in the same order as with the default strips: results are equal
// bit for bit.  (For larger volumes the strip width changes where a line's staged
// union is split and which pieces pair up around a fold, hence the last bits.)
struct FwdSmall {
    static constexpr int P = 16, NW = 8, UMAX = 128;
};

template <int HL, typename C = FwdCfg<HL>>
hipError_t forward_single(const OpArgs& g, hipStream_t s)
{
    int nbp, nba, nbh;
    forward_grid<C::P, C::NW, HL>(g.A, g.H, g.PX, nbp, nba, nbh);
    const int64_t nb = (int64_t)nbp * nba * nbh;
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((forward_z0_kernel<C::P, C::NW, HL, C::UMAX>), dim3((unsigned)nb),
                       dim3(64 * C::NW), 0, s, g);
    return hipGetLastError();
}

template <int HL>
hipError_t forward_batched(const BatchItem* host, const BatchItem* dev, int n, int grid,
                           hipStream_t s)
{
    using C = FwdCfg<HL>;
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        int nbp, nba, nbh;
        forward_grid<C::P, C::NW, HL>(host[i].A, host[i].H, host[i].PX, nbp, nba, nbh);
        total += (int64_t)nbp * nba * nbh;
    }
    if (total == 0) return hipSuccess;
    const int64_t gsz = grid > 0 ? std::min<int64_t>(grid, total) : total;
    hipLaunchKernelGGL((forward_batched_kernel<C::P, C::NW, HL, C::UMAX>), dim3((unsigned)gsz),
                       dim3(64 * C::NW), 0, s, dev, n, total);
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
    a.scale = ldexp(1.0, g.bits);
    a.inv_scale = ldexp(1.0, -g.bits);
    a.du = g.du;
    switch (single_variant(0, g.A, g.SY, g.SX, g.H, g.PX)) {
        case 1:
            if (std::max(g.SY, g.SX) <= FwdSmall::UMAX && forward_blocks(1, g.A, g.H, g.PX) < device_cus())
                return forward_single<1, FwdSmall>(a, s);
            return forward_single<1>(a, s);
        case 2: return forward_single<2>(a, s);
        default: return forward_single<4>(a, s);
    }
}

hipError_t launch_forward_batched(int variant, const BatchItem* items_host,
                                  const BatchItem* items_dev, int n, int grid, hipStream_t s)
{
    switch (variant) {
        case 1: return forward_batched<1>(items_host, items_dev, n, grid, s);
        case 2: return forward_batched<2>(items_host, items_dev, n, grid, s);
        case 4: return forward_batched<4>(items_host, items_dev, n, grid, s);
        default: return hipErrorInvalidValue;
    }
}

int64_t forward_blocks(int variant, int A, int64_t H, int64_t PX)
{
    int nbp = 0, nba = 0, nbh = 0;
    switch (variant) {
        case 1: forward_grid<FwdCfg<1>::P, FwdCfg<1>::NW, 1>(A, H, PX, nbp, nba, nbh); break;
        case 2: forward_grid<FwdCfg<2>::P, FwdCfg<2>::NW, 2>(A, H, PX, nbp, nba, nbh); break;
        case 4: forward_grid<FwdCfg<4>::P, FwdCfg<4>::NW, 4>(A, H, PX, nbp, nba, nbh); break;
        default: return -1;
    }
    return (int64_t)nbp * nba * nbh;
}

hipError_t launch_transpose(const float* in, float* out, int64_t R, int64_t C, int accumulate,
                            hipStream_t s)
{
    // workgroups stride over the tiles (one-dimensional grid: no limit on
    // either extent's tile count)
    const int64_t tiles = ((R + kTile - 1) / kTile) * ((C + kTile - 1) / kTile);
    if (tiles == 0) return hipSuccess;
    const unsigned grid = (unsigned)std::min<int64_t>(tiles, 8192);
    hipLaunchKernelGGL(transpose_kernel, dim3(grid), dim3(256), 0, s, in, out, R, C, accumulate);
    return hipGetLastError();
}

}  // namespace vp
