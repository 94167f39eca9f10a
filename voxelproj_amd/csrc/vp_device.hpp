// Device-side geometry and footprint math shared by the projector kernels.
//
// Geometry (reference: voxelproj/csrc/include/sp_backward_z0.cuh:24-26,84 and
// sp_forward_z0.cuh:132-156): voxel size = detector pixel size = 1, rotation
// centre at the volume centre, detector centre on the rotation axis (unless
// translated).  The detector coordinate of voxel (ix, iy) at angle phi is
//     u = (ix - xc) cos(phi) + (iy - yc) sin(phi) + pc - du,
// xc = (SX-1)/2, yc = (SY-1)/2, pc = (PX-1)/2, du the detector translation
// of that angle in pixels (0 in the reference geometry).  The weight of pixel p is the
// area of the unit voxel inside the unit detector strip around p:
//     W = F(p + 1/2 - u) - F(p - 1/2 - u),
// F the CDF of the unit-area trapezoid with plateau half-width
// a = ||c|-|s||/2, base half-width b = (|c|+|s|)/2 and height h = 1/max(|c|,|s|)
// (projection_common.cuh:139-166 evaluates the same F).  The support is at
// most the three pixels r-1, r, r+1 around r = round(u).
//
// Numerics.  u reaches several hundred, so forming it in FP32 would put an
// absolute error of ~1e-5 on every footprint position.  Instead cos and sin
// are split on the host side of each angle (in FP64) into a "hi" part with
// B fractional bits and a small "lo" rest.  With B chosen from the volume
// extents, (ix-xc)*c_hi + (iy-yc)*s_hi + pc is exact in FP32, and only the
// tiny lo term and the final fractional offset round.  Positions are then
// good to ~3e-8 and the weights to ~1e-7 absolute.  A translation du enters
// the same way: its integer part joins pc in the exact hi sum, its fraction
// (|.| <= 1/2) the lo sum.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vp {

// Per-angle constants, rounded once from FP64.
struct AngleK {
    float chi, clo;  // cos = chi + clo, chi a multiple of 2^-B
    float shi, slo;  // sin = shi + slo
    float a, b;      // plateau / base half-widths of the footprint trapezoid
    float h;         // plateau height 1/max(|c|,|s|)
    float cd;        // h / (2 (b - a)); 0 when a == b (quadratic flanks vanish)
    float pch, plo;  // pc - du = pch + plo: pch = pc - round(du) (exact), plo the fraction
};

// pc = (PX-1)/2; du = detector translation of this angle (pixels).
__device__ __forceinline__ AngleK make_angle(float phi, float du, double pc, double scale,
                                             double inv_scale)
{
    const double c = cos((double)phi);
    const double s = sin((double)phi);
    const double chi = rint(c * scale) * inv_scale;
    const double shi = rint(s * scale) * inv_scale;
    const double ac = fabs(c), as = fabs(s);
    const double a = 0.5 * fabs(ac - as);
    const double b = 0.5 * (ac + as);
    const double h = 1.0 / fmax(ac, as);
    AngleK k;
    k.chi = (float)chi;
    k.clo = (float)(c - chi);
    k.shi = (float)shi;
    k.slo = (float)(s - shi);
    k.a = (float)a;
    k.b = (float)b;
    k.h = (float)h;
    const float bf = (float)b, af = (float)a;
    k.cd = (bf > af) ? 0.5f * k.h / (bf - af) : 0.0f;
    const double dh = rint((double)du);
    k.pch = (float)(pc - dh);
    k.plo = (float)(dh - (double)du);
    return k;
}

// Q(z) = F(-z) for z >= 0 (left tail of the footprint CDF); also valid for
// z slightly below 0, where the linear plateau piece continues.
__device__ __forceinline__ float tail(float z, const AngleK& k)
{
    const float lin = fmaf(-k.h, z, 0.5f);
    const float e = fmaxf(k.b - z, 0.0f);
    const float quad = (k.cd * e) * e;
    return (z <= k.a) ? lin : quad;
}

// Footprint of one voxel at one angle: pixel r and weights of r-1, r, r+1.
struct Taps {
    int r;
    float wm, w0, wp;
};

// r = round(u), computed from (uhi + 1/2) + ulo: uhi + 1/2 is exact and the
// single rounding of the sum moves r only when u is within ~1e-4 of a
// half-integer, where either neighbour is a valid centre (|u - r| stays
// within 1/2 + 1e-4, and the footprint tails are linear there to second
// order).
__device__ __forceinline__ float centre_pixel(float uhi, float ulo)
{
    return floorf((uhi + 0.5f) + ulo);
}

// X = ix - xc, Y = iy - yc (exact half-integers).
__device__ __forceinline__ Taps voxel_taps(float X, float Y, const AngleK& k)
{
    const float uhi = fmaf(X, k.chi, fmaf(Y, k.shi, k.pch));  // exact
    const float ulo = fmaf(X, k.clo, fmaf(Y, k.slo, k.plo));  // small
    const float rf = centre_pixel(uhi, ulo);
    const float t = (uhi - rf) + ulo;                      // u - r; uhi - rf exact
    Taps o;
    o.r = (int)rf;
    o.wm = tail(0.5f + t, k);   // pixel r-1: F(-1/2 - t)
    o.wp = tail(0.5f - t, k);   // pixel r+1: 1 - F(1/2 - t)
    o.w0 = (1.0f - o.wm) - o.wp;
    return o;
}

// The same footprint anchored at its first nonzero pixel l: pixel l gets wl,
// l+1 gets wr and l+2 gets w3.  The centre pixel always has a positive weight
// (it contains the footprint's centre), so l = r-1 when pixel r-1 is hit and
// l = r otherwise; w3 is nonzero only when the footprint covers three pixels
// (about one voxel-angle in eight at 512^3 x 192), so the projectors apply
// two taps always and the third under a wave-uniform test.  Dropping the zero
// tap leaves every sum unchanged (x*0 adds +-0).
struct Taps2 {
    int l;
    float wl, wr, w3;
};

__device__ __forceinline__ Taps2 anchor_left(const Taps& t)
{
    const bool left = t.wm != 0.0f;
    Taps2 o;
    o.l = left ? t.r - 1 : t.r;
    o.wl = left ? t.wm : t.w0;
    o.wr = left ? t.w0 : t.wp;
    o.w3 = left ? t.wp : 0.0f;
    return o;
}

// Centre pixel only, for window and segment bounds.
__device__ __forceinline__ int voxel_r(float X, float Y, const AngleK& k)
{
    const float uhi = fmaf(X, k.chi, fmaf(Y, k.shi, k.pch));
    const float ulo = fmaf(X, k.clo, fmaf(Y, k.slo, k.plo));
    return (int)centre_pixel(uhi, ulo);
}

// Number of fractional bits of the hi split for a volume/detector extent and
// a bound on |du|.  Every partial sum of (ix-xc)*c_hi + (iy-yc)*s_hi + pch is
// a multiple of 2^-(B+1) below 2^E in magnitude, so E + B + 1 <= 24 keeps it
// exact.
__host__ __device__ inline int hi_bits(int64_t SY, int64_t SX, int64_t PX, double du_max = 0.0)
{
    const double m = 0.5 * (double)(SX - 1) + 0.5 * (double)(SY - 1) + 0.5 * (double)(PX - 1) +
                     du_max + 2.0;
    int e = 0;
    while ((double)(1ll << e) < m) ++e;
    return 23 - e;
}

// Global -> LDS DMA of BYTES (4 or 16) per lane: lane data from base + voff
// (bytes) lands at the lane's slot of the instruction's LDS destination.
// `base` must be wave-uniform; reads at voff >= range return 0.  This is the
// buffer (MUBUF) form of the transfer: the compiler tracks it with vmcnt
// only, so LDS reads issued while it is in flight keep their pipelined
// lgkmcnt waits.  With the FLAT form (global_load_lds) in flight it treats
// the LDS counter as unordered and makes every LDS wait lgkmcnt(0).
template <int BYTES>
__device__ __forceinline__ void dma_to_lds(const void* base, uint32_t range, uint32_t voff,
                                           __attribute__((address_space(3))) void* lds)
{
    constexpr int kRsrcWord3 = 0x00020000;  // raw 32-bit buffer, gfx9 family
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)range, kRsrcWord3);
    static_assert(BYTES == 4 || BYTES == 16, "DMA moves 4 or 16 bytes per lane");
    if constexpr (BYTES == 16)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(r, lds, 16, voff, 0, 0, 0);
    else
        __builtin_amdgcn_raw_ptr_buffer_load_lds(r, lds, 4, voff, 0, 0, 0);
}

// The same transfer from a descriptor built once for many instructions
// (dma_rsrc), with a wave-uniform byte offset `soff` (an SGPR) added to
// every lane's address: one scalar add per instruction instead of a new
// descriptor.  The caller keeps soff + voff + BYTES within `range`, so no
// lane depends on how the range check treats soff.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t dma_rsrc(const void* base, uint32_t range)
{
    constexpr int kRsrcWord3 = 0x00020000;  // raw 32-bit buffer, gfx9 family
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)range, kRsrcWord3);
}

__device__ __forceinline__ void dma16_to_lds(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff,
                                             __attribute__((address_space(3))) void* lds)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, lds, 16, voff, soff, 0, 0);
}

// An LDS array that a wave reads uniformly in an unrolled loop (one element
// per step), addressed from a VGPR that holds its base.  LDS reads take a
// VGPR address; left to itself the compiler keeps a uniform base in an SGPR
// and copies it into a VGPR before every read (one VALU instruction per
// step).  The copy made here is an asm result, which the compiler keeps
// rather than recomputes, and each read folds its element offset into the
// instruction.  `p` must point into LDS and be wave-uniform.  (Components of
// these vector types are read as values: __builtin_bit_cast of a component
// expression reinterprets the vector's first element instead.)
typedef float pair_f32 __attribute__((ext_vector_type(2)));
typedef float quad_f32 __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ __attribute__((address_space(3))) const T* lds_vgpr_base(const void* p)
{
    const uint32_t s = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const void*)p;
    uint32_t v;
    asm("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
    return (__attribute__((address_space(3))) const T*)(uintptr_t)v;
}

}  // namespace vp
