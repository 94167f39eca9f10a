/*
 * CPU oracle for the exact-trapezoid parallel-beam projector pair (A / AT).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library calls, links or
 * loads this file; only tests/, the smoke check and bench.py's CPU-baseline
 * leg use it, and only as the checker.
 *
 * Two restatements live here:
 *
 *  1. vpo_forward_f64 / vpo_backward_f64 -- the exact operator in FP64.
 *     W_a[p, v] = F_a(p + 1/2 - u_a(v)) - F_a(p - 1/2 - u_a(v)),
 *     u_a(ix, iy) = (ix - xc) cos + (iy - yc) sin + pc, F_a the CDF of the
 *     unit-area trapezoid footprint of a unit voxel.  Follows the FP64 formula
 *     of the reference's dpiece_wise_integrated
 *     (voxelproj/csrc/include/projection_common.cuh:61-85) and the 3-tap
 *     support / border-zero rule of the voxel-driven backprojector
 *     (voxelproj/csrc/include/sp_backward_z0.cuh:84-98).  A is evaluated
 *     voxel-driven here as the exact transpose of AT.
 *
 *  2. vpo_ref_forward_f32 / vpo_ref_backward_f32 -- an FP32 transcription of
 *     the reference CUDA kernels' arithmetic (ray-driven A with box clipping
 *     and incremental stepping, voxel-driven 3-tap AT), with IEEE libm
 *     sinf/cosf and no FMA contraction.  It stands in for the reference CUDA
 *     operator, which needs nvcc and an NVIDIA GPU and cannot run here.
 *       ray sum / traversal:  voxelproj/csrc/include/sp_forward_z0.cuh:8-185
 *                             voxelproj/csrc/include/sp_forward_z2.cuh:8-174
 *       clipping:             voxelproj/csrc/include/projection_common.cuh:168-272
 *       footprint CDF:        voxelproj/csrc/include/projection_common.cuh:139-166
 *       backprojection:       voxelproj/csrc/include/sp_backward_z0.cuh:8-100
 *                             voxelproj/csrc/include/sp_backward_z2.cuh:8-101
 *
 * Layouts (element strides are derived from z_order):
 *   z_order 0: x (SY, SX, H),  y (A, PX, H)   -- H fastest
 *   z_order 2: x (H, SY, SX),  y (H, A, PX)   -- H slowest
 *
 * All functions only touch H slices [h0, h1); slices are independent.
 * Outputs are ACCUMULATED into (out += op(in)), as the reference does.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <omp.h>

#define VPO_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* layout helpers                                                     */
/* ------------------------------------------------------------------ */
typedef struct {
    int64_t vy, vx, vh; /* volume strides for iy, ix, h */
    int64_t pa, pp, ph; /* sinogram strides for angle, pixel, h */
} vpo_strides;

static int vpo_make_strides(int z_order, int64_t SY, int64_t SX, int64_t H,
                            int64_t A, int64_t PX, vpo_strides *st)
{
    if (z_order == 0) {
        st->vy = SX * H; st->vx = H; st->vh = 1;
        st->pa = PX * H; st->pp = H; st->ph = 1;
        return 0;
    }
    if (z_order == 2) {
        st->vy = SX; st->vx = 1; st->vh = SY * SX;
        st->pa = PX; st->pp = 1; st->ph = A * PX;
        return 0;
    }
    return -1;
}

static inline double vpo_load(const void *p, int is_f64, int64_t i)
{
    return is_f64 ? ((const double *)p)[i] : (double)((const float *)p)[i];
}

/* ------------------------------------------------------------------ */
/* FP64 exact footprint                                               */
/* ------------------------------------------------------------------ */
typedef struct {
    double c, s;   /* cos, sin of the angle */
    double a, b;   /* plateau / base half-widths of the trapezoid */
    double h;      /* plateau height = 1 / max(|c|, |s|) */
} vpo_angle64;

static vpo_angle64 vpo_angle_consts64(float phi)
{
    vpo_angle64 k;
    double ac, as;
    k.c = cos((double)phi);
    k.s = sin((double)phi);
    ac = fabs(k.c);
    as = fabs(k.s);
    k.a = 0.5 * fabs(ac - as);
    k.b = 0.5 * (ac + as);
    k.h = 1.0 / (ac > as ? ac : as);
    return k;
}

/* CDF of the unit-area trapezoid centred at 0 (projection_common.cuh:61-85). */
static inline double vpo_cdf64(double t, const vpo_angle64 *k)
{
    const double a = k->a, b = k->b, h = k->h;
    if (t <= -b) return 0.0;
    if (t < -a) return 0.5 * h * (t + b) * (t + b) / (b - a);
    if (t < a) return 0.5 + h * t;
    if (t < b) return 1.0 - 0.5 * h * (b - t) * (b - t) / (b - a);
    return 1.0;
}

/* Weight of pixel p for a voxel whose centre projects to u. */
static inline double vpo_weight64(double p, double u, const vpo_angle64 *k)
{
    return vpo_cdf64(p + 0.5 - u, k) - vpo_cdf64(p - 0.5 - u, k);
}

VPO_EXPORT void vpo_angle_table64(const float *angles, int n, double *out)
{
    for (int i = 0; i < n; ++i) {
        vpo_angle64 k = vpo_angle_consts64(angles[i]);
        out[5 * i + 0] = k.c;
        out[5 * i + 1] = k.s;
        out[5 * i + 2] = k.a;
        out[5 * i + 3] = k.b;
        out[5 * i + 4] = k.h;
    }
}

/* Single weight, exposed for closed-form tests. */
VPO_EXPORT double vpo_weight_f64(float phi, double p, double u)
{
    vpo_angle64 k = vpo_angle_consts64(phi);
    return vpo_weight64(p, u, &k);
}

VPO_EXPORT double vpo_cdf_f64(float phi, double t)
{
    vpo_angle64 k = vpo_angle_consts64(phi);
    return vpo_cdf64(t, &k);
}

/*
 * A (exact, FP64): y[a, p, h] += sum_v W_a[p, v] x[v, h] for h in [h0, h1).
 * Parallel over angles: each angle owns its sinogram rows, so there are no
 * write races.  x may be float32 or float64 (x_is_f64); y is float64.
 */
VPO_EXPORT int vpo_forward_f64(const void *x, int x_is_f64, double *y,
                               const float *angles, int n_angles,
                               int64_t SY, int64_t SX, int64_t H, int64_t PX,
                               int z_order, int64_t h0, int64_t h1, int nthreads)
{
    vpo_strides st;
    if (vpo_make_strides(z_order, SY, SX, H, n_angles, PX, &st)) return -1;
    if (h0 < 0 || h1 > H || h0 > h1) return -2;
    const double xc = 0.5 * (double)(SX - 1);
    const double yc = 0.5 * (double)(SY - 1);
    const double pc = 0.5 * (double)(PX - 1);
    if (nthreads <= 0) nthreads = omp_get_max_threads();

#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int ia = 0; ia < n_angles; ++ia) {
        const vpo_angle64 k = vpo_angle_consts64(angles[ia]);
        for (int64_t iy = 0; iy < SY; ++iy) {
            for (int64_t ix = 0; ix < SX; ++ix) {
                const double u = ((double)ix - xc) * k.c + ((double)iy - yc) * k.s + pc;
                const double r = nearbyint(u);
                for (int t = -1; t <= 1; ++t) {
                    const double p = r + t;
                    if (p < 0.0 || p > (double)(PX - 1)) continue;
                    const double w = vpo_weight64(p, u, &k);
                    if (w == 0.0) continue;
                    const int64_t ip = (int64_t)p;
                    const int64_t vb = iy * st.vy + ix * st.vx;
                    const int64_t pb = (int64_t)ia * st.pa + ip * st.pp;
                    for (int64_t hh = h0; hh < h1; ++hh)
                        y[pb + hh * st.ph] += w * vpo_load(x, x_is_f64, vb + hh * st.vh);
                }
            }
        }
    }
    return 0;
}

/*
 * AT (exact, FP64): x[v, h] += sum_a sum_p W_a[p, v] y[a, p, h].
 * Voxel-driven gather over the 3-pixel support; pixels outside [0, PX) read
 * zero (sp_backward_z0.cuh:93-95 with border addressing, projectors.cu:23-25).
 */
VPO_EXPORT int vpo_backward_f64(const void *y, int y_is_f64, double *x,
                                const float *angles, int n_angles,
                                int64_t SY, int64_t SX, int64_t H, int64_t PX,
                                int z_order, int64_t h0, int64_t h1, int nthreads)
{
    vpo_strides st;
    if (vpo_make_strides(z_order, SY, SX, H, n_angles, PX, &st)) return -1;
    if (h0 < 0 || h1 > H || h0 > h1) return -2;
    const double xc = 0.5 * (double)(SX - 1);
    const double yc = 0.5 * (double)(SY - 1);
    const double pc = 0.5 * (double)(PX - 1);
    if (nthreads <= 0) nthreads = omp_get_max_threads();

    vpo_angle64 *tab = (vpo_angle64 *)__builtin_alloca(sizeof(vpo_angle64) * (size_t)(n_angles > 0 ? n_angles : 1));
    for (int ia = 0; ia < n_angles; ++ia) tab[ia] = vpo_angle_consts64(angles[ia]);

#pragma omp parallel for collapse(2) schedule(static) num_threads(nthreads)
    for (int64_t iy = 0; iy < SY; ++iy) {
        for (int64_t ix = 0; ix < SX; ++ix) {
            const int64_t vb = iy * st.vy + ix * st.vx;
            for (int ia = 0; ia < n_angles; ++ia) {
                const vpo_angle64 *k = &tab[ia];
                const double u = ((double)ix - xc) * k->c + ((double)iy - yc) * k->s + pc;
                const double r = nearbyint(u);
                for (int t = -1; t <= 1; ++t) {
                    const double p = r + t;
                    if (p < 0.0 || p > (double)(PX - 1)) continue;
                    const double w = vpo_weight64(p, u, k);
                    if (w == 0.0) continue;
                    const int64_t pb = (int64_t)ia * st.pa + (int64_t)p * st.pp;
                    for (int64_t hh = h0; hh < h1; ++hh)
                        x[vb + hh * st.vh] += w * vpo_load(y, y_is_f64, pb + hh * st.ph);
                }
            }
        }
    }
    return 0;
}

/*
 * Translated detector (an extension; the reference lists translations as
 * future work, README.md:65-68), from the definition:
 *   u_a(ix, iy) = (ix - xc) cos + (iy - yc) sin + pc - du[a],
 *   sinogram slice h of angle a sees the volume at slice position h + dv[a],
 *   linearly interpolated: (1 - f) x[., k + h] + f x[., k + h + 1],
 *   k = floor(dv[a]), f = dv[a] - k, slices outside [0, H) are zero.
 * du or dv may be NULL (no translation).  Output slices [h0, h1).  The
 * backprojector is written as the transpose of the same sum (gather over the
 * sinogram slices that read volume slice j), not as a second derivation.
 */
static inline double vpo_vol_at(const void *x, int is_f64, int64_t base, int64_t stride_h,
                                int64_t H, int64_t j)
{
    return (j < 0 || j >= H) ? 0.0 : vpo_load(x, is_f64, base + j * stride_h);
}

VPO_EXPORT int vpo_forward_geom_f64(const void *x, int x_is_f64, double *y,
                                    const float *angles, int n_angles,
                                    int64_t SY, int64_t SX, int64_t H, int64_t PX,
                                    int z_order, int64_t h0, int64_t h1,
                                    const double *du, const double *dv, int nthreads)
{
    vpo_strides st;
    if (vpo_make_strides(z_order, SY, SX, H, n_angles, PX, &st)) return -1;
    if (h0 < 0 || h1 > H || h0 > h1) return -2;
    const double xc = 0.5 * (double)(SX - 1);
    const double yc = 0.5 * (double)(SY - 1);
    if (nthreads <= 0) nthreads = omp_get_max_threads();

#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int ia = 0; ia < n_angles; ++ia) {
        const vpo_angle64 k = vpo_angle_consts64(angles[ia]);
        const double pc = 0.5 * (double)(PX - 1) - (du ? du[ia] : 0.0);
        const double d = dv ? dv[ia] : 0.0;
        const double kf = floor(d), f = d - kf;
        const int64_t kv = (int64_t)kf;
        for (int64_t iy = 0; iy < SY; ++iy) {
            for (int64_t ix = 0; ix < SX; ++ix) {
                const double u = ((double)ix - xc) * k.c + ((double)iy - yc) * k.s + pc;
                const double r = nearbyint(u);
                for (int t = -1; t <= 1; ++t) {
                    const double p = r + t;
                    if (p < 0.0 || p > (double)(PX - 1)) continue;
                    const double w = vpo_weight64(p, u, &k);
                    if (w == 0.0) continue;
                    const int64_t vb = iy * st.vy + ix * st.vx;
                    const int64_t pb = (int64_t)ia * st.pa + (int64_t)p * st.pp;
                    for (int64_t hh = h0; hh < h1; ++hh) {
                        const double xv = (1.0 - f) * vpo_vol_at(x, x_is_f64, vb, st.vh, H, hh + kv) +
                                          f * vpo_vol_at(x, x_is_f64, vb, st.vh, H, hh + kv + 1);
                        y[pb + hh * st.ph] += w * xv;
                    }
                }
            }
        }
    }
    return 0;
}

VPO_EXPORT int vpo_backward_geom_f64(const void *y, int y_is_f64, double *x,
                                     const float *angles, int n_angles,
                                     int64_t SY, int64_t SX, int64_t H, int64_t PX,
                                     int z_order, int64_t h0, int64_t h1,
                                     const double *du, const double *dv, int nthreads)
{
    vpo_strides st;
    if (vpo_make_strides(z_order, SY, SX, H, n_angles, PX, &st)) return -1;
    if (h0 < 0 || h1 > H || h0 > h1) return -2;
    const double xc = 0.5 * (double)(SX - 1);
    const double yc = 0.5 * (double)(SY - 1);
    if (nthreads <= 0) nthreads = omp_get_max_threads();

    vpo_angle64 *tab = (vpo_angle64 *)__builtin_alloca(sizeof(vpo_angle64) * (size_t)(n_angles > 0 ? n_angles : 1));
    for (int ia = 0; ia < n_angles; ++ia) tab[ia] = vpo_angle_consts64(angles[ia]);

#pragma omp parallel for collapse(2) schedule(static) num_threads(nthreads)
    for (int64_t iy = 0; iy < SY; ++iy) {
        for (int64_t ix = 0; ix < SX; ++ix) {
            const int64_t vb = iy * st.vy + ix * st.vx;
            for (int ia = 0; ia < n_angles; ++ia) {
                const vpo_angle64 *k = &tab[ia];
                const double pc = 0.5 * (double)(PX - 1) - (du ? du[ia] : 0.0);
                const double d = dv ? dv[ia] : 0.0;
                const double kf = floor(d), f = d - kf;
                const int64_t kv = (int64_t)kf;
                const double u = ((double)ix - xc) * k->c + ((double)iy - yc) * k->s + pc;
                const double r = nearbyint(u);
                for (int t = -1; t <= 1; ++t) {
                    const double p = r + t;
                    if (p < 0.0 || p > (double)(PX - 1)) continue;
                    const double w = vpo_weight64(p, u, k);
                    if (w == 0.0) continue;
                    const int64_t pb = (int64_t)ia * st.pa + (int64_t)p * st.pp;
                    /* volume slice j is read by sinogram slice hh = j - kv
                     * (weight 1 - f) and hh = j - kv - 1 (weight f) */
                    for (int64_t j = h0; j < h1; ++j) {
                        const double g = (1.0 - f) * vpo_vol_at(y, y_is_f64, pb, st.ph, H, j - kv) +
                                         f * vpo_vol_at(y, y_is_f64, pb, st.ph, H, j - kv - 1);
                        x[vb + j * st.vh] += w * g;
                    }
                }
            }
        }
    }
    return 0;
}

/*
 * Exact nonzero count of W over all angles (pixels restricted to [0, PX)),
 * used for the algorithmic flop count 2 * H * nnz.
 */
VPO_EXPORT int64_t vpo_nnz(const float *angles, int n_angles, int64_t SY, int64_t SX,
                           int64_t PX, int nthreads)
{
    const double xc = 0.5 * (double)(SX - 1);
    const double yc = 0.5 * (double)(SY - 1);
    const double pc = 0.5 * (double)(PX - 1);
    int64_t total = 0;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total) num_threads(nthreads)
    for (int ia = 0; ia < n_angles; ++ia) {
        const vpo_angle64 k = vpo_angle_consts64(angles[ia]);
        int64_t cnt = 0;
        for (int64_t iy = 0; iy < SY; ++iy)
            for (int64_t ix = 0; ix < SX; ++ix) {
                const double u = ((double)ix - xc) * k.c + ((double)iy - yc) * k.s + pc;
                const double r = nearbyint(u);
                for (int t = -1; t <= 1; ++t) {
                    const double p = r + t;
                    if (p < 0.0 || p > (double)(PX - 1)) continue;
                    if (vpo_weight64(p, u, &k) != 0.0) ++cnt;
                }
            }
        total += cnt;
    }
    return total;
}

/* ------------------------------------------------------------------ */
/* FP32 transcription of the reference kernels                         */
/* ------------------------------------------------------------------ */
typedef struct { float x, y; } vpo_f2;
typedef struct { float x, y, z, w; } vpo_f4;

/* __float2int_rn: round half to even (default rounding mode). */
static inline int vpo_rn(float v) { return (int)nearbyintf(v); }

/* Footprint constants as the reference kernels build them
 * (sp_forward_z0.cuh:136-149, sp_backward_z0.cuh:45-58). */
static vpo_f4 vpo_ref_pwi(float c, float s)
{
    const float cms = fabsf(0.5f * (c - s));
    const float cps = fabsf(0.5f * (c + s));
    const float lo = cms < cps ? cms : cps;
    const float hi = cms < cps ? cps : cms;
    const float lmax = 1.0f / ((hi - lo) + 2.0f * lo);
    const float coeff = 0.5f * lmax;
    const float coeffdiv = coeff / (hi - lo);
    vpo_f4 r = {lo, hi, coeff, coeffdiv};
    return r;
}

/* spiece_wise_integrated_precalc (projection_common.cuh:139-166). */
static inline float vpo_ref_cdf(float t, vpo_f4 q)
{
    const float a = q.x, b = q.y, coeff = q.z, coeffdiv = q.w;
    if (t <= -b) return 0.0f;
    if (t < -a) { float u = coeffdiv * (t + b); return u * (t + b); }
    if (t < a) { float l = coeff * (b + a); float m = 2.0f * coeff; return l + m * t; }
    if (t < b) { float u = coeffdiv * (b - t); return 1.0f - u * (b - t); }
    return 1.0f;
}

/* SideIntersectVec (projection_common.cuh:168-192): solve a1 + t1 d1 = b1 + t2 d2. */
static vpo_f2 vpo_ref_side(vpo_f2 a1, vpo_f2 d1, vpo_f2 b1, vpo_f2 d2)
{
    const vpo_f2 ba = {b1.x - a1.x, b1.y - a1.y};
    const float d11 = d1.x * d1.x + d1.y * d1.y;
    const float d12 = d1.x * d2.x + d1.y * d2.y;
    const float d22 = d2.x * d2.x + d2.y * d2.y;
    const float ba1 = ba.x * d1.x + ba.y * d1.y;
    const float ba2 = ba.x * d2.x + ba.y * d2.y;
    const float det = d11 * d22 - d12 * d12;
    const float n1 = ba1 * d22 - d12 * ba2;
    const float n2 = ba1 * d12 - d11 * ba2;
    vpo_f2 r = {-1.0f, -1.0f};
    if (fabsf(det) > 1e-15f) { r.x = n1 / det; r.y = n2 / det; }
    return r;
}

/* RectangleIntersectVec (projection_common.cuh:194-272): entry / exit of the
 * line p0 + t*vec with the box (x0, y0, x1, y1); sides tried in the order
 * bottom, top, left, right. */
static int vpo_ref_clip(vpo_f2 p0, vpo_f2 vec, vpo_f4 box, vpo_f4 *out)
{
    const vpo_f2 lo = {box.x, box.y};
    const vpo_f2 ex = {box.z - box.x, 0.0f};
    const vpo_f2 ey = {0.0f, box.w - box.y};
    const vpo_f2 c_br = {box.z, box.y};
    const vpo_f2 c_tl = {box.x, box.w};
    vpo_f2 hit[4];
    hit[0] = vpo_ref_side(p0, vec, lo, ex);   /* bottom */
    hit[1] = vpo_ref_side(p0, vec, c_tl, ex); /* top */
    hit[2] = vpo_ref_side(p0, vec, lo, ey);   /* left */
    hit[3] = vpo_ref_side(p0, vec, c_br, ey); /* right */
    int on[4];
    for (int i = 0; i < 4; ++i) on[i] = (0 <= hit[i].y && hit[i].y <= 1);
    for (int first = 0; first < 3; ++first) {
        if (!on[first]) continue;
        for (int second = first + 1; second < 4; ++second) {
            if (!on[second]) continue;
            out->x = p0.x + hit[first].x * vec.x;
            out->y = p0.y + hit[first].x * vec.y;
            out->z = p0.x + hit[second].x * vec.x;
            out->w = p0.y + hit[second].x * vec.y;
            return 1;
        }
    }
    return 0;
}

/* Ray traversal for one (angle, pixel): calls visit(iqx, iqy, w) for every
 * voxel on the ray in the reference's order.  Returns 0 if the ray misses.
 * stride along y: sp_forward_z0.cuh:58-103 ; along x: :8-56. */
typedef struct {
    int n;
    int cap;
    int *ix, *iy;
    float *w;
} vpo_visits;

static int vpo_ref_ray(float phi, int ip, int nx, int ny, int px, vpo_visits *vis)
{
    const float xcent = 0.5f * (float)(nx - 1);
    const float ycent = 0.5f * (float)(ny - 1);
    const float pcent = 0.5f * (float)(px - 1);
    const float c = cosf(phi), s = sinf(phi);
    const vpo_f4 q = vpo_ref_pwi(c, s);
    const float vx = fabsf(s), vy = fabsf(c);
    const vpo_f2 dir = {-s, c};
    const vpo_f2 org = {((float)ip - pcent) * c + xcent, ((float)ip - pcent) * s + ycent};
    const float rw = 0.5f + 0.5f / (vx > vy ? vx : vy);
    vpo_f4 bb;
    vis->n = 0;
    if (vy > vx) {
        const vpo_f4 box = {-rw, -0.5f, (float)(nx - 1) + rw, (float)ny - 0.5f};
        if (!vpo_ref_clip(org, dir, box, &bb)) return 0;
        const float phw = 0.5f * (1.0f + vx) / vy;
        if (bb.y > bb.w) { float t = bb.x; bb.x = bb.z; bb.z = t; t = bb.y; bb.y = bb.w; bb.w = t; }
        const float rvx = copysignf(vx, bb.z - bb.x);
        const float dx = rvx / vy;
        int y0 = vpo_rn(bb.y); if (y0 < 0) y0 = 0;
        int y1 = vpo_rn(bb.w); if (y1 > ny - 1) y1 = ny - 1;
        float xr = bb.x + dx * ((float)y0 - bb.y);
        for (int iqy = y0; iqy <= y1; ++iqy) {
            int xl = vpo_rn(xr - phw); if (xl < 0) xl = 0;
            int xh = vpo_rn(xr + phw); if (xh > nx - 1) xh = nx - 1;
            for (int iqx = xl; iqx <= xh; ++iqx) {
                const float d = ((float)iqx - xr) * vy;
                const float w = vpo_ref_cdf(0.5f - d, q) - vpo_ref_cdf(-0.5f - d, q);
                if (vis->n < vis->cap) {
                    vis->ix[vis->n] = iqx; vis->iy[vis->n] = iqy; vis->w[vis->n] = w;
                }
                vis->n++;
            }
            xr += dx;
        }
    } else {
        const vpo_f4 box = {-0.5f, -rw, (float)nx - 0.5f, (float)(ny - 1) + rw};
        if (!vpo_ref_clip(org, dir, box, &bb)) return 0;
        const float phw = 0.5f * (1.0f + vy) / vx;
        if (bb.x > bb.z) { float t = bb.x; bb.x = bb.z; bb.z = t; t = bb.y; bb.y = bb.w; bb.w = t; }
        const float rvy = copysignf(vy, bb.w - bb.y);
        const float dy = rvy / vx;
        int x0 = vpo_rn(bb.x); if (x0 < 0) x0 = 0;
        int x1 = vpo_rn(bb.z); if (x1 > nx - 1) x1 = nx - 1;
        float yr = bb.y + dy * ((float)x0 - bb.x);
        for (int iqx = x0; iqx <= x1; ++iqx) {
            int yl = vpo_rn(yr - phw); if (yl < 0) yl = 0;
            int yh = vpo_rn(yr + phw); if (yh > ny - 1) yh = ny - 1;
            for (int iqy = yl; iqy <= yh; ++iqy) {
                const float d = ((float)iqy - yr) * vx;
                const float w = vpo_ref_cdf(0.5f - d, q) - vpo_ref_cdf(-0.5f - d, q);
                if (vis->n < vis->cap) {
                    vis->ix[vis->n] = iqx; vis->iy[vis->n] = iqy; vis->w[vis->n] = w;
                }
                vis->n++;
            }
            yr += dy;
        }
    }
    return 1;
}

/*
 * Reference A in FP32: y += A x, ray-driven, one ray per (angle, pixel); the
 * per-h sums run over the ray's visits in the reference's order.
 */
VPO_EXPORT int vpo_ref_forward_f32(const float *x, float *y, const float *angles, int n_angles,
                                   int64_t SY, int64_t SX, int64_t H, int64_t PX,
                                   int z_order, int64_t h0, int64_t h1, int nthreads)
{
    vpo_strides st;
    if (vpo_make_strides(z_order, SY, SX, H, n_angles, PX, &st)) return -1;
    if (h0 < 0 || h1 > H || h0 > h1) return -2;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    const int64_t nh = h1 - h0;
    const int cap = (int)(4 * (SX + SY) + 16);
    int bad = 0;

#pragma omp parallel num_threads(nthreads) reduction(| : bad)
    {
        int *ixs = (int *)__builtin_alloca(sizeof(int) * (size_t)cap);
        int *iys = (int *)__builtin_alloca(sizeof(int) * (size_t)cap);
        float *ws = (float *)__builtin_alloca(sizeof(float) * (size_t)cap);
        float *acc = (float *)__builtin_alloca(sizeof(float) * (size_t)(nh > 0 ? nh : 1));
        vpo_visits vis = {0, cap, ixs, iys, ws};
#pragma omp for schedule(dynamic, 16)
        for (int64_t ray = 0; ray < (int64_t)n_angles * PX; ++ray) {
            const int ia = (int)(ray / PX);
            const int ip = (int)(ray % PX);
            if (!vpo_ref_ray(angles[ia], ip, (int)SX, (int)SY, (int)PX, &vis)) continue;
            if (vis.n > cap) { bad |= 1; continue; }
            for (int64_t j = 0; j < nh; ++j) acc[j] = 0.0f;
            for (int v = 0; v < vis.n; ++v) {
                const int64_t vb = (int64_t)vis.iy[v] * st.vy + (int64_t)vis.ix[v] * st.vx;
                const float w = vis.w[v];
                for (int64_t j = 0; j < nh; ++j) {
                    const float t = x[vb + (h0 + j) * st.vh] * w;
                    acc[j] = acc[j] + t;
                }
            }
            const int64_t pb = (int64_t)ia * st.pa + (int64_t)ip * st.pp;
            for (int64_t j = 0; j < nh; ++j) y[pb + (h0 + j) * st.ph] += acc[j];
        }
    }
    return bad ? -3 : 0;
}

/* Reference AT in FP32: x += AT y (sp_backward_z0.cuh:69-98). */
VPO_EXPORT int vpo_ref_backward_f32(const float *y, float *x, const float *angles, int n_angles,
                                    int64_t SY, int64_t SX, int64_t H, int64_t PX,
                                    int z_order, int64_t h0, int64_t h1, int nthreads)
{
    vpo_strides st;
    if (vpo_make_strides(z_order, SY, SX, H, n_angles, PX, &st)) return -1;
    if (h0 < 0 || h1 > H || h0 > h1) return -2;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    const int64_t nh = h1 - h0;
    const float xcent = 0.5f * (float)(SX - 1);
    const float ycent = 0.5f * (float)(SY - 1);
    const float pcent = 0.5f * (float)(PX - 1);

    float *cs = (float *)__builtin_alloca(sizeof(float) * 2 * (size_t)(n_angles > 0 ? n_angles : 1));
    vpo_f4 *qs = (vpo_f4 *)__builtin_alloca(sizeof(vpo_f4) * (size_t)(n_angles > 0 ? n_angles : 1));
    for (int ia = 0; ia < n_angles; ++ia) {
        const float c = cosf(angles[ia]), s = sinf(angles[ia]);
        cs[2 * ia] = c; cs[2 * ia + 1] = s;
        qs[ia] = vpo_ref_pwi(c, s);
    }

#pragma omp parallel num_threads(nthreads)
    {
        float *acc = (float *)__builtin_alloca(sizeof(float) * (size_t)(nh > 0 ? nh : 1));
#pragma omp for collapse(2) schedule(static)
        for (int64_t iy = 0; iy < SY; ++iy) {
            for (int64_t ix = 0; ix < SX; ++ix) {
                const float rx = (float)ix - xcent;
                const float ry = (float)iy - ycent;
                for (int64_t j = 0; j < nh; ++j) acc[j] = 0.0f;
                for (int ia = 0; ia < n_angles; ++ia) {
                    const float up = rx * cs[2 * ia] + ry * cs[2 * ia + 1];
                    const float u = up + pcent;
                    const int r = vpo_rn(u);
                    const float fr = (float)r - u;
                    const float v1 = vpo_ref_cdf(fr - 0.5f, qs[ia]);
                    const float v2 = vpo_ref_cdf(fr + 0.5f, qs[ia]);
                    const float w[3] = {v1, v2 - v1, 1.0f - v2};
                    for (int t = 0; t < 3; ++t) {
                        const int p = r - 1 + t;
                        const int inside = (p >= 0 && p < PX);
                        const int64_t pb = (int64_t)ia * st.pa + (int64_t)(inside ? p : 0) * st.pp;
                        for (int64_t j = 0; j < nh; ++j) {
                            const float val = inside ? y[pb + (h0 + j) * st.ph] : 0.0f;
                            const float prod = w[t] * val;
                            acc[j] = acc[j] + prod;
                        }
                    }
                }
                const int64_t vb = iy * st.vy + ix * st.vx;
                for (int64_t j = 0; j < nh; ++j) x[vb + (h0 + j) * st.vh] += acc[j];
            }
        }
    }
    return 0;
}

VPO_EXPORT int vpo_max_threads(void) { return omp_get_max_threads(); }
