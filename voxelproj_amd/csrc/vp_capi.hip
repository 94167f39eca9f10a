// C ABI of the projector library (declared in include/voxelproj_hip.h).
// Validation happens here, before anything touches the device; kernels are
// launched stream-ordered on the caller's stream and every launch is checked.
#include <stddef.h>
#include <stdio.h>
#include <string>

#include "../../include/voxelproj_hip.h"
#include "vp_device.hpp"
#include "vp_launch.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(VP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

constexpr int64_t kMaxExtent = 32768;  // keeps the hi/lo position split accurate

int validate(int z_order, const void* in, const void* out, const float* angles, int n_angles,
             int64_t SY, int64_t SX, int64_t H, int64_t PX, vp::Geom* g, bool* empty)
{
    if (z_order != 0 && z_order != 2)
        return fail(VP_ERR_ARG, "z_order must be 0 or 2, got " + std::to_string(z_order));
    if (SY < 0 || SX < 0 || H < 0 || PX < 0 || n_angles < 0)
        return fail(VP_ERR_ARG, "extents must be non-negative");
    if (SY > kMaxExtent || SX > kMaxExtent || PX > kMaxExtent)
        return fail(VP_ERR_LIMIT, "SY, SX and PX must be <= 32768");
    if (n_angles > 65535)
        return fail(VP_ERR_LIMIT, "at most 65535 angles per call");
    if (H > (int64_t)64 * 65535)
        return fail(VP_ERR_LIMIT, "H must be <= 4194240");
    *empty = (SY == 0 || SX == 0 || H == 0 || PX == 0 || n_angles == 0);
    if (!*empty && (in == nullptr || out == nullptr || angles == nullptr))
        return fail(VP_ERR_ARG, "null device pointer");
    g->SY = SY;
    g->SX = SX;
    g->H = H;
    g->PX = PX;
    g->A = n_angles;
    g->bits = vp::hi_bits(SY, SX, PX);
    return VP_OK;
}

static_assert(sizeof(vp_batch_item) == sizeof(vp::BatchItem), "batch item layout");
static_assert(offsetof(vp_batch_item, n_angles) == offsetof(vp::BatchItem, A), "batch item layout");
static_assert(offsetof(vp_batch_item, accumulate) == offsetof(vp::BatchItem, accumulate),
              "batch item layout");

int validate_batch(int op, int variant, const vp_batch_item* items, int n)
{
    if (n < 0) return fail(VP_ERR_ARG, "negative batch size");
    if (n > 0 && items == nullptr) return fail(VP_ERR_ARG, "null descriptor array");
    if ((op == 0 ? vp::forward_blocks(variant, 1, 1, 1) : vp::backward_blocks(variant, 1, 1, 1)) < 0)
        return fail(VP_ERR_ARG, "unknown kernel variant " + std::to_string(variant));
    for (int i = 0; i < n; ++i) {
        const vp_batch_item& it = items[i];
        vp::Geom g;
        bool empty = false;
        const void* in = it.in;
        const void* out = it.out;
        int rc = validate(0, in, out, it.angles, it.n_angles, it.SY, it.SX, it.H, it.PX, &g, &empty);
        if (rc != VP_OK) return fail(rc, "batch item " + std::to_string(i) + ": " + g_last_error);
        if (empty) return fail(VP_ERR_ARG, "batch item " + std::to_string(i) + " is empty");
        if (it.accumulate != 0 && it.accumulate != 1)
            return fail(VP_ERR_ARG, "batch item " + std::to_string(i) + ": accumulate must be 0 or 1");
        // The forward kernels with 1 or 2 slices per lane stage 4 or 2 voxels
        // per DMA instruction through one 32-bit buffer offset.
        if (op == 0 && (variant == 1 || variant == 2)) {
            const int64_t span = (int64_t)(4 / variant - 1) * it.SX * it.H + it.H;
            if (span >= ((int64_t)1 << 29))
                return fail(VP_ERR_LIMIT, "batch item " + std::to_string(i) + ": SX * H too large for variant " +
                                              std::to_string(variant) + " (use variant 4)");
        }
    }
    return VP_OK;
}

}  // namespace

extern "C" {

int vp_abi_version(void) { return 2; }

const char* vp_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

namespace {

// Detector translation arguments: du / dv device arrays of n_angles floats
// (or null); du_max bounds |du| (it sizes the exact hi/lo position split).
int validate_shift(const float* du, double du_max, vp::Geom* g)
{
    if (!(du_max >= 0.0) || du_max > 1.0e6)
        return fail(VP_ERR_ARG, "du_max must be a finite bound in [0, 1e6]");
    g->du = du;
    g->bits = vp::hi_bits(g->SY, g->SX, g->PX, du ? du_max : 0.0);
    return VP_OK;
}

template <typename T>
struct Workspace {
    T* p = nullptr;
    hipStream_t s;
    hipError_t alloc(int64_t n, hipStream_t st)
    {
        s = st;
        return hipMallocAsync((void**)&p, (size_t)n * sizeof(T), s);
    }
    hipError_t release()
    {
        hipError_t e = p ? hipFreeAsync(p, s) : hipSuccess;
        p = nullptr;
        return e;
    }
};

// y += A x (float32), any layout, optional translations.
int forward_f32(int z_order, const float* x, float* y, const float* angles, const vp::Geom& g,
                const float* dv, hipStream_t s)
{
    const int64_t nx = g.SY * g.SX * g.H;
    const int64_t ny = (int64_t)g.A * g.PX * g.H;
    hipError_t e;
    if (z_order == 0 && !dv) {
        e = vp::launch_forward_z0(x, y, angles, g, 1, s);
        return e == hipSuccess ? VP_OK : hip_fail(e, "forward launch");
    }
    // Workspace: the volume in slice-fastest order (z_order 2 only), the
    // untranslated sinogram, and (z_order 2 with dv) the translated one.
    const int64_t wx = z_order == 2 ? nx : 0;
    const int64_t wy2 = (z_order == 2 && dv) ? ny : 0;
    Workspace<float> ws;
    e = ws.alloc(wx + ny + wy2, s);
    if (e != hipSuccess) return hip_fail(e, "forward workspace");
    const float* x0 = x;
    if (z_order == 2) {
        e = vp::launch_transpose(x, ws.p, g.H, g.SY * g.SX, 0, s);
        x0 = ws.p;
    }
    float* y0 = ws.p + wx;
    if (e == hipSuccess) e = vp::launch_forward_z0(x0, y0, angles, g, 0, s);
    if (z_order == 0) {
        if (e == hipSuccess)
            e = vp::launch_vshift<float>(0, 0, y0, y, g.A, g.PX, g.H, 0, g.H, 0, dv, 1, s);
    } else {
        const float* yz = y0;
        if (dv && e == hipSuccess) {
            e = vp::launch_vshift<float>(0, 0, y0, y0 + ny, g.A, g.PX, g.H, 0, g.H, 0, dv, 0, s);
            yz = y0 + ny;
        }
        if (e == hipSuccess) e = vp::launch_transpose(yz, y, (int64_t)g.A * g.PX, g.H, 1, s);
    }
    hipError_t ef = ws.release();
    if (e != hipSuccess) return hip_fail(e, "forward launch");
    if (ef != hipSuccess) return hip_fail(ef, "forward workspace release");
    return VP_OK;
}

// x += AT y (float32), any layout, optional translations.
int backward_f32(int z_order, const float* y, float* x, const float* angles, const vp::Geom& g,
                 const float* dv, hipStream_t s)
{
    const int64_t nx = g.SY * g.SX * g.H;
    const int64_t ny = (int64_t)g.A * g.PX * g.H;
    hipError_t e;
    if (z_order == 0 && !dv) {
        e = vp::launch_backward_z0(y, x, angles, g, 1, s);
        return e == hipSuccess ? VP_OK : hip_fail(e, "backward launch");
    }
    const int64_t wx = z_order == 2 ? nx : 0;
    const int64_t wy = ny + ((z_order == 2 && dv) ? ny : 0);
    Workspace<float> ws;
    e = ws.alloc(wx + wy, s);
    if (e != hipSuccess) return hip_fail(e, "backward workspace");
    float* x0 = ws.p;
    float* yb = ws.p + wx;
    const float* yz = y;  // sinogram in slice-fastest order
    if (z_order == 2) {
        e = vp::launch_transpose(y, yb, g.H, (int64_t)g.A * g.PX, 0, s);
        yz = yb;
    }
    if (dv && e == hipSuccess) {
        float* g0 = (z_order == 2) ? yb + ny : yb;
        e = vp::launch_vshift<float>(0, 1, yz, g0, g.A, g.PX, g.H, 0, g.H, 0, dv, 0, s);
        yz = g0;
    }
    if (z_order == 0) {
        if (e == hipSuccess) e = vp::launch_backward_z0(yz, x, angles, g, 1, s);
    } else {
        if (e == hipSuccess) e = vp::launch_backward_z0(yz, x0, angles, g, 0, s);
        if (e == hipSuccess) e = vp::launch_transpose(x0, x, g.SY * g.SX, g.H, 1, s);
    }
    hipError_t ef = ws.release();
    if (e != hipSuccess) return hip_fail(e, "backward launch");
    if (ef != hipSuccess) return hip_fail(ef, "backward workspace release");
    return VP_OK;
}

// y += A x (float64).  With dv the untranslated sinogram goes to a zeroed
// workspace first (the FP64 kernels accumulate).
int forward_f64(int z_order, const double* x, double* y, const float* angles, const vp::Geom& g,
                const float* dv, hipStream_t s)
{
    hipError_t e;
    if (!dv) {
        e = vp::launch_forward_f64(z_order, x, y, angles, g, s);
        return e == hipSuccess ? VP_OK : hip_fail(e, "forward f64 launch");
    }
    const int64_t ny = (int64_t)g.A * g.PX * g.H;
    Workspace<double> ws;
    e = ws.alloc(ny, s);
    if (e != hipSuccess) return hip_fail(e, "forward f64 workspace");
    e = vp::launch_fill<double>(ws.p, ny, 0.0, s);
    if (e == hipSuccess) e = vp::launch_forward_f64(z_order, x, ws.p, angles, g, s);
    if (e == hipSuccess)
        e = vp::launch_vshift<double>(z_order, 0, ws.p, y, g.A, g.PX, g.H, 0, g.H, 0, dv, 1, s);
    hipError_t ef = ws.release();
    if (e != hipSuccess) return hip_fail(e, "forward f64 launch");
    if (ef != hipSuccess) return hip_fail(ef, "forward f64 workspace release");
    return VP_OK;
}

int backward_f64(int z_order, const double* y, double* x, const float* angles, const vp::Geom& g,
                 const float* dv, hipStream_t s)
{
    hipError_t e;
    if (!dv) {
        e = vp::launch_backward_f64(z_order, y, x, angles, g, s);
        return e == hipSuccess ? VP_OK : hip_fail(e, "backward f64 launch");
    }
    const int64_t ny = (int64_t)g.A * g.PX * g.H;
    Workspace<double> ws;
    e = ws.alloc(ny, s);
    if (e != hipSuccess) return hip_fail(e, "backward f64 workspace");
    e = vp::launch_vshift<double>(z_order, 1, y, ws.p, g.A, g.PX, g.H, 0, g.H, 0, dv, 0, s);
    if (e == hipSuccess) e = vp::launch_backward_f64(z_order, ws.p, x, angles, g, s);
    hipError_t ef = ws.release();
    if (e != hipSuccess) return hip_fail(e, "backward f64 launch");
    if (ef != hipSuccess) return hip_fail(ef, "backward f64 workspace release");
    return VP_OK;
}

template <typename T, typename Op>
int run_op(Op op, int z_order, const T* in, T* out, const float* angles, int n_angles, int64_t SY,
           int64_t SX, int64_t H, int64_t PX, const float* du, const float* dv, double du_max,
           void* stream)
{
    g_last_error.clear();
    vp::Geom g;
    bool empty = false;
    int rc = validate(z_order, in, out, angles, n_angles, SY, SX, H, PX, &g, &empty);
    if (rc != VP_OK || empty) return rc;
    rc = validate_shift(du, du_max, &g);
    if (rc != VP_OK) return rc;
    return op(z_order, in, out, angles, g, dv, (hipStream_t)stream);
}

}  // namespace

extern "C" {

int vp_forward(int z_order, const float* x, float* y, const float* angles, int n_angles,
               int64_t SY, int64_t SX, int64_t H, int64_t PX, void* stream)
{
    return run_op<float>(forward_f32, z_order, x, y, angles, n_angles, SY, SX, H, PX, nullptr,
                         nullptr, 0.0, stream);
}

int vp_backward(int z_order, const float* y, float* x, const float* angles, int n_angles,
                int64_t SY, int64_t SX, int64_t H, int64_t PX, void* stream)
{
    return run_op<float>(backward_f32, z_order, y, x, angles, n_angles, SY, SX, H, PX, nullptr,
                         nullptr, 0.0, stream);
}

int vp_forward_f64(int z_order, const double* x, double* y, const float* angles, int n_angles,
                   int64_t SY, int64_t SX, int64_t H, int64_t PX, void* stream)
{
    return run_op<double>(forward_f64, z_order, x, y, angles, n_angles, SY, SX, H, PX, nullptr,
                          nullptr, 0.0, stream);
}

int vp_backward_f64(int z_order, const double* y, double* x, const float* angles, int n_angles,
                    int64_t SY, int64_t SX, int64_t H, int64_t PX, void* stream)
{
    return run_op<double>(backward_f64, z_order, y, x, angles, n_angles, SY, SX, H, PX, nullptr,
                          nullptr, 0.0, stream);
}

int vp_forward_geom(int z_order, const float* x, float* y, const float* angles, int n_angles,
                    int64_t SY, int64_t SX, int64_t H, int64_t PX, const float* du,
                    const float* dv, double du_max, void* stream)
{
    return run_op<float>(forward_f32, z_order, x, y, angles, n_angles, SY, SX, H, PX, du, dv,
                         du_max, stream);
}

int vp_backward_geom(int z_order, const float* y, float* x, const float* angles, int n_angles,
                     int64_t SY, int64_t SX, int64_t H, int64_t PX, const float* du,
                     const float* dv, double du_max, void* stream)
{
    return run_op<float>(backward_f32, z_order, y, x, angles, n_angles, SY, SX, H, PX, du, dv,
                         du_max, stream);
}

int vp_forward_geom_f64(int z_order, const double* x, double* y, const float* angles,
                        int n_angles, int64_t SY, int64_t SX, int64_t H, int64_t PX,
                        const float* du, const float* dv, double du_max, void* stream)
{
    return run_op<double>(forward_f64, z_order, x, y, angles, n_angles, SY, SX, H, PX, du, dv,
                          du_max, stream);
}

int vp_backward_geom_f64(int z_order, const double* y, double* x, const float* angles,
                         int n_angles, int64_t SY, int64_t SX, int64_t H, int64_t PX,
                         const float* du, const float* dv, double du_max, void* stream)
{
    return run_op<double>(backward_f64, z_order, y, x, angles, n_angles, SY, SX, H, PX, du, dv,
                          du_max, stream);
}

int vp_vshift(int z_order, int transpose, const float* in, float* out, int n_angles, int64_t PX,
              int64_t H_in, int64_t in_offset, int64_t H_out, int64_t out_offset,
              const float* dv, int accumulate, void* stream)
{
    g_last_error.clear();
    if (z_order != 0 && z_order != 2) return fail(VP_ERR_ARG, "z_order must be 0 or 2");
    if (transpose != 0 && transpose != 1) return fail(VP_ERR_ARG, "transpose must be 0 or 1");
    if (accumulate != 0 && accumulate != 1) return fail(VP_ERR_ARG, "accumulate must be 0 or 1");
    if (n_angles < 0 || PX < 0 || H_in < 0 || H_out < 0)
        return fail(VP_ERR_ARG, "extents must be non-negative");
    if ((int64_t)n_angles * PX * H_out == 0) return VP_OK;
    if (out == nullptr || dv == nullptr || (H_in > 0 && in == nullptr))
        return fail(VP_ERR_ARG, "null device pointer");
    hipError_t e = vp::launch_vshift<float>(z_order, transpose, in, out, n_angles, PX, H_in,
                                            in_offset, H_out, out_offset, dv, accumulate,
                                            (hipStream_t)stream);
    return e == hipSuccess ? VP_OK : hip_fail(e, "vshift launch");
}

int vp_default_variant(int op, int64_t H) { return vp::default_variant(op, H); }

int64_t vp_batch_blocks(int op, int variant, int n_angles, int64_t SY, int64_t SX, int64_t H,
                        int64_t PX)
{
    if (n_angles <= 0 || SY <= 0 || SX <= 0 || H <= 0 || PX <= 0) return -1;
    if (op == 0) return vp::forward_blocks(variant, n_angles, H, PX);
    if (op == 1) return vp::backward_blocks(variant, SY, SX, H);
    return -1;
}

int vp_forward_batched(int variant, const vp_batch_item* items_host, const vp_batch_item* items_dev,
                       int n, int grid, void* stream)
{
    g_last_error.clear();
    int rc = validate_batch(0, variant, items_host, n);
    if (rc != VP_OK || n == 0) return rc;
    if (items_dev == nullptr) return fail(VP_ERR_ARG, "null device descriptor array");
    hipError_t e = vp::launch_forward_batched(
        variant, reinterpret_cast<const vp::BatchItem*>(items_host),
        reinterpret_cast<const vp::BatchItem*>(items_dev), n, grid, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "batched forward launch");
    return VP_OK;
}

int vp_backward_batched(int variant, const vp_batch_item* items_host, const vp_batch_item* items_dev,
                        int n, int grid, void* stream)
{
    g_last_error.clear();
    int rc = validate_batch(1, variant, items_host, n);
    if (rc != VP_OK || n == 0) return rc;
    if (items_dev == nullptr) return fail(VP_ERR_ARG, "null device descriptor array");
    hipError_t e = vp::launch_backward_batched(
        variant, reinterpret_cast<const vp::BatchItem*>(items_host),
        reinterpret_cast<const vp::BatchItem*>(items_dev), n, grid, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "batched backward launch");
    return VP_OK;
}

}  // extern "C"
