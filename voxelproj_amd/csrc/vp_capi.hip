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
    }
    return VP_OK;
}

}  // namespace

extern "C" {

int vp_abi_version(void) { return 1; }

const char* vp_last_error(void) { return g_last_error.c_str(); }

int vp_forward(int z_order, const float* x, float* y, const float* angles, int n_angles,
               int64_t SY, int64_t SX, int64_t H, int64_t PX, void* stream)
{
    g_last_error.clear();
    vp::Geom g;
    bool empty = false;
    int rc = validate(z_order, x, y, angles, n_angles, SY, SX, H, PX, &g, &empty);
    if (rc != VP_OK || empty) return rc;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if (z_order == 0) {
        e = vp::launch_forward_z0(x, y, angles, g, 1, s);
        if (e != hipSuccess) return hip_fail(e, "forward z_order 0 launch");
        return VP_OK;
    }
    // z_order 2: transpose the volume to slice-fastest order, project, and
    // add the transposed sinogram back.
    const int64_t nx = SY * SX * H;
    const int64_t ny = (int64_t)n_angles * PX * H;
    float* ws = nullptr;
    e = hipMallocAsync((void**)&ws, (size_t)(nx + ny) * sizeof(float), s);
    if (e != hipSuccess) return hip_fail(e, "forward z_order 2 workspace");
    float* x0 = ws;
    float* y0 = ws + nx;
    e = vp::launch_transpose(x, x0, H, SY * SX, 0, s);
    if (e == hipSuccess) e = vp::launch_forward_z0(x0, y0, angles, g, 0, s);
    if (e == hipSuccess) e = vp::launch_transpose(y0, y, (int64_t)n_angles * PX, H, 1, s);
    hipError_t ef = hipFreeAsync(ws, s);
    if (e != hipSuccess) return hip_fail(e, "forward z_order 2 launch");
    if (ef != hipSuccess) return hip_fail(ef, "forward z_order 2 workspace release");
    return VP_OK;
}

int vp_backward(int z_order, const float* y, float* x, const float* angles, int n_angles,
                int64_t SY, int64_t SX, int64_t H, int64_t PX, void* stream)
{
    g_last_error.clear();
    vp::Geom g;
    bool empty = false;
    int rc = validate(z_order, y, x, angles, n_angles, SY, SX, H, PX, &g, &empty);
    if (rc != VP_OK || empty) return rc;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if (z_order == 0) {
        e = vp::launch_backward_z0(y, x, angles, g, 1, s);
        if (e != hipSuccess) return hip_fail(e, "backward z_order 0 launch");
        return VP_OK;
    }
    // z_order 2: transpose the sinogram to slice-fastest order, backproject,
    // and add the transposed volume back.
    const int64_t nx = SY * SX * H;
    const int64_t ny = (int64_t)n_angles * PX * H;
    float* ws = nullptr;
    e = hipMallocAsync((void**)&ws, (size_t)(nx + ny) * sizeof(float), s);
    if (e != hipSuccess) return hip_fail(e, "backward z_order 2 workspace");
    float* x0 = ws;
    float* y0 = ws + nx;
    e = vp::launch_transpose(y, y0, H, (int64_t)n_angles * PX, 0, s);
    if (e == hipSuccess) e = vp::launch_backward_z0(y0, x0, angles, g, 0, s);
    if (e == hipSuccess) e = vp::launch_transpose(x0, x, SY * SX, H, 1, s);
    hipError_t ef = hipFreeAsync(ws, s);
    if (e != hipSuccess) return hip_fail(e, "backward z_order 2 launch");
    if (ef != hipSuccess) return hip_fail(ef, "backward z_order 2 workspace release");
    return VP_OK;
}

int vp_forward_f64(int z_order, const double* x, double* y, const float* angles, int n_angles,
                   int64_t SY, int64_t SX, int64_t H, int64_t PX, void* stream)
{
    g_last_error.clear();
    vp::Geom g;
    bool empty = false;
    int rc = validate(z_order, x, y, angles, n_angles, SY, SX, H, PX, &g, &empty);
    if (rc != VP_OK || empty) return rc;
    hipError_t e = vp::launch_forward_f64(z_order, x, y, angles, g, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "forward f64 launch");
    return VP_OK;
}

int vp_backward_f64(int z_order, const double* y, double* x, const float* angles, int n_angles,
                    int64_t SY, int64_t SX, int64_t H, int64_t PX, void* stream)
{
    g_last_error.clear();
    vp::Geom g;
    bool empty = false;
    int rc = validate(z_order, y, x, angles, n_angles, SY, SX, H, PX, &g, &empty);
    if (rc != VP_OK || empty) return rc;
    hipError_t e = vp::launch_backward_f64(z_order, y, x, angles, g, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "backward f64 launch");
    return VP_OK;
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
