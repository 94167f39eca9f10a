// C ABI of the projector library (declared in include/voxelproj_hip.h).
// Validation happens here, before anything touches the device; kernels are
// launched stream-ordered on the caller's stream and every launch is checked.
#include <stddef.h>
#include <stdio.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <utility>

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

static_assert(sizeof(vp_transpose_item) == sizeof(vp::TransposeItem), "transpose item layout");
static_assert(offsetof(vp_transpose_item, rows) == offsetof(vp::TransposeItem, R), "transpose item layout");
static_assert(offsetof(vp_transpose_item, accumulate) == offsetof(vp::TransposeItem, accumulate),
              "transpose item layout");

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
        // per DMA instruction (variant 32: 8) through one 32-bit buffer offset.
        if (op == 0 && (variant == 1 || variant == 2 || variant == 32)) {
            const int64_t per = variant == 32 ? 8 : 4 / variant;  // voxels per instruction
            const int64_t span = (per - 1) * it.SX * it.H + it.H;
            if (span >= ((int64_t)1 << 29))
                return fail(VP_ERR_LIMIT, "batch item " + std::to_string(i) + ": SX * H too large for variant " +
                                              std::to_string(variant) + " (use variant 4)");
        }
    }
    return VP_OK;
}

}  // namespace

extern "C" {

int vp_abi_version(void) { return 4; }

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

// ---------------------------------------------------------------- workspace
// Device memory beyond the inputs and outputs: the layout transposes of
// z_order 2 and the untranslated sinogram of the dv path.  The caller passes
// it (vp_forward_ws / vp_backward_ws: memory the caller owns, for instance
// from its framework's allocator), or the library keeps one buffer per
// (device, stream) that grows only when a call needs more, so steady-state
// calls allocate nothing.  The cache's lock is held for the whole call, so a
// concurrent call on the same stream cannot regrow (free) the buffer between
// another call's use of it and the enqueueing of that call's kernels.
struct Span {
    void* p = nullptr;
    int64_t bytes = 0;
};

std::mutex g_ws_mutex;
std::map<std::pair<int, hipStream_t>, Span> g_ws;

// Buffer of at least `bytes` for stream s on the current device (caller
// holds g_ws_mutex).
hipError_t cached_workspace(int64_t bytes, hipStream_t s, void** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    Span& b = g_ws[{dev, s}];
    if (b.bytes < bytes) {
        if (b.p) {
            e = hipFreeAsync(b.p, s);  // after the stream's earlier uses
            b.p = nullptr;
            b.bytes = 0;
            if (e != hipSuccess) return e;
        }
        const int64_t nb = std::max(bytes, (int64_t)256);
        e = hipMallocAsync(&b.p, (size_t)nb, s);
        if (e != hipSuccess) {
            b.p = nullptr;
            return e;
        }
        b.bytes = nb;
    }
    *out = b.p;
    return hipSuccess;
}

// Workspace (elements of T) that a call of operator op (0 = A, 1 = AT, -1 =
// the larger of the two) needs: z_order 2 transposes the volume and the
// sinogram (|x| + |y| at full H, or H slabs of a smaller buffer), except that
// AT's 4-slice kernel writes a z_order 2 volume itself (|y|); dv needs the
// untranslated sinogram (|y|, and with z_order 2 a second one).  Float32
// paths only use the z_order 0 kernels; FP64 has native z_order 2 kernels and
// needs a workspace only for dv.
int64_t workspace_elems(bool f64, int z_order, const vp::Geom& g, bool dv, int op)
{
    const int64_t nx = g.SY * g.SX * g.H;
    const int64_t ny = (int64_t)g.A * g.PX * g.H;
    if (f64) return dv ? ny : 0;
    const int64_t vol = (op == 1 && vp::backward_writes_z2(g)) ? 0 : nx;
    if (!dv) return z_order == 2 ? vol + ny : 0;
    return z_order == 2 ? vol + 2 * ny : ny;
}

// Minimum workspace of the slab path: one slice of volume and sinogram.
int64_t slab_min_elems(const vp::Geom& g) { return g.SY * g.SX + (int64_t)g.A * g.PX; }

// out (+)= op(in) in z_order 2 without translations, through H slabs of the
// workspace (z_order 2 slabs are contiguous): the input slab is transposed
// into slice-fastest order, the z_order 0 kernel runs on it, and its result
// is transposed and added into the output slab -- except where the
// backprojection's 4-slice kernel adds its sums straight into the z_order 2
// volume (vp::backward_writes_z2), which needs no volume workspace.  Slab
// heights are multiples of 256 slices (a whole slice block of every kernel
// variant) unless the whole H fits or fewer than 256 slices do.
hipError_t z2_by_slabs(int op, const float* in, float* out, const float* angles,
                       const vp::Geom& g, float* ws, int64_t ws_elems, hipStream_t s)
{
    const int64_t nv = g.SY * g.SX, ns = (int64_t)g.A * g.PX;  // floats per slice
    const bool direct = op == 1 && vp::backward_writes_z2(g);
    int64_t hs = ws_elems / (direct ? ns : nv + ns);
    if (hs >= g.H)
        hs = g.H;
    else if (hs >= 256)
        hs -= hs % 256;
    hipError_t e = hipSuccess;
    for (int64_t h0 = 0; h0 < g.H && e == hipSuccess; h0 += hs) {
        vp::Geom gs = g;
        gs.H = std::min(hs, g.H - h0);
        float* wx = ws;                // (SY, SX, hs)
        float* wy = ws + nv * gs.H;    // (A, PX, hs)
        if (op == 0) {
            e = vp::launch_transpose(in + h0 * nv, wx, gs.H, nv, 0, s);
            if (e == hipSuccess) e = vp::launch_forward_z0(wx, wy, angles, gs, 0, s);
            if (e == hipSuccess) e = vp::launch_transpose(wy, out + h0 * ns, ns, gs.H, 1, s);
        } else if (direct) {
            e = vp::launch_transpose(in + h0 * ns, ws, gs.H, ns, 0, s);
            if (e == hipSuccess) e = vp::launch_backward_to_z2(ws, out + h0 * nv, angles, gs, s);
        } else {
            e = vp::launch_transpose(in + h0 * ns, wy, gs.H, ns, 0, s);
            if (e == hipSuccess) e = vp::launch_backward_z0(wy, wx, angles, gs, 0, s);
            if (e == hipSuccess) e = vp::launch_transpose(wx, out + h0 * nv, nv, gs.H, 1, s);
        }
    }
    return e;
}

// y += A x (float32), any layout, optional translations; ws holds ws_elems
// floats (see workspace_elems; z_order 2 without dv also runs in slabs).
hipError_t forward_f32(int z_order, const float* x, float* y, const float* angles,
                       const vp::Geom& g, const float* dv, float* ws, int64_t ws_elems,
                       hipStream_t s)
{
    const int64_t nx = g.SY * g.SX * g.H;
    const int64_t ny = (int64_t)g.A * g.PX * g.H;
    if (!dv) {
        if (z_order == 0) return vp::launch_forward_z0(x, y, angles, g, 1, s);
        return z2_by_slabs(0, x, y, angles, g, ws, ws_elems, s);
    }
    // dv: the untranslated sinogram y0 (slice-fastest), then its H interpolation
    hipError_t e = hipSuccess;
    const float* x0 = x;
    float* y0 = ws;
    if (z_order == 2) {
        e = vp::launch_transpose(x, ws, g.H, g.SY * g.SX, 0, s);
        x0 = ws;
        y0 = ws + nx;
    }
    if (e == hipSuccess) e = vp::launch_forward_z0(x0, y0, angles, g, 0, s);
    if (z_order == 0) {
        if (e == hipSuccess)
            e = vp::launch_vshift<float>(0, 0, y0, y, g.A, g.PX, g.H, 0, g.H, 0, dv, 1, s);
        return e;
    }
    if (e == hipSuccess)
        e = vp::launch_vshift<float>(0, 0, y0, y0 + ny, g.A, g.PX, g.H, 0, g.H, 0, dv, 0, s);
    if (e == hipSuccess) e = vp::launch_transpose(y0 + ny, y, (int64_t)g.A * g.PX, g.H, 1, s);
    return e;
}

// x += AT y (float32), any layout, optional translations.
hipError_t backward_f32(int z_order, const float* y, float* x, const float* angles,
                        const vp::Geom& g, const float* dv, float* ws, int64_t ws_elems,
                        hipStream_t s)
{
    const int64_t nx = g.SY * g.SX * g.H;
    const int64_t ny = (int64_t)g.A * g.PX * g.H;
    if (!dv) {
        if (z_order == 0) return vp::launch_backward_z0(y, x, angles, g, 1, s);
        return z2_by_slabs(1, y, x, angles, g, ws, ws_elems, s);
    }
    // dv: the adjoint H interpolation first, then AT of the untranslated
    // sinogram.  Workspace: [the volume before its transpose (z_order 2, not
    // when the kernel writes z_order 2 itself)] [transposed sinogram
    // (z_order 2)] [interpolated sinogram].
    hipError_t e = hipSuccess;
    const bool direct = z_order == 2 && vp::backward_writes_z2(g);
    const float* yz = y;  // sinogram in slice-fastest order
    float* x0 = ws;
    float* yb = (z_order == 2 && !direct) ? ws + nx : ws;
    if (z_order == 2) {
        e = vp::launch_transpose(y, yb, g.H, (int64_t)g.A * g.PX, 0, s);
        yz = yb;
    }
    float* g0 = z_order == 2 ? yb + ny : yb;
    if (e == hipSuccess) e = vp::launch_vshift<float>(0, 1, yz, g0, g.A, g.PX, g.H, 0, g.H, 0, dv, 0, s);
    if (z_order == 0) {
        if (e == hipSuccess) e = vp::launch_backward_z0(g0, x, angles, g, 1, s);
        return e;
    }
    if (direct) {
        if (e == hipSuccess) e = vp::launch_backward_to_z2(g0, x, angles, g, s);
        return e;
    }
    if (e == hipSuccess) e = vp::launch_backward_z0(g0, x0, angles, g, 0, s);
    if (e == hipSuccess) e = vp::launch_transpose(x0, x, g.SY * g.SX, g.H, 1, s);
    return e;
}

// y += A x (float64).  With dv the untranslated sinogram goes to a zeroed
// workspace first (the FP64 kernels accumulate).
hipError_t forward_f64(int z_order, const double* x, double* y, const float* angles,
                       const vp::Geom& g, const float* dv, double* ws, int64_t /*ws_elems*/,
                       hipStream_t s)
{
    if (!dv) return vp::launch_forward_f64(z_order, x, y, angles, g, s);
    const int64_t ny = (int64_t)g.A * g.PX * g.H;
    hipError_t e = vp::launch_fill<double>(ws, ny, 0.0, s);
    if (e == hipSuccess) e = vp::launch_forward_f64(z_order, x, ws, angles, g, s);
    if (e == hipSuccess)
        e = vp::launch_vshift<double>(z_order, 0, ws, y, g.A, g.PX, g.H, 0, g.H, 0, dv, 1, s);
    return e;
}

hipError_t backward_f64(int z_order, const double* y, double* x, const float* angles,
                        const vp::Geom& g, const float* dv, double* ws, int64_t /*ws_elems*/,
                        hipStream_t s)
{
    if (!dv) return vp::launch_backward_f64(z_order, y, x, angles, g, s);
    hipError_t e = vp::launch_vshift<double>(z_order, 1, y, ws, g.A, g.PX, g.H, 0, g.H, 0, dv, 0, s);
    if (e == hipSuccess) e = vp::launch_backward_f64(z_order, ws, x, angles, g, s);
    return e;
}

template <typename T, typename Op>
int run_op(Op op, int which, const char* what, int z_order, const T* in, T* out, const float* angles,
           int n_angles, int64_t SY, int64_t SX, int64_t H, int64_t PX, const float* du,
           const float* dv, double du_max, T* ws, int64_t ws_elems, void* stream)
{
    g_last_error.clear();
    vp::Geom g;
    bool empty = false;
    int rc = validate(z_order, in, out, angles, n_angles, SY, SX, H, PX, &g, &empty);
    if (rc != VP_OK || empty) return rc;
    rc = validate_shift(du, du_max, &g);
    if (rc != VP_OK) return rc;
    if (ws_elems < 0 || (ws_elems > 0 && ws == nullptr))
        return fail(VP_ERR_ARG, "workspace: negative size or null pointer");
    const bool f64 = sizeof(T) == 8;
    const int64_t need = workspace_elems(f64, z_order, g, dv != nullptr, which);
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if (need == 0) {
        e = op(z_order, in, out, angles, g, dv, nullptr, 0, s);
    } else if (ws != nullptr) {
        // a caller's workspace: z_order 2 without dv runs in H slabs of it
        const int64_t least = (!f64 && !dv) ? slab_min_elems(g) : need;
        if (ws_elems < least)
            return fail(VP_ERR_LIMIT, std::string(what) + ": workspace of " + std::to_string(ws_elems) +
                                          " elements, this call needs at least " + std::to_string(least));
        e = op(z_order, in, out, angles, g, dv, ws, ws_elems, s);
    } else {
        std::lock_guard<std::mutex> lock(g_ws_mutex);
        void* p = nullptr;
        e = cached_workspace(need * (int64_t)sizeof(T), s, &p);
        if (e != hipSuccess) return hip_fail(e, (std::string(what) + " workspace").c_str());
        e = op(z_order, in, out, angles, g, dv, (T*)p, need, s);
    }
    if (e != hipSuccess) return hip_fail(e, (std::string(what) + " launch").c_str());
    return VP_OK;
}

// ------------------------------------------------ allocation-free variant
// Scratch of vp_project_noalloc: one file-scope device array, allocated
// with the code object on each device the library runs on and held until
// the process ends.  Host code never uses the array's name as an address:
// its device address is looked up once per device (hipGetSymbolAddress,
// with that device current) and kept.  Every call on a device shares that
// device's array, so calls that overlap on one device (different streams,
// or a captured graph replayed beside a call) race on it.  Nothing in it
// needs to start at zero: each slab's transposes and kernel overwrite what
// they later read.
constexpr int64_t kScratchElems = (int64_t)16 << 20;  // 16 Mi floats = 64 MiB
__device__ __align__(256) float scratch[kScratchElems];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};

hipError_t scratch_of_current_device(float** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

extern "C" {

int vp_forward(int z_order, const float* x, float* y, const float* angles, int n_angles,
               int64_t SY, int64_t SX, int64_t H, int64_t PX, void* stream)
{
    return run_op<float>(forward_f32, 0, "forward", z_order, x, y, angles, n_angles, SY, SX, H, PX,
                         nullptr, nullptr, 0.0, nullptr, 0, stream);
}

int vp_backward(int z_order, const float* y, float* x, const float* angles, int n_angles,
                int64_t SY, int64_t SX, int64_t H, int64_t PX, void* stream)
{
    return run_op<float>(backward_f32, 1, "backward", z_order, y, x, angles, n_angles, SY, SX, H, PX,
                         nullptr, nullptr, 0.0, nullptr, 0, stream);
}

int vp_forward_f64(int z_order, const double* x, double* y, const float* angles, int n_angles,
                   int64_t SY, int64_t SX, int64_t H, int64_t PX, void* stream)
{
    return run_op<double>(forward_f64, 0, "forward f64", z_order, x, y, angles, n_angles, SY, SX, H,
                          PX, nullptr, nullptr, 0.0, nullptr, 0, stream);
}

int vp_backward_f64(int z_order, const double* y, double* x, const float* angles, int n_angles,
                    int64_t SY, int64_t SX, int64_t H, int64_t PX, void* stream)
{
    return run_op<double>(backward_f64, 1, "backward f64", z_order, y, x, angles, n_angles, SY, SX, H,
                          PX, nullptr, nullptr, 0.0, nullptr, 0, stream);
}

int vp_forward_geom(int z_order, const float* x, float* y, const float* angles, int n_angles,
                    int64_t SY, int64_t SX, int64_t H, int64_t PX, const float* du,
                    const float* dv, double du_max, void* stream)
{
    return run_op<float>(forward_f32, 0, "forward", z_order, x, y, angles, n_angles, SY, SX, H, PX,
                         du, dv, du_max, nullptr, 0, stream);
}

int vp_backward_geom(int z_order, const float* y, float* x, const float* angles, int n_angles,
                     int64_t SY, int64_t SX, int64_t H, int64_t PX, const float* du,
                     const float* dv, double du_max, void* stream)
{
    return run_op<float>(backward_f32, 1, "backward", z_order, y, x, angles, n_angles, SY, SX, H, PX,
                         du, dv, du_max, nullptr, 0, stream);
}

int vp_forward_geom_f64(int z_order, const double* x, double* y, const float* angles,
                        int n_angles, int64_t SY, int64_t SX, int64_t H, int64_t PX,
                        const float* du, const float* dv, double du_max, void* stream)
{
    return run_op<double>(forward_f64, 0, "forward f64", z_order, x, y, angles, n_angles, SY, SX, H,
                          PX, du, dv, du_max, nullptr, 0, stream);
}

int vp_backward_geom_f64(int z_order, const double* y, double* x, const float* angles,
                         int n_angles, int64_t SY, int64_t SX, int64_t H, int64_t PX,
                         const float* du, const float* dv, double du_max, void* stream)
{
    return run_op<double>(backward_f64, 1, "backward f64", z_order, y, x, angles, n_angles, SY, SX, H,
                          PX, du, dv, du_max, nullptr, 0, stream);
}

int vp_forward_ws(int z_order, const float* x, float* y, const float* angles, int n_angles,
                  int64_t SY, int64_t SX, int64_t H, int64_t PX, const float* du, const float* dv,
                  double du_max, float* workspace, int64_t workspace_elems, void* stream)
{
    return run_op<float>(forward_f32, 0, "forward", z_order, x, y, angles, n_angles, SY, SX, H, PX,
                         du, dv, du_max, workspace, workspace_elems, stream);
}

int vp_backward_ws(int z_order, const float* y, float* x, const float* angles, int n_angles,
                   int64_t SY, int64_t SX, int64_t H, int64_t PX, const float* du, const float* dv,
                   double du_max, float* workspace, int64_t workspace_elems, void* stream)
{
    return run_op<float>(backward_f32, 1, "backward", z_order, y, x, angles, n_angles, SY, SX, H, PX,
                         du, dv, du_max, workspace, workspace_elems, stream);
}

int64_t vp_workspace_elems(int f64, int z_order, int n_angles, int64_t SY, int64_t SX, int64_t H,
                           int64_t PX, int with_dv)
{
    if ((z_order != 0 && z_order != 2) || n_angles < 0 || SY < 0 || SX < 0 || H < 0 || PX < 0)
        return -1;
    vp::Geom g;
    g.SY = SY;
    g.SX = SX;
    g.H = H;
    g.PX = PX;
    g.A = n_angles;
    return workspace_elems(f64 != 0, z_order, g, with_dv != 0, -1);
}

int vp_release_workspaces(void)
{
    g_last_error.clear();
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return hip_fail(e, "release workspaces");
    int rc = VP_OK;
    for (auto& kv : g_ws) {
        if (!kv.second.p) continue;
        e = hipSetDevice(kv.first.first);
        if (e == hipSuccess) e = hipStreamSynchronize(kv.first.second);
        if (e == hipSuccess) e = hipFree(kv.second.p);
        if (e != hipSuccess && rc == VP_OK) rc = hip_fail(e, "release workspaces");
        kv.second = Span();
    }
    g_ws.clear();
    e = hipSetDevice(cur);
    if (e != hipSuccess && rc == VP_OK) rc = hip_fail(e, "release workspaces");
    return rc;
}

int vp_project_noalloc(int op, int z_order, const float* in, float* out, const float* angles,
                       int n_angles, int64_t SY, int64_t SX, int64_t H, int64_t PX, void* stream)
{
    g_last_error.clear();
    if (op != 0 && op != 1) return fail(VP_ERR_ARG, "op must be 0 (A) or 1 (AT), got " + std::to_string(op));
    vp::Geom g;
    bool empty = false;
    int rc = validate(z_order, in, out, angles, n_angles, SY, SX, H, PX, &g, &empty);
    if (rc != VP_OK || empty) return rc;
    if (z_order == 2 && slab_min_elems(g) > kScratchElems)
        return fail(VP_ERR_LIMIT, "vp_project_noalloc: one slice of volume and sinogram (SY*SX + "
                                  "n_angles*PX = " + std::to_string(slab_min_elems(g)) +
                                  " floats) exceeds the scratch of " + std::to_string(kScratchElems) +
                                  " floats");
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if (z_order == 0) {
        e = op == 0 ? vp::launch_forward_z0(in, out, angles, g, 1, s)
                    : vp::launch_backward_z0(in, out, angles, g, 1, s);
    } else {
        float* ws = nullptr;
        e = scratch_of_current_device(&ws);
        if (e != hipSuccess) return hip_fail(e, "vp_project_noalloc scratch");
        e = z2_by_slabs(op, in, out, angles, g, ws, kScratchElems, s);
    }
    if (e != hipSuccess) return hip_fail(e, "vp_project_noalloc launch");
    return VP_OK;
}

int64_t vp_noalloc_scratch_elems(void) { return kScratchElems; }

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

int vp_launch_variant(int op, int n_angles, int64_t SY, int64_t SX, int64_t H, int64_t PX)
{
    if ((op != 0 && op != 1) || n_angles <= 0 || SY <= 0 || SX <= 0 || H <= 0 || PX <= 0) return -1;
    return vp::single_variant(op, n_angles, SY, SX, H, PX);
}

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

int vp_transpose_batched(const vp_transpose_item* items_host, const vp_transpose_item* items_dev,
                         int n, int grid, void* stream)
{
    g_last_error.clear();
    if (n < 0) return fail(VP_ERR_ARG, "negative batch size");
    if (n == 0) return VP_OK;
    if (items_host == nullptr || items_dev == nullptr) return fail(VP_ERR_ARG, "null descriptor array");
    for (int i = 0; i < n; ++i) {
        const vp_transpose_item& it = items_host[i];
        const std::string who = "transpose item " + std::to_string(i);
        if (it.rows < 0 || it.cols < 0) return fail(VP_ERR_ARG, who + ": negative extent");
        if (it.rows > ((int64_t)1 << 40) / std::max<int64_t>(it.cols, 1))
            return fail(VP_ERR_LIMIT, who + ": matrix too large");
        if (it.rows * it.cols > 0 && (it.in == nullptr || it.out == nullptr))
            return fail(VP_ERR_ARG, who + ": null device pointer");
        if (it.accumulate != 0 && it.accumulate != 1) return fail(VP_ERR_ARG, who + ": accumulate must be 0 or 1");
        if (it.reserved != 0) return fail(VP_ERR_ARG, who + ": reserved field must be 0");
    }
    hipError_t e = vp::launch_transpose_batched(reinterpret_cast<const vp::TransposeItem*>(items_host),
                                                reinterpret_cast<const vp::TransposeItem*>(items_dev), n, grid,
                                                (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "batched transpose launch");
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
