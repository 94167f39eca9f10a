// Internal launch interface between the C-ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vp {

struct Geom {
    int64_t SY, SX, H;  // volume extents (z_order independent)
    int64_t PX;         // detector pixels
    int A;              // number of angles (sinogram angle extent)
    int bits;           // fractional bits of the cos/sin hi split
    const float* du = nullptr;  // per-angle detector translation (pixels), or null
};

// Arguments of one projector instance as the per-block device code sees them.
struct OpArgs {
    const float* in;
    float* out;
    const float* angles;
    int A, SY, SX, H, PX;
    int accumulate;  // 1: out += op(in), 0: out = op(in)
    double scale, inv_scale;  // 2^B and 2^-B of the cos/sin hi split
    const float* du;          // per-angle detector translation (pixels), or null
    int out_z2;               // AT, 4 slices per lane: out is (H, SY, SX) instead of (SY, SX, H)
    // Slices h0 .. h0 + hn - 1 of the H (hn = 0: h0 .. H - 1) are the ones
    // this launch covers; strides and bounds still use H.  Honoured by AT's
    // 4-slice and row kernels (single launches split a ragged H between
    // them); every other launch passes 0, 0.
    int h0, hn;
};

// One instance of a batched launch (layout shared with the C ABI's vp_batch_item).
struct BatchItem {
    const float* in;
    float* out;
    const float* angles;
    int32_t A, SY, SX, H, PX;
    int32_t accumulate;
};

// Kernel variants (the autotuner's search space): 1, 2, 4 = slices per lane
// (blocks of 64, 128, 256 slices); for AT also 32, 64, 128 = blocks of that
// many slices at 4 slices per lane, a wave's lanes split into rows of 8, 16,
// 32 lanes that hold one voxel each; for A also 416, 418, 420 = 4 slices per
// lane with 16-, 18- and 20-pixel strips (4 itself: 22 pixels), and 32 = 1
// slice per lane with two angles per sweep wave.
constexpr int kForwardVariants[] = {1, 2, 4, 416, 418, 420, 32};
constexpr int kBackwardVariants[] = {1, 2, 4, 32, 64, 128};
int default_variant(int op, int64_t H);  // op 0 = A, 1 = AT
// Variant of a single (non-batched) launch: the H-based default, except that
// 2 slices per lane drops to 1 for A when its grid has fewer workgroups than
// the device has CUs, and that A's 4 slices per lane takes the strip width
// whose grid fills the device best (forward_strip_variant).
int single_variant(int op, int A, int64_t SY, int64_t SX, int64_t H, int64_t PX);
// A at 4 slices per lane: 4, 416, 418 or 420, the strip width P that
// minimises rounds of resident workgroups x (P + 4), the per-workgroup cost
// measured to grow about as P + 4 (one workgroup per CU).
int forward_strip_variant(int A, int64_t H, int64_t PX);
int device_cus();  // compute units of the current device (cached per device)
// Work items (blocks) of one instance under a variant; -1 for an unknown variant.
int64_t forward_blocks(int variant, int A, int64_t H, int64_t PX);
int64_t backward_blocks(int variant, int64_t SY, int64_t SX, int64_t H);

// Persistent launch over n instances (z_order 0); items_dev holds the same
// descriptors as items_host.
hipError_t launch_forward_batched(int variant, const BatchItem* items_host,
                                  const BatchItem* items_dev, int n, int grid, hipStream_t s);
hipError_t launch_backward_batched(int variant, const BatchItem* items_host,
                                   const BatchItem* items_dev, int n, int grid, hipStream_t s);

// x (+)= AT y in z_order 0 (accumulate != 0 adds, else overwrites).
hipError_t launch_backward_z0(const float* y, float* x, const float* angles, const Geom& g,
                              int accumulate, hipStream_t s);
// x += AT y with y in z_order 0 and x in z_order 2 (H, SY, SX), written
// directly by the 4-slice kernel; backward_writes_z2 tells whether that is
// the single launch's kernel for g (else the row kernels, with transposes,
// are faster).
bool backward_writes_z2(const Geom& g);
hipError_t launch_backward_to_z2(const float* y, float* x, const float* angles, const Geom& g,
                                 hipStream_t s);

// y (+)= A x in z_order 0 (accumulate != 0 adds, else overwrites).
hipError_t launch_forward_z0(const float* x, float* y, const float* angles, const Geom& g,
                             int accumulate, hipStream_t s);

// FP64 operator pair, either layout (outputs accumulate).
hipError_t launch_forward_f64(int z_order, const double* x, double* y, const float* angles,
                              const Geom& g, hipStream_t s);
hipError_t launch_backward_f64(int z_order, const double* y, double* x, const float* angles,
                               const Geom& g, hipStream_t s);

// Linear interpolation along H of a sinogram, per angle (detector translation
// dv along H, in slices).  Forward (transpose = 0): out at global slice i
// samples in at global position i + dv[a]; transpose = 1 applies the adjoint.
// in / out hold slices [in_off, in_off + Hin) / [out_off, out_off + Hout) of
// the full sinogram; positions outside the input slab read 0.  z_order 0:
// (A, PX, H); 2: (H, A, PX).  Outputs accumulate when accumulate != 0.
template <typename T>
hipError_t launch_vshift(int z_order, int transpose, const T* in, T* out, int A, int64_t PX,
                         int64_t Hin, int64_t in_off, int64_t Hout, int64_t out_off,
                         const float* dv, int accumulate, hipStream_t s);

// p[0..n) = v (stream-ordered fill kernel; usable under graph capture).
template <typename T>
hipError_t launch_fill(T* p, int64_t n, T v, hipStream_t s);

// out[c][r] (+)= in[r][c] for an R x C row-major matrix.
hipError_t launch_transpose(const float* in, float* out, int64_t R, int64_t C, int accumulate,
                            hipStream_t s);

// One matrix of a batched transpose (layout shared with the C ABI's
// vp_transpose_item).
struct TransposeItem {
    const float* in;
    float* out;
    int64_t R, C;
    int32_t accumulate, reserved;
};

// out[c][r] (+)= in[r][c] for each of n matrices in one persistent launch;
// items_dev holds the same descriptors as items_host.
hipError_t launch_transpose_batched(const TransposeItem* items_host, const TransposeItem* items_dev,
                                    int n, int grid, hipStream_t s);

}  // namespace vp
