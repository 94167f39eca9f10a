/*
 * voxelproj_hip -- C ABI of the MI355X (gfx950) exact-trapezoid parallel-beam
 * projector pair.  Plain pointers and sizes; no framework types.
 *
 * Geometry: voxel = detector pixel = 1, rotation centre at the volume
 * centre, detector centre on the rotation axis.  All arrays are float32,
 * C-contiguous, resident in device memory of the current HIP device.
 *
 *   z_order 0:  volume x (SY, SX, H),  sinogram y (A, PX, H)   -- H fastest
 *   z_order 2:  volume x (H, SY, SX),  sinogram y (H, A, PX)   -- H slowest
 *
 * Every entry point ACCUMULATES into its output (out += op(in)), as the
 * reference kernels do, and is stream-ordered on `stream` (a hipStream_t;
 * NULL = the null stream).  Calls return 0 on success or a negative status;
 * vp_last_error() then describes the failure (per calling thread).
 */
#ifndef VOXELPROJ_HIP_H
#define VOXELPROJ_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VP_OK 0
#define VP_ERR_ARG -1     /* invalid argument (z_order, extents, null pointer) */
#define VP_ERR_LIMIT -2   /* extents beyond what the kernels support */
#define VP_ERR_HIP -3     /* HIP runtime error (launch, allocation) */

#if defined(__GNUC__)
#define VP_API __attribute__((visibility("default")))
#else
#define VP_API
#endif

/*
 * Forward projection  y += A x.
 * Replaces the reference binding _CU.forward_z0 / _CU.forward_z2
 * (voxelproj/csrc/voxelproj.cpp:15-18,25-28) and the host launchers
 * forward_z0 / forward_z2 (voxelproj/csrc/projectors.cu:34-104, 171-234).
 * `angles` holds n_angles float32 radians on the device; n_angles is the
 * sinogram's angle extent.
 */
VP_API int vp_forward(int z_order, const float *x, float *y, const float *angles, int n_angles,
               int64_t SY, int64_t SX, int64_t H, int64_t PX, void *stream);

/*
 * Backprojection  x += AT y  (exact adjoint of vp_forward).
 * Replaces _CU.backward_z0 / _CU.backward_z2 (voxelproj/csrc/voxelproj.cpp:20-23,30-33)
 * and the host launchers backward_z0 / backward_z2
 * (voxelproj/csrc/projectors.cu:106-169, 236-299).
 */
VP_API int vp_backward(int z_order, const float *y, float *x, const float *angles, int n_angles,
                int64_t SY, int64_t SX, int64_t H, int64_t PX, void *stream);

/*
 * Double-precision variants of vp_forward / vp_backward (same layouts and
 * accumulate semantics, float64 data, float32 angles).  No counterpart in
 * the reference, which is float32-only (voxelproj/csrc/projectors.cu:37);
 * they serve FP64 gradient checks and solvers.
 */
VP_API int vp_forward_f64(int z_order, const double *x, double *y, const float *angles, int n_angles,
                          int64_t SY, int64_t SX, int64_t H, int64_t PX, void *stream);
VP_API int vp_backward_f64(int z_order, const double *y, double *x, const float *angles, int n_angles,
                           int64_t SY, int64_t SX, int64_t H, int64_t PX, void *stream);

/*
 * Translated detector geometry (the reference lists detector translations as
 * future work: README.md:65-68; no counterpart in its code).  Same as
 * vp_forward / vp_backward and the _f64 variants, plus two optional device
 * arrays of n_angles float32 values (NULL = no translation):
 *   du: detector translation along the pixel axis, in pixels; pixel p of angle
 *       a then sees u = (ix-xc) cos + (iy-yc) sin + (PX-1)/2 - du[a];
 *   dv: detector translation along H, in slices; sinogram slice h of angle a
 *       then samples the volume at slice position h + dv[a] (linear
 *       interpolation between slices, zero outside [0, H)).
 * du_max must bound max |du[a]| (it sizes the exact position split; ignored
 * when du is NULL).  The backprojector is the exact adjoint.  A translation
 * (dx, dy) of the volume in its plane equals du[a] = -(dx cos a + dy sin a).
 */
VP_API int vp_forward_geom(int z_order, const float *x, float *y, const float *angles, int n_angles,
                           int64_t SY, int64_t SX, int64_t H, int64_t PX, const float *du,
                           const float *dv, double du_max, void *stream);
VP_API int vp_backward_geom(int z_order, const float *y, float *x, const float *angles, int n_angles,
                            int64_t SY, int64_t SX, int64_t H, int64_t PX, const float *du,
                            const float *dv, double du_max, void *stream);
VP_API int vp_forward_geom_f64(int z_order, const double *x, double *y, const float *angles,
                               int n_angles, int64_t SY, int64_t SX, int64_t H, int64_t PX,
                               const float *du, const float *dv, double du_max, void *stream);
VP_API int vp_backward_geom_f64(int z_order, const double *y, double *x, const float *angles,
                                int n_angles, int64_t SY, int64_t SX, int64_t H, int64_t PX,
                                const float *du, const float *dv, double du_max, void *stream);

/*
 * Workspace.  z_order 2 (and dv) need device memory beyond the inputs and
 * outputs: the layout transposes and the untranslated sinogram.  The entry
 * points above take it from a buffer the library keeps per (device, stream),
 * which grows only when a call needs more, so steady-state calls allocate
 * nothing; vp_release_workspaces() frees those buffers (it synchronizes the
 * streams that used them).  vp_forward_ws / vp_backward_ws take the
 * workspace from the caller instead (for example from a framework's caching
 * allocator): `workspace` holds `workspace_elems` floats on the current
 * device, used in stream order on `stream` and free again once the call's
 * work on that stream has run.  z_order 2 without dv processes H in slabs
 * when the workspace holds less than vp_workspace_elems(): it needs at least
 * SY*SX + n_angles*PX floats (one slice).  With dv the full size is needed.
 * No counterpart in the reference, which allocates a texture array per call
 * (voxelproj/csrc/projectors.cu:58,102-103).
 */
VP_API int vp_forward_ws(int z_order, const float *x, float *y, const float *angles, int n_angles,
                         int64_t SY, int64_t SX, int64_t H, int64_t PX, const float *du,
                         const float *dv, double du_max, float *workspace, int64_t workspace_elems,
                         void *stream);
VP_API int vp_backward_ws(int z_order, const float *y, float *x, const float *angles, int n_angles,
                          int64_t SY, int64_t SX, int64_t H, int64_t PX, const float *du,
                          const float *dv, double du_max, float *workspace, int64_t workspace_elems,
                          void *stream);

/* Workspace elements (floats, or doubles when f64 != 0) that one call of the
 * given layout and extents needs without H slabs, for either operator (the
 * larger need: vp_backward in z_order 2 above 128 slices needs no volume
 * part, since its kernel writes the z_order 2 volume itself); 0 when it needs
 * none (z_order 0 without dv); < 0 on invalid input. */
VP_API int64_t vp_workspace_elems(int f64, int z_order, int n_angles, int64_t SY, int64_t SX,
                                  int64_t H, int64_t PX, int with_dv);

/* Free the library's cached workspaces on every device (0 on success). */
VP_API int vp_release_workspaces(void);

/*
 * Allocation-free projector call, for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  No counterpart in the reference.  op 0: out += A in
 * (in = volume, out = sinogram); op 1: out += AT in.  The same kernels and
 * semantics as vp_forward / vp_backward; z_order 2 runs its layout
 * transposes through H slabs of a static device scratch array of
 * vp_noalloc_scratch_elems() floats (64 MiB per device, held until the
 * process ends), so it needs SY*SX + n_angles*PX <= that size and otherwise
 * returns VP_ERR_LIMIT.  Calls that overlap on one device share the scratch
 * and race: serialize them (one stream, or events).
 */
VP_API int vp_project_noalloc(int op, int z_order, const float *in, float *out,
                              const float *angles, int n_angles, int64_t SY, int64_t SX,
                              int64_t H, int64_t PX, void *stream);

/* Floats in vp_project_noalloc's scratch. */
VP_API int64_t vp_noalloc_scratch_elems(void);

/*
 * The H translation on its own, on slabs of a sinogram, for drivers that
 * stream data larger than device memory through the projectors in H slabs
 * (no counterpart in the reference).  `in` holds global slices
 * [in_offset, in_offset + H_in), `out` [out_offset, out_offset + H_out);
 * transpose 0 gives out(h) = interpolation of in at h + dv[a] (the forward
 * translation), transpose 1 its adjoint.  Input slices outside the slab read
 * zero; accumulate 1 adds to out, 0 overwrites it.
 */
VP_API int vp_vshift(int z_order, int transpose, const float *in, float *out, int n_angles,
                     int64_t PX, int64_t H_in, int64_t in_offset, int64_t H_out,
                     int64_t out_offset, const float *dv, int accumulate, void *stream);

/*
 * Batched form: many projector instances of uneven size and angle count in
 * one persistent launch (z_order 0 layout).  No counterpart in the
 * reference, which projects one volume per call.
 *
 * Each instance is described by a vp_batch_item.  The caller passes the
 * descriptor array twice: `items_host` (host memory, used for validation and
 * to size the launch) and `items_dev` (an identical copy in device memory,
 * read by the kernel).  `variant` selects the kernel configuration (1, 2 or 4
 * slices per lane for either operator; 32, 64 or 128 slices per block for
 * AT; for A 416, 418, 420: 4 slices per lane with 16-, 18- or 20-pixel
 * strips, and 32: 1 slice per lane and two angles per wave, 32 slices per
 * block); see vp_default_variant and vp_launch_variant.  `grid` is the number of persistent workgroups
 * (<= 0: one per work item).  Results do not depend on the grid or on the
 * other instances in the batch.  Because the launch reads only device
 * memory, it can be captured into a HIP graph and replayed.
 */
typedef struct vp_batch_item {
    const float *in;      /* A: volume x (SY, SX, H); AT: sinogram y (A, PX, H) */
    float *out;           /* A: sinogram y;           AT: volume x            */
    const float *angles;  /* n_angles float32 radians, device memory          */
    int32_t n_angles, SY, SX, H, PX;
    int32_t accumulate;   /* 1: out += op(in); 0: out = op(in)                */
} vp_batch_item;

VP_API int vp_forward_batched(int variant, const vp_batch_item *items_host,
                              const vp_batch_item *items_dev, int n, int grid, void *stream);
VP_API int vp_backward_batched(int variant, const vp_batch_item *items_host,
                               const vp_batch_item *items_dev, int n, int grid, void *stream);

/*
 * Batched layout conversion, for batched launches of z_order 2 instances:
 * out (+)= transpose(in) for n row-major matrices in one persistent launch
 * (in: rows x cols; out: cols x rows; accumulate 1 adds to out, 0
 * overwrites it).  A z_order 2 volume (H, SY, SX) is the H x SY*SX matrix
 * whose transpose is the z_order 0 volume (SY, SX, H); a z_order 2 sinogram
 * (H, A, PX) likewise transposes to (A, PX, H).  Descriptors are passed as
 * for vp_forward_batched (host copy for validation, identical device copy
 * read by the kernel); `grid` persistent workgroups, <= 0 for one per 64 x 64
 * tile (at most 8192).  No counterpart in the reference, whose z_order 2
 * kernels read their layout directly (voxelproj/csrc/include/sp_forward_z2.cuh,
 * sp_backward_z2.cuh).
 */
typedef struct vp_transpose_item {
    const float *in;
    float *out;
    int64_t rows, cols;
    int32_t accumulate;
    int32_t reserved;     /* 0 */
} vp_transpose_item;

VP_API int vp_transpose_batched(const vp_transpose_item *items_host,
                                const vp_transpose_item *items_dev, int n, int grid, void *stream);

/* Default kernel variant of op 0 (A) / 1 (AT) for H slices, the grouping key
 * of batched launches.  Variants 1, 2, 4: 1, 2 or 4 consecutive slices per
 * lane (blocks of 64, 128, 256 slices); AT also has 32, 64, 128: blocks of
 * that many slices at 4 slices per lane, the wave's lanes in rows of 8, 16,
 * 32 that hold one voxel each (its default below 129 slices).  vp_forward /
 * vp_backward start from the default (see vp_launch_variant).  Variants 1
 * and 2 give identical results; AT's 4, 32, 64 and 128 identical results
 * among themselves; the two families add the same terms in another order
 * (agreement to float32 rounding).  A also has 416, 418 and 420 (4 slices
 * per lane with 16-, 18- and 20-pixel strips instead of 22) and 32 (1 slice
 * per lane and two angles per wave, never a default: for the autotuner on
 * batches of H <= 32; equal to variant 1 on volumes of at most 128 voxels per
 * side).  A's 4-slice variants agree with each other, and 32 with 1, to
 * float32 rounding. */
VP_API int vp_default_variant(int op, int64_t H);

/* The variant a single vp_forward / vp_backward call of this geometry uses,
 * so that a batched launch grouped by it gives the same results bit for bit:
 * the default of vp_default_variant, except that for A 2 slices per lane
 * drops to 1 when its grid has fewer workgroups than the device has compute
 * units, and 4 slices per lane takes the strip width (4, 416, 418, 420) whose
 * grid fills the device best.  Uses the current device's compute-unit count
 * (256 without a device).  -1 on invalid input.  No counterpart in the
 * reference. */
VP_API int vp_launch_variant(int op, int n_angles, int64_t SY, int64_t SX, int64_t H, int64_t PX);

/* Number of work items (workgroups) one instance contributes to a batched
 * launch of `op` (0 = A, 1 = AT) under `variant`; < 0 on invalid input. */
VP_API int64_t vp_batch_blocks(int op, int variant, int n_angles, int64_t SY, int64_t SX,
                               int64_t H, int64_t PX);

/* Description of the last failure on the calling thread ("" if none). */
VP_API const char *vp_last_error(void);

/* ABI version of this header (incremented whenever the exported interface changes). */
VP_API int vp_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* VOXELPROJ_HIP_H */
