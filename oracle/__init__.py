"""CPU oracle for the exact-trapezoid projector pair (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this package, and only as the checker.  The product package
``voxelproj_amd`` never imports it and has no CPU fallback.

Functions
---------
forward_f64 / backward_f64
    Exact operator in FP64 (vp_oracle.c: vpo_forward_f64 / vpo_backward_f64),
    restating the FP64 footprint of the reference's ``dpiece_wise_integrated``
    (voxelproj/csrc/include/projection_common.cuh:61-85) with the 3-tap
    support and border-zero rule of voxelproj/csrc/include/sp_backward_z0.cuh:84-98.
ref_forward_f32 / ref_backward_f32
    FP32 transcription of the reference CUDA kernels' arithmetic
    (voxelproj/csrc/include/sp_forward_z0.cuh:8-185, sp_backward_z0.cuh:8-100
    and their z_order 2 twins), standing in for the CUDA operator.
forward_geom_f64 / backward_geom_f64
    The exact operator with detector translations du (pixels, along the
    detector) and dv (slices, along H, linear interpolation), written from
    the definition (an extension: the reference lists translations as future
    work, README.md:65-68).
nnz
    Exact nonzero count of W, for the algorithmic flop count 2*H*nnz.

Layouts: z_order 0 -> x (SY, SX, H), y (A, PX, H); z_order 2 -> x (H, SY, SX),
y (H, A, PX).
"""
from .oracle import (  # noqa: F401
    load,
    build,
    forward_f64,
    backward_f64,
    forward_geom_f64,
    backward_geom_f64,
    ref_forward_f32,
    ref_backward_f32,
    nnz,
    weight_f64,
    cdf_f64,
    angle_table_f64,
    max_threads,
    x_shape,
    y_shape,
)
