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
};

// x (+)= AT y in z_order 0 (accumulate != 0 adds, else overwrites).
hipError_t launch_backward_z0(const float* y, float* x, const float* angles, const Geom& g,
                              int accumulate, hipStream_t s);

// y (+)= A x in z_order 0 (accumulate != 0 adds, else overwrites).
hipError_t launch_forward_z0(const float* x, float* y, const float* angles, const Geom& g,
                             int accumulate, hipStream_t s);

// FP64 operator pair, either layout (outputs accumulate).
hipError_t launch_forward_f64(int z_order, const double* x, double* y, const float* angles,
                              const Geom& g, hipStream_t s);
hipError_t launch_backward_f64(int z_order, const double* y, double* x, const float* angles,
                               const Geom& g, hipStream_t s);

// out[c][r] (+)= in[r][c] for an R x C row-major matrix.
hipError_t launch_transpose(const float* in, float* out, int64_t R, int64_t C, int accumulate,
                            hipStream_t s);

}  // namespace vp
