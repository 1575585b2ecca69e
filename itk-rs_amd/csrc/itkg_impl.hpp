// Internal declarations shared by the C-ABI host code and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "itkg.h"

namespace itkg {

constexpr int kMaxTaps = 2 * ITKG_MAX_RADIUS + 1;

// Thread-local error slot behind itkg_last_error().
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
void clear_error();

// Arguments of one clamp-to-edge pass along `axis` (the convolve_axis of
// src/filters/conv.rs:11-41) from a strided source view to a strided
// destination view.  Coordinates are absolute image indices.
template <typename T>
struct AxisPassArgs {
    const T* src;          // pixel at src_lo
    T* dst;                // pixel at dst_lo
    int64_t n[ITKG_MAX_DIM];         // pass output extents (1 beyond ndim)
    int64_t out_lo[ITKG_MAX_DIM];    // pass output region index
    int64_t src_lo[ITKG_MAX_DIM];
    int64_t src_stride[ITKG_MAX_DIM];
    int64_t dst_lo[ITKG_MAX_DIM];
    int64_t dst_stride[ITKG_MAX_DIM];
    int64_t clamp_lo, clamp_hi;      // pass input extent on `axis`
    int64_t total;                   // product of n
    int32_t axis, radius;
    int32_t accumulate;              // adjoint only: dst += result
    T w[kMaxTaps];                   // (T)w_k, k = 0..2r
    T tail[ITKG_MAX_RADIUS + 1];     // adjoint only: (T)(sum_{k<m} w_k), m = 0..r
};

template <typename T>
hipError_t launch_axis_forward(const AxisPassArgs<T>& a, int ncomp, itkg_arith arith,
                               hipStream_t s);
// Adjoint pass; [g_lo, g_hi] is the gradient-of-output extent on the axis.
template <typename T>
hipError_t launch_axis_adjoint_range(const AxisPassArgs<T>& a, int ncomp, int64_t g_lo,
                                     int64_t g_hi, hipStream_t s);

// Fused one-round-trip 3-D forward for f32 scalar images.
struct Fused3dArgs {
    const float* src;        // input view pixel at the pass-0 input region index
    float* dst;              // output view pixel at the requested region index
    int64_t src_stride[3];
    int64_t dst_stride[3];
    int32_t n_in[3];         // input region extents (the clamp bounds)
    int32_t n_out[3];        // requested region extents
    int32_t off[3];          // requested.index - input.index (0..r)
    int32_t tz;              // output planes per block
    int32_t tiles_x, tiles_y, tiles_z;
    float wx[2 * 32 + 1], wy[2 * 32 + 1], wz[2 * 32 + 1];
};

// True when the fused kernel has an instantiation for these radii.
bool fused3d_supported(int rx, int ry, int rz);
// Output tile of one fused workgroup in x and y.
int fused3d_tile_x();
int fused3d_tile_y(int rz);
hipError_t launch_fused3d(const Fused3dArgs& a, int rx, int ry, int rz, itkg_arith arith,
                          hipStream_t s);

}  // namespace itkg
