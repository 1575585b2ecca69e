// Pixel conversions of the UnaryFilter nodes that surround a Gaussian in
// itk-rs pipelines (src/filters/mod.rs:47-80; examples/blur_png.rs:35-37
// casts u8 -> f32 with `p as f32` and back with
// `p.round().clamp(0.0, 255.0) as u8`), with Rust's `as` semantics:
//   - integer -> integer: a narrower target keeps the low bits, a wider one
//     extends by the source's signedness;
//   - integer -> float, f64 -> f32: round to nearest, ties to even;
//   - float -> integer: truncation toward zero, saturated to the target
//     range, NaN -> 0;
//   - ROUND mode first rounds a float half away from zero (f32::round,
//     f64::round); for the u8 target the reference's clamp(0, 255) then
//     changes nothing that the saturating cast does not already do (NaN
//     stays NaN through clamp and casts to 0).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <type_traits>

#include "itkg_impl.hpp"

namespace itkg {

namespace {

template <class F>
__device__ __forceinline__ F round_half_away(F x) {
    // x - trunc(x) is exact, so the test is exact too
    const F t = x < F(0) ? ceil(x) : floor(x);
    const F d = x - t;
    return d >= F(0.5) ? t + F(1) : (d <= F(-0.5) ? t - F(1) : t);
}

template <class D, class S>
__device__ __forceinline__ D rust_as(S v) {
    if constexpr (std::is_floating_point_v<S> && std::is_integral_v<D>) {
        if (v != v) return D(0);
        constexpr S lo = (S)std::numeric_limits<D>::min();
        constexpr S hi = (S)std::numeric_limits<D>::max();
        if (v <= lo) return std::numeric_limits<D>::min();
        if (v >= hi) return std::numeric_limits<D>::max();
        return (D)v;  // in range: truncation toward zero
    } else if constexpr (std::is_integral_v<S> && std::is_integral_v<D>) {
        return (D)v;  // widening extends by the source's signedness, narrowing truncates
    } else {
        return (D)v;
    }
}

constexpr int kConvBlock = 256;

// One block per (row, chunk of kConvBlock elements of the row); rows
// enumerate axes 1..3 (axis 1 fastest), elements run over x * ncomp.
template <class S, class D, bool ROUND>
__global__ __launch_bounds__(kConvBlock) void convert_kernel(const ConvertArgs a) {
    const int64_t row_elems = a.n[0] * a.ncomp;
    const int64_t chunks = (row_elems + kConvBlock - 1) / kConvBlock;
    const int64_t rows = a.n[1] * a.n[2] * a.n[3];
    for (int64_t item = blockIdx.x; item < rows * chunks; item += gridDim.x) {
        const int64_t row = item / chunks, c = item - row * chunks;
        const int64_t y = row % a.n[1], zw = row / a.n[1];
        const int64_t z = zw % a.n[2], w = zw / a.n[2];
        const int64_t j = c * kConvBlock + threadIdx.x;
        if (j >= row_elems) continue;
        const int64_t x = j / a.ncomp, e = j - x * a.ncomp;
        const int64_t si = (x * a.in_stride[0] + y * a.in_stride[1] + z * a.in_stride[2] +
                            w * a.in_stride[3]) * a.ncomp + e;
        const int64_t di = (x * a.out_stride[0] + y * a.out_stride[1] + z * a.out_stride[2] +
                            w * a.out_stride[3]) * a.ncomp + e;
        S v = static_cast<const S*>(a.in)[si];
        if constexpr (ROUND && std::is_floating_point_v<S>) v = round_half_away(v);
        static_cast<D*>(a.out)[di] = rust_as<D>(v);
    }
}

template <class S, class D>
hipError_t launch_sd(const ConvertArgs& a, bool round, hipStream_t s) {
    const int64_t row_elems = a.n[0] * a.ncomp;
    const int64_t items = ((row_elems + kConvBlock - 1) / kConvBlock) * a.n[1] * a.n[2] * a.n[3];
    const unsigned grid = (unsigned)std::min<int64_t>(items, 1 << 20);
    if (round)
        hipLaunchKernelGGL((convert_kernel<S, D, true>), dim3(grid), dim3(kConvBlock), 0, s, a);
    else
        hipLaunchKernelGGL((convert_kernel<S, D, false>), dim3(grid), dim3(kConvBlock), 0, s, a);
    return hipGetLastError();
}

template <class S>
hipError_t launch_s(const ConvertArgs& a, bool round, hipStream_t s) {
    switch (a.out_type) {
        case ITKG_PT_U8: return launch_sd<S, uint8_t>(a, round, s);
        case ITKG_PT_U16: return launch_sd<S, uint16_t>(a, round, s);
        case ITKG_PT_I16: return launch_sd<S, int16_t>(a, round, s);
        case ITKG_PT_I32: return launch_sd<S, int32_t>(a, round, s);
        case ITKG_PT_F32: return launch_sd<S, float>(a, round, s);
        case ITKG_PT_F64: return launch_sd<S, double>(a, round, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_convert(const ConvertArgs& a, hipStream_t s) {
    if (a.n[0] * a.n[1] * a.n[2] * a.n[3] == 0) return hipSuccess;
    const bool round = a.mode == ITKG_CONVERT_ROUND;
    switch (a.in_type) {
        case ITKG_PT_U8: return launch_s<uint8_t>(a, round, s);
        case ITKG_PT_U16: return launch_s<uint16_t>(a, round, s);
        case ITKG_PT_I16: return launch_s<int16_t>(a, round, s);
        case ITKG_PT_I32: return launch_s<int32_t>(a, round, s);
        case ITKG_PT_F32: return launch_s<float>(a, round, s);
        case ITKG_PT_F64: return launch_s<double>(a, round, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace itkg
