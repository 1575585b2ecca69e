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
// enumerate axes 1..3 (axis 1 fastest), elements run over x * ncomp.  The
// row and chunk of an item are uniform across the block; per element only
// the component split remains, none for scalar pixels and in 32 bits when
// a row has fewer than 2^32 elements.
template <class S, class D, bool ROUND>
__global__ __launch_bounds__(kConvBlock) void convert_kernel(const ConvertArgs a) {
    const int64_t row_elems = a.n[0] * a.ncomp;
    const int64_t chunks = (row_elems + kConvBlock - 1) / kConvBlock;
    const int64_t rows = a.n[1] * a.n[2] * a.n[3];
    const bool narrow = row_elems <= (int64_t)UINT32_MAX;
    const uint32_t nc = (uint32_t)a.ncomp;
    for (int64_t item = blockIdx.x; item < rows * chunks; item += gridDim.x) {
        const int64_t row = item / chunks, c = item - row * chunks;
        const int64_t y = row % a.n[1], zw = row / a.n[1];
        const int64_t z = zw % a.n[2], w = zw / a.n[2];
        const int64_t j = c * kConvBlock + threadIdx.x;
        if (j >= row_elems) continue;
        int64_t x, e;
        if (nc == 1) {
            x = j;
            e = 0;
        } else if (narrow) {
            const uint32_t q = (uint32_t)j / nc;
            x = q;
            e = (int64_t)((uint32_t)j - q * nc);
        } else {
            x = j / nc;
            e = j - x * nc;
        }
        const int64_t si = (x * a.in_stride[0] + y * a.in_stride[1] + z * a.in_stride[2] +
                            w * a.in_stride[3]) * nc + e;
        const int64_t di = (x * a.out_stride[0] + y * a.out_stride[1] + z * a.out_stride[2] +
                            w * a.out_stride[3]) * nc + e;
        S v = static_cast<const S*>(a.in)[si];
        if constexpr (ROUND && std::is_floating_point_v<S>) v = round_half_away(v);
        static_cast<D*>(a.out)[di] = rust_as<D>(v);
    }
}

// The common case: blockIdx.y is the row's axis-1 index, blockIdx.z its
// axis-2/3 plane, blockIdx.x a span of U * kConvBlock elements of the row
// (U = 1..4 chosen so that short rows do not leave threads idle), so no
// index needs a division beyond the component split.  Needs n[1] and
// n[2] * n[3] within the grid limits and rows under 2^31 elements.
constexpr int kConvMaxU = 4;

template <class S, class D, bool ROUND>
__global__ __launch_bounds__(kConvBlock) void convert_rows_kernel(const ConvertArgs a,
                                                                  const int u) {
    const uint32_t nc = (uint32_t)a.ncomp;
    const uint32_t n2 = (uint32_t)a.n[2];
    const int64_t y = blockIdx.y, z = blockIdx.z % n2, w = blockIdx.z / n2;
    const int64_t in_row = y * a.in_stride[1] + z * a.in_stride[2] + w * a.in_stride[3];
    const int64_t out_row = y * a.out_stride[1] + z * a.out_stride[2] + w * a.out_stride[3];
    const uint32_t row_elems = (uint32_t)(a.n[0] * nc);
    const uint32_t j0 = blockIdx.x * (uint32_t)(u * kConvBlock) + threadIdx.x;
    const S* in = static_cast<const S*>(a.in);
    D* out = static_cast<D*>(a.out);
#pragma unroll
    for (int k = 0; k < kConvMaxU; ++k) {
        const uint32_t j = j0 + (uint32_t)(k * kConvBlock);
        if (k >= u || j >= row_elems) break;
        const uint32_t x = nc == 1 ? j : j / nc;
        const uint32_t e = j - x * nc;
        S v = in[((int64_t)x * a.in_stride[0] + in_row) * nc + e];
        if constexpr (ROUND && std::is_floating_point_v<S>) v = round_half_away(v);
        out[((int64_t)x * a.out_stride[0] + out_row) * nc + e] = rust_as<D>(v);
    }
}

// Four consecutive elements per lane and iteration, as one load and one
// store each: dense scalar rows whose starts are all aligned to four
// elements (the host checks the pointers and strides).
template <class T>
struct alignas(4 * sizeof(T)) Quad {
    T v[4];
};

template <class S, class D, bool ROUND>
__global__ __launch_bounds__(kConvBlock) void convert_rows4_kernel(const ConvertArgs a,
                                                                   const int u) {
    const uint32_t n2 = (uint32_t)a.n[2];
    const int64_t y = blockIdx.y, z = blockIdx.z % n2, w = blockIdx.z / n2;
    const int64_t in_row = y * a.in_stride[1] + z * a.in_stride[2] + w * a.in_stride[3];
    const int64_t out_row = y * a.out_stride[1] + z * a.out_stride[2] + w * a.out_stride[3];
    const uint32_t quads = (uint32_t)(a.n[0] / 4);
    const uint32_t q0 = blockIdx.x * (uint32_t)(u * kConvBlock) + threadIdx.x;
    const Quad<S>* in = reinterpret_cast<const Quad<S>*>(static_cast<const S*>(a.in) + in_row);
    Quad<D>* out = reinterpret_cast<Quad<D>*>(static_cast<D*>(a.out) + out_row);
#pragma unroll
    for (int k = 0; k < kConvMaxU; ++k) {
        const uint32_t q = q0 + (uint32_t)(k * kConvBlock);
        if (k >= u || q >= quads) break;
        const Quad<S> v = in[q];
        Quad<D> r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            S t = v.v[i];
            if constexpr (ROUND && std::is_floating_point_v<S>) t = round_half_away(t);
            r.v[i] = rust_as<D>(t);
        }
        out[q] = r;
    }
}

// Views that are one dense run of elements on both sides (whole contiguous
// images, the usual CastFilter call): a flat range of quads, kConvFlatU per
// lane, kConvFlatU * kConvBlock per block, so a wave moves several rows'
// worth of bytes where a row-per-block grid (a 512-pixel row is 128 quads)
// would leave half of each block idle and spend its time in workgroup
// dispatch.  The last n % 4 elements, if any, go one per lane in block 0.
constexpr int kConvFlatU = 2;

template <class S, class D, bool ROUND>
__global__ __launch_bounds__(kConvBlock) void convert_flat4_kernel(const S* in, D* out,
                                                                   const uint32_t quads,
                                                                   const uint32_t tail) {
    const uint32_t q0 = blockIdx.x * (uint32_t)(kConvFlatU * kConvBlock) + threadIdx.x;
    const Quad<S>* in4 = reinterpret_cast<const Quad<S>*>(in);
    Quad<D>* out4 = reinterpret_cast<Quad<D>*>(out);
    Quad<S> v[kConvFlatU];
#pragma unroll
    for (int k = 0; k < kConvFlatU; ++k) {
        const uint32_t q = q0 + (uint32_t)(k * kConvBlock);
        if (q < quads) v[k] = in4[q];
    }
#pragma unroll
    for (int k = 0; k < kConvFlatU; ++k) {
        const uint32_t q = q0 + (uint32_t)(k * kConvBlock);
        if (q < quads) {
            Quad<D> r;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                S t = v[k].v[i];
                if constexpr (ROUND && std::is_floating_point_v<S>) t = round_half_away(t);
                r.v[i] = rust_as<D>(t);
            }
            out4[q] = r;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < tail) {
        const int64_t j = (int64_t)quads * 4 + threadIdx.x;
        S t = in[j];
        if constexpr (ROUND && std::is_floating_point_v<S>) t = round_half_away(t);
        out[j] = rust_as<D>(t);
    }
}

// True when the view's elements (pixels times components) are one dense
// run in the order x, then axes 1..3: every axis of more than one pixel
// sits at the stride of the axes below it.
inline bool dense_run(const int64_t* n, const int64_t* stride) {
    int64_t p = 1;
    for (int d = 0; d < ITKG_MAX_DIM; ++d) {
        if (n[d] > 1 && stride[d] != p) return false;
        p *= n[d];
    }
    return true;
}

template <class S, class D>
bool quad_ok(const ConvertArgs& a) {
    if (a.ncomp != 1 || a.in_stride[0] != 1 || a.out_stride[0] != 1 || a.n[0] % 4) return false;
    if ((uintptr_t)a.in % (4 * sizeof(S)) || (uintptr_t)a.out % (4 * sizeof(D))) return false;
    for (int d = 1; d < ITKG_MAX_DIM; ++d)
        if ((a.n[d] > 1) && (a.in_stride[d] % 4 || a.out_stride[d] % 4)) return false;
    return true;
}

template <class S, class D>
hipError_t launch_sd(const ConvertArgs& a, bool round, hipStream_t s) {
    const int64_t row_elems = a.n[0] * a.ncomp;
    const int64_t elems = row_elems * a.n[1] * a.n[2] * a.n[3];
    if (dense_run(a.n, a.in_stride) && dense_run(a.n, a.out_stride) &&
        (uintptr_t)a.in % (4 * sizeof(S)) == 0 && (uintptr_t)a.out % (4 * sizeof(D)) == 0 &&
        elems / 4 < (INT64_C(1) << 32) - kConvFlatU * kConvBlock) {
        const int64_t quads = elems / 4;
        const int64_t per_block = (int64_t)kConvFlatU * kConvBlock;
        const unsigned g = (unsigned)std::max<int64_t>(1, (quads + per_block - 1) / per_block);
        const S* in = static_cast<const S*>(a.in);
        D* out = static_cast<D*>(a.out);
        if (round)
            hipLaunchKernelGGL((convert_flat4_kernel<S, D, true>), dim3(g), dim3(kConvBlock), 0, s,
                               in, out, (uint32_t)quads, (uint32_t)(elems % 4));
        else
            hipLaunchKernelGGL((convert_flat4_kernel<S, D, false>), dim3(g), dim3(kConvBlock), 0, s,
                               in, out, (uint32_t)quads, (uint32_t)(elems % 4));
        return hipGetLastError();
    }
    if (row_elems < (INT64_C(1) << 31) && a.n[1] <= 65535 && a.n[2] * a.n[3] <= 65535 &&
        quad_ok<S, D>(a)) {
        const int64_t quads = a.n[0] / 4;
        const int u = (int)std::min<int64_t>(kConvMaxU, (quads + kConvBlock - 1) / kConvBlock);
        const dim3 g((unsigned)((quads + (int64_t)u * kConvBlock - 1) / ((int64_t)u * kConvBlock)),
                     (unsigned)a.n[1], (unsigned)(a.n[2] * a.n[3]));
        if (round)
            hipLaunchKernelGGL((convert_rows4_kernel<S, D, true>), g, dim3(kConvBlock), 0, s, a, u);
        else
            hipLaunchKernelGGL((convert_rows4_kernel<S, D, false>), g, dim3(kConvBlock), 0, s, a, u);
        return hipGetLastError();
    }
    if (row_elems < (INT64_C(1) << 31) && a.n[1] <= 65535 && a.n[2] * a.n[3] <= 65535) {
        const int u = (int)std::min<int64_t>(kConvMaxU, (row_elems + kConvBlock - 1) / kConvBlock);
        const dim3 g((unsigned)((row_elems + (int64_t)u * kConvBlock - 1) / ((int64_t)u * kConvBlock)),
                     (unsigned)a.n[1], (unsigned)(a.n[2] * a.n[3]));
        if (round)
            hipLaunchKernelGGL((convert_rows_kernel<S, D, true>), g, dim3(kConvBlock), 0, s, a, u);
        else
            hipLaunchKernelGGL((convert_rows_kernel<S, D, false>), g, dim3(kConvBlock), 0, s, a, u);
        return hipGetLastError();
    }
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
