// Per-axis clamp-to-edge pass (forward) and its transpose (adjoint).
//
// Forward: the convolve_axis of src/filters/conv.rs:11-41 for any ndim <= 4,
// f32/f64, 1..4 interleaved components, on strided views.  One thread per
// output pixel; it is the general path (any D, any radius up to
// ITKG_MAX_RADIUS) and the arithmetic reference the fused kernels match.
//
// EXACT arithmetic reproduces conv.rs:27-35 with pixel.rs:21: the
// accumulator starts at +0 and each tap adds fl(x * (T)w_k), taps in order,
// no FMA (this file is compiled with -ffp-contract=off and the multiply and
// add are written as separate statements).
#include "itkg_impl.hpp"

namespace itkg {

namespace {

template <typename T, int NC, bool FAST>
__global__ __launch_bounds__(256) void axis_forward_kernel(const AxisPassArgs<T> a) {
    const int ax = a.axis;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t lin = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; lin < a.total;
         lin += step) {
        int64_t rem = lin;
        int64_t c[ITKG_MAX_DIM];
#pragma unroll
        for (int d = 0; d < ITKG_MAX_DIM - 1; ++d) {
            c[d] = rem % a.n[d];
            rem /= a.n[d];
        }
        c[ITKG_MAX_DIM - 1] = rem;
        int64_t soff = 0, doff = 0, oax = 0;
#pragma unroll
        for (int d = 0; d < ITKG_MAX_DIM; ++d) {
            const int64_t o = a.out_lo[d] + c[d];
            doff += (o - a.dst_lo[d]) * a.dst_stride[d];
            if (d == ax) {
                oax = o;
            } else {
                soff += (o - a.src_lo[d]) * a.src_stride[d];
            }
        }
        T acc[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[q] = T(0);
        const int taps = 2 * a.radius + 1;
        for (int k = 0; k < taps; ++k) {
            int64_t v = oax + k - a.radius;
            v = v < a.clamp_lo ? a.clamp_lo : v;
            v = v > a.clamp_hi ? a.clamp_hi : v;
            const T* p = a.src + (soff + (v - a.src_lo[ax]) * a.src_stride[ax]) * NC;
            const T w = a.w[k];
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                if constexpr (FAST) {
                    acc[q] = __builtin_fma(p[q], w, acc[q]);
                } else {
                    const T prod = p[q] * w;
                    acc[q] = acc[q] + prod;
                }
            }
        }
        T* out = a.dst + doff * NC;
#pragma unroll
        for (int q = 0; q < NC; ++q) out[q] = acc[q];
    }
}

// Transposed pass: the destination is the forward pass's source region.
// g_in[j] = sum_k g_out[j - k + r] * w_k over outputs that exist, plus, at
// the two clamp edges, every output's clamped taps folded onto the edge:
// at j = lo the outputs i < lo + r contribute g_out[i] * sum_{k < lo-i+r} w_k,
// at j = hi the outputs i > hi - r contribute g_out[i] * sum_{k < i+r-hi} w_k
// (the weights are symmetric, so the top-m sum equals the first-m sum).
template <typename T, int NC>
__global__ __launch_bounds__(256) void axis_adjoint_kernel(const AxisPassArgs<T> a,
                                                           int64_t g_lo, int64_t g_hi) {
    const int ax = a.axis;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t lin = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; lin < a.total;
         lin += step) {
        int64_t rem = lin;
        int64_t c[ITKG_MAX_DIM];
#pragma unroll
        for (int d = 0; d < ITKG_MAX_DIM - 1; ++d) {
            c[d] = rem % a.n[d];
            rem /= a.n[d];
        }
        c[ITKG_MAX_DIM - 1] = rem;
        int64_t soff = 0, doff = 0, j = 0;
#pragma unroll
        for (int d = 0; d < ITKG_MAX_DIM; ++d) {
            const int64_t o = a.out_lo[d] + c[d];
            doff += (o - a.dst_lo[d]) * a.dst_stride[d];
            if (d == ax) {
                j = o;
            } else {
                soff += (o - a.src_lo[d]) * a.src_stride[d];
            }
        }
        const int64_t gs = a.src_stride[ax];
        T acc[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[q] = T(0);
        const int r = a.radius;
        for (int k = 0; k <= 2 * r; ++k) {
            const int64_t i = j - k + r;
            if (i < g_lo || i > g_hi) continue;
            const T* p = a.src + (soff + (i - a.src_lo[ax]) * gs) * NC;
#pragma unroll
            for (int q = 0; q < NC; ++q) acc[q] = acc[q] + p[q] * a.w[k];
        }
        if (j == a.clamp_lo) {
            const int64_t ie = a.clamp_lo + r - 1 < g_hi ? a.clamp_lo + r - 1 : g_hi;
            for (int64_t i = g_lo; i <= ie; ++i) {
                const T t = a.tail[a.clamp_lo - i + r];
                const T* p = a.src + (soff + (i - a.src_lo[ax]) * gs) * NC;
#pragma unroll
                for (int q = 0; q < NC; ++q) acc[q] = acc[q] + p[q] * t;
            }
        }
        if (j == a.clamp_hi) {
            const int64_t ib = a.clamp_hi - r + 1 > g_lo ? a.clamp_hi - r + 1 : g_lo;
            for (int64_t i = ib; i <= g_hi; ++i) {
                const T t = a.tail[i + r - a.clamp_hi];
                const T* p = a.src + (soff + (i - a.src_lo[ax]) * gs) * NC;
#pragma unroll
                for (int q = 0; q < NC; ++q) acc[q] = acc[q] + p[q] * t;
            }
        }
        T* out = a.dst + doff * NC;
#pragma unroll
        for (int q = 0; q < NC; ++q) out[q] = a.accumulate ? out[q] + acc[q] : acc[q];
    }
}

inline unsigned grid_for(int64_t total) {
    int64_t b = (total + 255) / 256;
    const int64_t cap = int64_t(1) << 20;
    return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

template <typename T, bool FAST>
hipError_t fwd_nc(const AxisPassArgs<T>& a, int ncomp, hipStream_t s) {
    const dim3 grid(grid_for(a.total)), block(256);
    switch (ncomp) {
        case 1: hipLaunchKernelGGL((axis_forward_kernel<T, 1, FAST>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((axis_forward_kernel<T, 2, FAST>), grid, block, 0, s, a); break;
        case 3: hipLaunchKernelGGL((axis_forward_kernel<T, 3, FAST>), grid, block, 0, s, a); break;
        case 4: hipLaunchKernelGGL((axis_forward_kernel<T, 4, FAST>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_axis_forward(const AxisPassArgs<T>& a, int ncomp, itkg_arith arith,
                               hipStream_t s) {
    if (a.total == 0) return hipSuccess;
    return arith == ITKG_FAST ? fwd_nc<T, true>(a, ncomp, s) : fwd_nc<T, false>(a, ncomp, s);
}

template <typename T>
hipError_t launch_axis_adjoint_range(const AxisPassArgs<T>& a, int ncomp, int64_t g_lo,
                                     int64_t g_hi, hipStream_t s) {
    if (a.total == 0) return hipSuccess;
    const dim3 grid(grid_for(a.total)), block(256);
    switch (ncomp) {
        case 1: hipLaunchKernelGGL((axis_adjoint_kernel<T, 1>), grid, block, 0, s, a, g_lo, g_hi); break;
        case 2: hipLaunchKernelGGL((axis_adjoint_kernel<T, 2>), grid, block, 0, s, a, g_lo, g_hi); break;
        case 3: hipLaunchKernelGGL((axis_adjoint_kernel<T, 3>), grid, block, 0, s, a, g_lo, g_hi); break;
        case 4: hipLaunchKernelGGL((axis_adjoint_kernel<T, 4>), grid, block, 0, s, a, g_lo, g_hi); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_axis_forward<float>(const AxisPassArgs<float>&, int, itkg_arith,
                                               hipStream_t);
template hipError_t launch_axis_forward<double>(const AxisPassArgs<double>&, int, itkg_arith,
                                                hipStream_t);
template hipError_t launch_axis_adjoint_range<float>(const AxisPassArgs<float>&, int, int64_t,
                                                     int64_t, hipStream_t);
template hipError_t launch_axis_adjoint_range<double>(const AxisPassArgs<double>&, int, int64_t,
                                                      int64_t, hipStream_t);

}  // namespace itkg
