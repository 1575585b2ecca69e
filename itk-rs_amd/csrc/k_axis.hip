// Per-axis clamp-to-edge pass (forward) and its transpose (adjoint).
//
// Forward: the convolve_axis of src/filters/conv.rs:11-41 for any ndim <= 4,
// f32/f64, 1..4 interleaved components, on strided views.  It is the
// general path (any D, any radius up to ITKG_MAX_RADIUS, anisotropic radii)
// and the arithmetic reference the fused kernels match.
//
// EXACT arithmetic reproduces conv.rs:27-35 with pixel.rs:21: the
// accumulator starts at +0 and each tap adds fl(x * (T)w_k), taps in order,
// no FMA (this file is compiled with -ffp-contract=off and the multiply and
// add are written as separate statements).  FAST: one FMA per tap.
//
// Work decomposition (one wave = one work item, no workgroup barriers):
//   - pass axis != 0 ("strided"): the 64 lanes take 64 consecutive pixels
//     along axis 0 and each walks a run of CPT outputs along the pass axis,
//     reading the CPT + 2r inputs of the run once (coalesced across lanes,
//     clamped to the pass input extent);
//   - pass axis 0 ("row"): the wave stages the 64 * CPT + 2r inputs of a row
//     segment in LDS (coalesced), then each lane computes a run of CPT
//     consecutive outputs, reading LDS through a padded layout (one pad
//     pixel per CPT) that keeps the lanes' strided reads conflict-free.
// Every output still receives its taps k = 0..2r in increasing order.
#include "itkg_impl.hpp"

#include <type_traits>
#include <utility>

namespace itkg {

namespace {

template <class F, int... I>
__device__ __forceinline__ void static_for_impl(std::integer_sequence<int, I...>, F&& f) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(std::make_integer_sequence<int, N>{}, f);
}

// One tap of one output component.
template <typename T, bool FAST>
__device__ __forceinline__ T tap(T acc, T x, T w) {
    if constexpr (FAST) {
        if constexpr (std::is_same_v<T, float>)
            return __builtin_fmaf(x, w, acc);  // single rounding in f32
        else
            return __builtin_fma(x, w, acc);
    } else {
        const T prod = x * w;
        return acc + prod;
    }
}

// A run of CPT consecutive outputs: output c takes input c + j as its tap j
// (j = 0..2r, increasing, so each output's sum has the reference order).
// The loop runs over taps; a ring of CPT registers holds inputs j .. j+CPT-1
// (input i in slot i mod CPT), and a block of CPT taps is unrolled so every
// slot index is a constant: per tap one weight (uniform), one input load
// (the input CPT ahead, into the slot just consumed) and CPT multiply-adds.
// load(jb, off, v) fetches input jb + off, jb a multiple of CPT (runtime),
// off < 2 CPT (compile-time); inputs 0 .. 2r + CPT are fetched.
template <typename T, int NC, bool FAST, int CPT, class Load>
__device__ __forceinline__ void conv_run(const T* w, const int r, Load&& load, T (&acc)[CPT][NC]) {
    const int taps = 2 * r + 1;
    T ring[CPT][NC];
    static_for<CPT>([&](auto c_) {
        constexpr int c = decltype(c_)::value;
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[c][q] = T(0);
        load(0, std::integral_constant<int, c>{}, ring[c]);
    });
    auto block = [&](const int jb, const int limit) {
        static_for<CPT>([&](auto jj_) {
            constexpr int jj = decltype(jj_)::value;
            if (jj < limit) {
                const T wj = w[jb + jj];
#pragma unroll
                for (int c = 0; c < CPT; ++c)
#pragma unroll
                    for (int q = 0; q < NC; ++q)
                        acc[c][q] = tap<T, FAST>(acc[c][q], ring[(jj + c) % CPT][q], wj);
                load(jb, std::integral_constant<int, jj + CPT>{}, ring[jj]);
            }
        });
    };
    int jb = 0;
    for (; jb + CPT <= taps; jb += CPT) block(jb, CPT);
    if (jb < taps) block(jb, taps - jb);
}

// XCD-aware block order: hardware deals workgroups to the 8 XCDs round
// robin, so block b works on item block (b mod 8) * ceil(G / 8) + b / 8 and
// neighbouring runs (which share 2r inputs) read through one XCD's L2.
__device__ __forceinline__ int64_t xcd_block() {
    const int64_t g = gridDim.x, b = blockIdx.x;
    return (b % 8) * ((g + 7) / 8) + b / 8;
}

// Adjoint edge terms of one output at pass-axis position j, added to o:
// at j = clamp_lo the outputs i < clamp_lo + r fold their clamped taps
// onto it with weight tail[clamp_lo - i + r], at j = clamp_hi the outputs
// i > clamp_hi - r with weight tail[i + r - clamp_hi] (the weights are
// symmetric, so the top-m sum equals the first-m sum).  `at(i)` reads the
// incoming gradient at pass-axis position i.
template <typename T, int NC, class At>
__device__ __forceinline__ void adjoint_edges(const AxisPassArgs<T>& a, int64_t j, At&& at,
                                              T (&o)[NC]) {
    const int r = a.radius;
    if (j == a.clamp_lo) {
        const int64_t ie = a.clamp_lo + r - 1 < a.g_hi ? a.clamp_lo + r - 1 : a.g_hi;
#pragma unroll 4
        for (int64_t i = a.g_lo; i <= ie; ++i) {
            T v[NC];
            at(i, v);
            const T t = a.tail[a.clamp_lo - i + r];
#pragma unroll
            for (int q = 0; q < NC; ++q) o[q] = tap<T, true>(o[q], v[q], t);
        }
    }
    if (j == a.clamp_hi) {
        const int64_t ib = a.clamp_hi - r + 1 > a.g_lo ? a.clamp_hi - r + 1 : a.g_lo;
#pragma unroll 4
        for (int64_t i = ib; i <= a.g_hi; ++i) {
            T v[NC];
            at(i, v);
            const T t = a.tail[i + r - a.clamp_hi];
#pragma unroll
            for (int q = 0; q < NC; ++q) o[q] = tap<T, true>(o[q], v[q], t);
        }
    }
}

// Pass along a strided axis (a.axis >= 1).  Forward: inputs clamped to
// [clamp_lo, clamp_hi] on the pass axis.  Adjoint (ADJ): the incoming
// gradient is zero outside [g_lo, g_hi] (a zero-padded correlation, which
// equals the convolution for the symmetric weights), plus the edge terms.
template <typename T, int NC, bool FAST, bool ADJ, int CPT>
__global__ __launch_bounds__(256) void axis_strided_kernel([[maybe_unused]] const AxisPassArgs<T> args) {
    // Read in place from the kernel-argument segment: the weights are indexed
    // with runtime values (scalar loads, no private copy of the block).
    const AxisPassArgs<T>& a = *(const AxisPassArgs<T>*)__builtin_amdgcn_kernarg_segment_ptr();
    const int ax = a.axis;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t item = xcd_block() * 4 + wave;  // wave-uniform from here on
    // the two axes that are neither 0 nor the pass axis
    const int b1 = ax == 1 ? 2 : 1;
    const int b2 = ax == 3 ? 2 : 3;
    const int64_t n0c = (a.n[0] + 63) / 64;
    const int64_t nac = (a.n[ax] + CPT - 1) / CPT;
    int64_t it = item;
    const int64_t i0 = it % n0c;
    it /= n0c;
    const int64_t ia = it % nac;
    it /= nac;
    const int64_t o1 = it % a.n[b1];
    const int64_t o2 = it / a.n[b1];
    if (o2 >= a.n[b2]) return;  // whole wave: past the last item
    const int64_t col = i0 * 64 + lane;
    const bool col_ok = col < a.n[0];
    const int64_t colc = col_ok ? col : a.n[0] - 1;
    const int64_t p0 = a.out_lo[0] + colc, p1 = a.out_lo[b1] + o1, p2 = a.out_lo[b2] + o2;
    const int r = a.radius;
    const int64_t first = a.out_lo[ax] + ia * CPT - r;  // pass-axis position of input 0
    // input bounds relative to input 0 (the extent along one axis fits int):
    // forward clamps into [clamp_lo, clamp_hi], the adjoint reads zeros
    // outside [g_lo, g_hi]
    const int lo_k = (int)((ADJ ? a.g_lo : a.clamp_lo) - first);
    const int hi_k = (int)((ADJ ? a.g_hi : a.clamp_hi) - first);
    const int64_t sa = a.src_stride[ax] * NC;
    const T* src = a.src + ((p0 - a.src_lo[0]) * a.src_stride[0] + (p1 - a.src_lo[b1]) * a.src_stride[b1] +
                            (p2 - a.src_lo[b2]) * a.src_stride[b2] + (first - a.src_lo[ax]) * a.src_stride[ax]) * NC;
    auto load = [&](int jb, auto off, T (&v)[NC]) {
        const int i = jb + decltype(off)::value;
        if constexpr (ADJ) {
            if (i < lo_k || i > hi_k) {
#pragma unroll
                for (int q = 0; q < NC; ++q) v[q] = T(0);
                return;
            }
        }
        const int ic = i < lo_k ? lo_k : (i > hi_k ? hi_k : i);
        const T* p = src + (int64_t)ic * sa;
#pragma unroll
        for (int q = 0; q < NC; ++q) v[q] = p[q];
    };
    T acc[CPT][NC];
    if constexpr (ADJ) {
        // Runs whose inputs 0 .. 2r + CPT all lie inside the gradient's
        // extent (all but the runs at the two ends of the axis) read without
        // the per-tap zero test, whose branches would otherwise hold each
        // block's loads to the block that issues them.
        auto load_in = [&](int jb, auto off, T (&v)[NC]) {
            const T* p = src + (int64_t)(jb + decltype(off)::value) * sa;
#pragma unroll
            for (int q = 0; q < NC; ++q) v[q] = p[q];
        };
        if (lo_k <= 0 && hi_k >= 2 * r + CPT)  // wave-uniform
            conv_run<T, NC, FAST, CPT>(a.w, r, load_in, acc);
        else
            conv_run<T, NC, FAST, CPT>(a.w, r, load, acc);
    } else {
        conv_run<T, NC, FAST, CPT>(a.w, r, load, acc);
    }
    const int64_t j0 = a.out_lo[ax] + ia * CPT;  // pass-axis position of output 0
    if constexpr (ADJ) {
        if (j0 <= a.clamp_lo + CPT - 1 || j0 + CPT - 1 >= a.clamp_hi) {  // uniform
            auto at = [&](int64_t i, T (&v)[NC]) {
                const T* p = src + (i - first) * sa;
#pragma unroll
                for (int q = 0; q < NC; ++q) v[q] = p[q];
            };
#pragma unroll
            for (int c = 0; c < CPT; ++c) adjoint_edges<T, NC>(a, j0 + c, at, acc[c]);
        }
    }
    if (!col_ok) return;
    T* dst = a.dst + ((a.out_lo[0] + col - a.dst_lo[0]) * a.dst_stride[0] +
                      (p1 - a.dst_lo[b1]) * a.dst_stride[b1] + (p2 - a.dst_lo[b2]) * a.dst_stride[b2]) * NC;
    const int64_t oa0 = j0 - a.dst_lo[ax];
    const int64_t da = a.dst_stride[ax] * NC;
    const int nval = (int)(a.n[ax] - ia * CPT);  // outputs of this run inside the region
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        if (c < nval) {
            T* o = dst + (oa0 + c) * da;
#pragma unroll
            for (int q = 0; q < NC; ++q) o[q] = (ADJ && a.accumulate) ? o[q] + acc[c][q] : acc[c][q];
        }
    }
}

// Staged row layout: runs of CPT inputs, each followed by one pad pixel, so
// that lane l (reading input l * CPT + k) hits bank l * (CPT + 1) + const.
template <int CPT>
__device__ __forceinline__ int staged(int i) { return (i % CPT) + (i / CPT) * (CPT + 1); }

// Pass along axis 0 (forward, or the adjoint with ADJ).  Dynamic LDS: 4
// waves x row_lds_bytes() / 4.
template <typename T, int NC, bool FAST, bool ADJ, int CPT>
__global__ __launch_bounds__(256) void axis_row_kernel([[maybe_unused]] const AxisPassArgs<T> args) {
    const AxisPassArgs<T>& a = *(const AxisPassArgs<T>*)__builtin_amdgcn_kernarg_segment_ptr();
    extern __shared__ __attribute__((aligned(16))) unsigned char row_lds_raw[];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = a.radius;
    // inputs of one segment, plus the ring's read-ahead past the last tap
    const int span = 64 * CPT + 2 * r + CPT;
    const int per_wave = ((span - 1) / CPT + 1) * (CPT + 1);
    T* buf = reinterpret_cast<T*>(row_lds_raw) + (int64_t)wave * per_wave * NC;
    const int64_t item = xcd_block() * 4 + wave;
    const int64_t nseg = (a.n[0] + 64 * CPT - 1) / (64 * CPT);
    int64_t it = item;
    const int64_t is = it % nseg;
    it /= nseg;
    const int64_t o1 = it % a.n[1];
    it /= a.n[1];
    const int64_t o2 = it % a.n[2];
    const int64_t o3 = it / a.n[2];
    const bool live = o3 < a.n[3];  // wave-uniform; dead waves only skip memory
    const int64_t q1 = a.out_lo[1] + o1, q2 = a.out_lo[2] + o2, q3 = a.out_lo[3] + o3;
    const int64_t first = a.out_lo[0] + is * 64 * CPT - r;  // axis-0 position of input 0
    const int lo_k = (int)((ADJ ? a.g_lo : a.clamp_lo) - first);
    const int hi_k = (int)((ADJ ? a.g_hi : a.clamp_hi) - first);
    const int64_t s0 = a.src_stride[0] * NC;
    const T* src = a.src + ((q1 - a.src_lo[1]) * a.src_stride[1] + (q2 - a.src_lo[2]) * a.src_stride[2] +
                            (q3 - a.src_lo[3]) * a.src_stride[3] + (first - a.src_lo[0]) * a.src_stride[0]) * NC;
    if (live) {
        // All of a lane's loads are issued before its first LDS write (one
        // memory latency per wave): KB rounds of 64 cover spans up to
        // 64 * KB, i.e. radii up to 32 * (KB - CPT) - CPT / 2; longer spans
        // finish in the loop after.  The loads need no guard (the clamp
        // keeps every address inside the input extent); the writes past the
        // span are skipped.
        constexpr int KB = CPT + 2;
        T v[KB][NC];
#pragma unroll
        for (int k = 0; k < KB; ++k) {
            const int i = lane + 64 * k;
            const int ic = i < lo_k ? lo_k : (i > hi_k ? hi_k : i);
            const T* p = src + (int64_t)ic * s0;
#pragma unroll
            for (int q = 0; q < NC; ++q) v[k][q] = p[q];
        }
#pragma unroll
        for (int k = 0; k < KB; ++k) {
            const int i = lane + 64 * k;
            if (i < span) {
                const bool in = i >= lo_k && i <= hi_k;
                T* d = buf + staged<CPT>(i) * NC;
#pragma unroll
                for (int q = 0; q < NC; ++q) d[q] = (ADJ && !in) ? T(0) : v[k][q];
            }
        }
#pragma unroll 2
        for (int i = lane + 64 * KB; i < span; i += 64) {
            const bool in = i >= lo_k && i <= hi_k;
            const int ic = i < lo_k ? lo_k : (i > hi_k ? hi_k : i);
            const T* p = src + (int64_t)ic * s0;
            T* d = buf + staged<CPT>(i) * NC;
#pragma unroll
            for (int q = 0; q < NC; ++q) d[q] = (ADJ && !in) ? T(0) : p[q];
        }
    }
    // The staged segment is private to this wave: order its LDS writes
    // before the other lanes' reads.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (!live) return;
    const int base = lane * CPT;
    const T* lane_buf = buf + lane * (CPT + 1) * NC;
    auto load = [&](int jb, auto off, T (&v)[NC]) {
        constexpr int o = decltype(off)::value;
        // input base + jb + o, jb a multiple of CPT
        const T* p = lane_buf + ((jb / CPT + o / CPT) * (CPT + 1) + o % CPT) * NC;
#pragma unroll
        for (int q = 0; q < NC; ++q) v[q] = p[q];
    };
    T acc[CPT][NC];
    conv_run<T, NC, FAST, CPT>(a.w, r, load, acc);
    const int64_t x0 = is * 64 * CPT + base;  // output index of acc[0] along axis 0
    if constexpr (ADJ) {
        const int64_t j0 = a.out_lo[0] + x0;
        // Edge terms (the two lanes holding the outputs at clamp_lo and
        // clamp_hi), in one loop for both lanes: each adds its terms to its
        // edge output in the order of adjoint_edges.  The gradient positions
        // they read lie within r of the edge output, inside this segment's
        // staged span [first, first + span): they are read from LDS.
        const int64_t lo_end = a.clamp_lo + r - 1 < a.g_hi ? a.clamp_lo + r - 1 : a.g_hi;
        const int64_t hi_beg = a.clamp_hi - r + 1 > a.g_lo ? a.clamp_hi - r + 1 : a.g_lo;
        const int64_t clo = a.clamp_lo - j0, chi = a.clamp_hi - j0;  // edge output index in the run
        const bool has_lo = clo >= 0 && clo < CPT, has_hi = chi >= 0 && chi < CPT;
        if (has_lo || has_hi) {  // at most two lanes
            // a lane holding both edges (an axis of at most CPT pixels) does
            // them one after the other, as adjoint_edges does
            auto edge = [&](const int64_t c, const int64_t i0, const int n, const int t0, const int dt) {
                T o[NC];
#pragma unroll
                for (int q = 0; q < NC; ++q) o[q] = acc[0][q];
#pragma unroll
                for (int cc = 1; cc < CPT; ++cc)
#pragma unroll
                    for (int q = 0; q < NC; ++q) o[q] = c == cc ? acc[cc][q] : o[q];
                const T* p0 = buf;
#pragma unroll 4
                for (int k = 0; k < n; ++k) {
                    const T* p = p0 + staged<CPT>((int)(i0 + k - first)) * NC;
                    const T t = a.tail[t0 + dt * k];
#pragma unroll
                    for (int q = 0; q < NC; ++q) o[q] = tap<T, true>(o[q], p[q], t);
                }
#pragma unroll
                for (int cc = 0; cc < CPT; ++cc)
#pragma unroll
                    for (int q = 0; q < NC; ++q) acc[cc][q] = c == cc ? o[q] : acc[cc][q];
            };
            // lo edge: i = g_lo .. lo_end, weight tail[clamp_lo - i + r];
            // hi edge: i = hi_beg .. g_hi, weight tail[i + r - clamp_hi].
            // The lane's first edge is chosen by value, so that the two edge
            // lanes run the same loop.
            const int n_lo = (int)(lo_end - a.g_lo + 1), n_hi = (int)(a.g_hi - hi_beg + 1);
            const int t_lo = (int)(a.clamp_lo - a.g_lo + r), t_hi = (int)(hi_beg + r - a.clamp_hi);
            edge(has_lo ? clo : chi, has_lo ? a.g_lo : hi_beg, has_lo ? n_lo : n_hi,
                 has_lo ? t_lo : t_hi, has_lo ? -1 : 1);
            if (has_lo && has_hi) edge(chi, hi_beg, n_hi, t_hi, 1);
        }
    }
    if (NC == 1 && a.dst_stride[0] == 1) {
        // Dense scalar rows: the segment's outputs go through this wave's LDS
        // (the same padded layout as the inputs, conflict-free both ways),
        // so that each store instruction writes 64 consecutive pixels
        // instead of one pixel every CPT.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();  // every lane is done reading its inputs
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int c = 0; c < CPT; ++c) buf[staged<CPT>(base + c)] = acc[c][0];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int64_t seg0 = is * 64 * CPT;  // output index of the segment's first pixel
        T* drow = a.dst + (a.out_lo[0] + seg0 - a.dst_lo[0]) + (q1 - a.dst_lo[1]) * a.dst_stride[1] +
                  (q2 - a.dst_lo[2]) * a.dst_stride[2] + (q3 - a.dst_lo[3]) * a.dst_stride[3];
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int i = lane + 64 * k;
            if (seg0 + i < a.n[0]) {
                const T v = buf[staged<CPT>(i)];
                drow[i] = (ADJ && a.accumulate) ? drow[i] + v : v;
            }
        }
        return;
    }
    T* dst = a.dst + ((a.out_lo[0] + x0 - a.dst_lo[0]) * a.dst_stride[0] +
                      (q1 - a.dst_lo[1]) * a.dst_stride[1] + (q2 - a.dst_lo[2]) * a.dst_stride[2] +
                      (q3 - a.dst_lo[3]) * a.dst_stride[3]) * NC;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        if (x0 + c < a.n[0]) {
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                T* o = dst + c * a.dst_stride[0] * NC + q;
                *o = (ADJ && a.accumulate) ? *o + acc[c][q] : acc[c][q];
            }
        }
    }
}


// Dynamic LDS of the row kernel (4 waves).
inline size_t row_lds_bytes(int cpt, int r, int nc, size_t elem) {
    const int span = 64 * cpt + 2 * r + cpt;
    return (size_t)4 * ((span - 1) / cpt + 1) * (cpt + 1) * nc * elem;
}

template <typename T, int NC, bool FAST, bool ADJ, int CPT>
hipError_t pass_launch(const AxisPassArgs<T>& a, hipStream_t s) {
    const int64_t other = a.total / a.n[a.axis];
    int64_t items;
    if constexpr (CPT <= 8) {
        if (a.axis == 0) {
            const int64_t seg = 64 * CPT;
            items = ((a.n[0] + seg - 1) / seg) * other;
            const size_t lds = row_lds_bytes(CPT, a.radius, NC, sizeof(T));
            const dim3 grid((unsigned)(((items + 3) / 4 + 7) / 8 * 8)), block(256);
            hipLaunchKernelGGL((axis_row_kernel<T, NC, FAST, ADJ, CPT>), grid, block, lds, s, a);
            return hipGetLastError();
        }
    }
    items = ((a.n[0] + 63) / 64) * ((a.n[a.axis] + CPT - 1) / CPT) * (other / a.n[0]);
    const dim3 grid((unsigned)(((items + 3) / 4 + 7) / 8 * 8)), block(256);
    hipLaunchKernelGGL((axis_strided_kernel<T, NC, FAST, ADJ, CPT>), grid, block, 0, s, a);
    return hipGetLastError();
}

// Strided f32 passes of scalar pixels from this radius on run 16 outputs per lane.
#ifndef ITKG_S16
#define ITKG_S16 8
#endif

template <typename T, int NC, bool FAST, bool ADJ>
hipError_t pass_cpt(const AxisPassArgs<T>& a, hipStream_t s) {
    if constexpr (std::is_same_v<T, float> && NC == 1) {
        if (a.axis != 0 && a.radius >= ITKG_S16) return pass_launch<T, NC, FAST, ADJ, 16>(a, s);
    }
    int cpt = 8;  // outputs per lane of a run
    // the row kernel's staging must fit the 64 KiB a workgroup gets by default
    while (a.axis == 0 && cpt > 1 && row_lds_bytes(cpt, a.radius, NC, sizeof(T)) > 65536) cpt /= 2;
    switch (cpt) {
        case 8: return pass_launch<T, NC, FAST, ADJ, 8>(a, s);
        case 4: return pass_launch<T, NC, FAST, ADJ, 4>(a, s);
        case 2: return pass_launch<T, NC, FAST, ADJ, 2>(a, s);
        default: return pass_launch<T, NC, FAST, ADJ, 1>(a, s);
    }
}

// A strided pass (axis != 0) over scalar f32 pixels whose rows are dense
// and line up in groups of 4 (axis-0 extent, axis-0 offsets between the
// output and the views, and every other stride multiples of 4, both views
// 16-byte aligned) runs as the 4-component kernel over groups of 4 adjacent
// pixels: each lane loads and stores 16 bytes (a wave 1 KiB of a row
// instead of 256 B), and every pixel gets the same taps in the same order.
// Measured at 512^3, radius 6: 285 -> 236 us per pass.  Not for radii from
// ITKG_S16 (the scalar kernel's 16-output runs are faster there: 300 vs 340
// us at radius 21) nor f64 (no change).
template <typename T>
bool quad_view(const AxisPassArgs<T>& a, int ncomp, AxisPassArgs<T>* q) {
    if (!std::is_same_v<T, float> || a.radius >= ITKG_S16) return false;
    if (ncomp != 1 || a.axis == 0 || a.src_stride[0] != 1 || a.dst_stride[0] != 1) return false;
    if (a.n[0] % 4 || (a.out_lo[0] - a.src_lo[0]) % 4 || (a.out_lo[0] - a.dst_lo[0]) % 4) return false;
    if (reinterpret_cast<uintptr_t>(a.src) % 16 || reinterpret_cast<uintptr_t>(a.dst) % 16) return false;
    for (int d = 1; d < ITKG_MAX_DIM; ++d)
        if (a.src_stride[d] % 4 || a.dst_stride[d] % 4) return false;
    *q = a;
    q->n[0] = a.n[0] / 4;
    q->total = a.total / 4;
    q->out_lo[0] = 0;  // axis 0 in groups of 4, counted from the output's first pixel
    q->src_lo[0] = -(a.out_lo[0] - a.src_lo[0]) / 4;
    q->dst_lo[0] = -(a.out_lo[0] - a.dst_lo[0]) / 4;
    for (int d = 1; d < ITKG_MAX_DIM; ++d) {
        q->src_stride[d] = a.src_stride[d] / 4;
        q->dst_stride[d] = a.dst_stride[d] / 4;
    }
    return true;
}

template <typename T, bool FAST, bool ADJ>
hipError_t pass_nc(const AxisPassArgs<T>& a, int ncomp, hipStream_t s) {
    AxisPassArgs<T> q;
    if (quad_view(a, ncomp, &q)) return pass_cpt<T, 4, FAST, ADJ>(q, s);
    switch (ncomp) {
        case 1: return pass_cpt<T, 1, FAST, ADJ>(a, s);
        case 2: return pass_cpt<T, 2, FAST, ADJ>(a, s);
        case 3: return pass_cpt<T, 3, FAST, ADJ>(a, s);
        case 4: return pass_cpt<T, 4, FAST, ADJ>(a, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

template <typename T>
hipError_t launch_axis_forward(const AxisPassArgs<T>& a, int ncomp, itkg_arith arith,
                               hipStream_t s) {
    if (a.total == 0) return hipSuccess;
    return arith == ITKG_FAST ? pass_nc<T, true, false>(a, ncomp, s)
                              : pass_nc<T, false, false>(a, ncomp, s);
}

// The adjoint pass runs one fused multiply-add per tap.
template <typename T>
hipError_t launch_axis_adjoint_range(const AxisPassArgs<T>& a, int ncomp, int64_t g_lo,
                                     int64_t g_hi, hipStream_t s) {
    if (a.total == 0) return hipSuccess;
    AxisPassArgs<T> b = a;
    b.g_lo = g_lo;
    b.g_hi = g_hi;
    return pass_nc<T, true, true>(b, ncomp, s);
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
