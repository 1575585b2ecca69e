// Fused one-round-trip 3-D Gaussian (f32 scalar pixels): forward and adjoint.
//
// Forward: GaussianFilter::generate_region (src/filters/gaussian.rs:59-90)
// for D = 3 (D = 2 is one plane with z radius 0) in one kernel.  Every pixel
// of the requested region gets exactly the convolve_axis passes of
// src/filters/conv.rs:11-41 (x, then y, then z), each rounded to f32, so the
// result is bit-identical to the per-axis kernel and to the reference
// (EXACT), or follows the same FMA chain as the per-axis FAST kernel.
//
// Adjoint: the transpose of that operator.  Per axis it is a zero-padded
// correlation with the (symmetric) weights plus, at the two clamp edges, the
// clamped taps folded onto the edge voxel: with P[m] = sum_{k<m} w_k,
//   g_in[0]   += sum_{m=1..r} P[m] * g_out[r - off - m]
//   g_in[n-1] += sum_{m=1..r} P[m] * g_out[n - 1 - off - r + m]
// (terms outside g_out are zero).  The transposed passes commute in exact
// arithmetic, so they run in the same x, y, z order as the forward.
//
// Workgroup: NW waves.  It owns a TX x TY destination tile in x/y (64 x TY,
// or 128 x 32 for the _W modes; one component of it, for interleaved
// pixels) and a run of TZ destination planes, and walks the source planes
// the run needs in z order, one step per plane (planes clamped at the image
// edge are filtered again; adjoint planes outside the source are zeros, so
// every step is the same).  The (TY+2r) x (TX+2r) source window of a plane
// is loaded as a flattened array
// with per-lane offsets computed once per tile, into one of NSET register
// sets (2, or 4 where the registers allow), NSET steps before it is stored
// to LDS.  Per step, with both LDS buffers (the window and the
// x-filtered tile) double-buffered:
//   - the y pass of plane u (CPT+2r LDS values per CPT outputs down a
//     column);
//   - one z step: every in-flight output of the lane's columns receives its
//     next tap (scatter form), so each output sums taps k = 0..2r in order --
//     EXACT: fl(v * w_k) added to an accumulator that starts at +0;
//     FAST/adjoint: one FMA per tap.  The ring of 2r+1 accumulators has fixed
//     slots: the slot that opens (zeroed) and the one that completes are
//     selected by a uniform index (one indexed register move each), and the
//     weights rotate instead of the data;
//   - the completed output plane is stored;
//   - plane u+2 goes from its register set into the window buffer plane u
//     has left, and the set is reloaded with plane u+2+NSET (with 4 sets,
//     after the x pass, which then has the set's registers);
//   - the x pass of plane u+1;
//   - one LDS-only barrier.
// The adjoint at small radii (Cfg::EARLY) issues the LDS reads of its first
// x-pass item right after the y pass, so that the z step covers their
// latency, and keeps all weights in scalar registers, with the next step's z
// weights loaded before the barrier: a scalar load is waited for with a full
// lgkmcnt drain, which would also wait for the LDS reads in flight.
//
// Addressing: window loads and dense output stores take a scalar base plus a
// 32-bit lane offset (the offsets are made opaque to the compiler at each
// use, so that it does not hoist their widening into 64-bit per-lane
// addresses).  Dense output is written non-temporally; interleaved
// components and the adjoint's accumulation take plain stores on a separate
// address path (a store merged from both kinds would lose the hint).  The
// persistent batch kernel holds more registers and uses these options, the
// 4 load sets and EARLY only where they fit without spilling.
#pragma once

#include <type_traits>
#include <utility>

#include "itkg_impl.hpp"

namespace itkg {

namespace {

// LDS-only workgroup barrier: waits for this wave's LDS traffic, not for
// global memory, so the prefetch loads stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// lds_barrier, then the same wait again where the compiler's wait tracking
// sees it: scalar loads issued before the barrier are known complete after
// it, so no later wait drains LDS reads for them.
__device__ __forceinline__ void lds_barrier_tracked() {
    lds_barrier();
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt: no wait
}

// Waves and output columns per lane, by z radius and mode: the z ring holds
// CPT * (2rz+1) accumulators per lane; the y tile is TY = NW * CPT rows.
//   FAST and adjoint: 16 waves; 4 columns (64-row tiles) up to rz 7, then 2;
//   their _S variants: 16 waves x 2 columns (32-row tiles) at every radius;
//   EXACT (a larger register need: the product and the sum round
//   separately): 16 waves x 4 columns (64-row tiles) up to rz 5, 16 x 2
//   at rz 6..8 and 15 (4 columns spill there: 20 VGPRs at rz 6); 8 x 4
//   (rz 9, 10) or 8 x 2 (rz 11..14), where 16 waves would spill.  The
//   64-row tiles measured 8-10% faster at 512^3 (rz 3, 5) and 22% on the
//   VecPixel radii (13, 13, 4), bit-identical.
template <int RXY, int RZ, int MODE>
struct Cfg {
    // the unsharp modes take the tiles and registers of their plain mode
    static constexpr int BM = fused_base_mode(MODE);
    static constexpr bool EXACT = fused_is_exact(MODE);
    static constexpr bool EXACT_MID = EXACT && RZ >= 9 && RZ <= 14;
    static constexpr int NW = !EXACT ? 16 : (EXACT_MID ? 8 : 16);
    static constexpr int NTH = 64 * NW;
    // Tile width: 64 (a wave's lanes are the tile's columns) or, for the
    // _W modes, 128: the waves alternate between the two 64-column
    // halves (wave w: half w % HALVES, row group w / HALVES), so each source
    // row of the window is one 128 + 2r segment.
    static constexpr int TX = fused_tile_x(BM);
    static constexpr int HALVES = TX / 64;
    static constexpr bool SMALL = BM == FUSED_FWD_FAST_S || BM == FUSED_ADJ_S;
    // radii from 18 up (FMA modes): one column per lane, so that the ring
    // (2r + 1 registers a column) fits
    // Anisotropic radii (x/y above z): the adjoint from x/y radius 11 with z
    // radius 6 or 7, and EXACT at x/y radius 11 or more with 8 waves, take
    // 2 columns where 4 would spill.
    static constexpr bool WIDE_XY = RXY >= 11 && RXY > RZ;
    static constexpr int CPT = EXACT
                                   ? ((EXACT_MID && RZ <= 10 && !WIDE_XY) || RZ <= 5 ? 4 : 2)
                                   : (RZ >= 18 ? 1
                                               : (RZ <= 7 && !SMALL &&
                                                          !(BM == FUSED_ADJ && WIDE_XY && RZ >= 6)
                                                      ? 4
                                                      : 2));
    static constexpr int TY = NW / HALVES * CPT;
    static constexpr int WX = TX + 2 * RXY;       // loaded window width = sIn row stride
    static constexpr int HY = TY + 2 * RXY;       // loaded window rows
    static constexpr int E = WX * HY;             // window elements per plane
    static constexpr int LJ = (E + NTH - 1) / NTH;  // window loads per lane and plane
    // x pass: each lane produces Q consecutive outputs of one row; a row
    // takes TX / Q lanes, so an item (one wave's 64 lanes) covers
    // RPI = 64 Q / TX rows.  Q in {4, 2, 1} is chosen to balance the
    // ceil(HY / RPI) items over the NW waves (every wave waits for the
    // slowest at the barrier that follows); Q-wide LDS reads need the row
    // stride WX to be a multiple of Q, and a row at most one wave's lanes.
    static constexpr int rpi(int q) { return 64 * q / TX; }
    static constexpr int xcost(int q) {
        if (WX % q || rpi(q) < 1) return 1 << 30;
        const int items = (HY + rpi(q) - 1) / rpi(q);
        const int per_wave = (items + NW - 1) / NW;
        return per_wave * (q * (2 * RXY + 1) + 2 * ((q + 2 * RXY + q - 1) / q));
    }
    static constexpr int Q = xcost(4) <= xcost(2) ? (xcost(4) <= xcost(1) ? 4 : 1)
                                                  : (xcost(2) <= xcost(1) ? 2 : 1);
    static constexpr int RPI = rpi(Q);                // rows per x item
    static constexpr int RG = (HY + RPI - 1) / RPI;   // x-pass items
    static constexpr int XQ = (Q + 2 * RXY + Q - 1) / Q;  // Q-wide reads per x item
    // Split x pass: full rounds of RPI-row items (every wave one item per
    // round), then the rows left over one row per item (QL = TX / 64
    // outputs per lane), dealt to the waves; used when its slowest wave has
    // less work.
    static constexpr int QL = HALVES;
    static constexpr int NA = (HY / RPI) / NW * NW;  // items of the full rounds
    static constexpr int RA = RPI * NA;              // rows they cover
    static constexpr int RL = HY - RA;               // rows left over
    static constexpr int xcost_split() {
        return (NA / NW) * (Q * (2 * RXY + 1) + 2 * XQ) +
               ((RL + NW - 1) / NW) * (QL * (2 * RXY + 1) + 2 * (QL + 2 * RXY));
    }
    static constexpr bool XSPLIT = Q > 1 && NA > 0 && xcost_split() < xcost(Q);
    // One sIn plane: the window (row stride WX), room for the unconditional
    // stores of the last loads (LJ * NTH) and for the x pass's reads of the
    // rows past HY in its last item (their outputs are never used).
    static constexpr int PLANE0 = LJ * NTH > RPI * RG * WX ? LJ * NTH : RPI * RG * WX;
    static constexpr int PLANE = (PLANE0 + 3) / 4 * 4;
    static constexpr int SX = TX;                 // sX row stride (floats)
    static constexpr int XPLANE = RPI * RG * SX;  // one sX buffer
    static constexpr int P = 2 * RZ + 1;
    // Adjoint at small radii (64-row tiles): the weights stay in scalar
    // registers for the whole tile and the first x-pass round's LDS reads
    // are issued before the z step.  Measured 4-6% faster for the adjoint;
    // no gain for forward FAST, and larger radii would spill.
    static constexpr bool EARLY = BM == FUSED_ADJ && RXY <= 7 && RZ <= 7;
    // Register sets of window loads: a plane is loaded NSET steps before it
    // is stored to LDS.  4 where the registers fit without spilling and it
    // measured faster: forward FAST with 64-row tiles up to radius 6 (126
    // VGPRs at r = 6; 4% at r = 6) and the FMA modes with 32-row tiles (2
    // columns per lane), forward and adjoint, up to radius 10 (2-3%); 2
    // elsewhere.
    static constexpr int NSET =
        !EXACT && RXY <= 10 && RZ <= 10 &&
                (CPT == 2 || ((BM == FUSED_FWD_FAST || BM == FUSED_FWD_FAST_W) && RXY <= 6 && RZ <= 6))
            ? 4
            : 2;
    static_assert(WX - TX <= 64, "halo wider than one wave");
    static_assert(HALVES == 1 ||
                      ((BM == FUSED_FWD_FAST_W || BM == FUSED_FWD_EXACT_W) && NW % HALVES == 0),
                  "128-wide tiles: forward FAST and EXACT only");
    static_assert((2 * PLANE + 2 * XPLANE) * 4 <= 160 * 1024, "LDS budget of one workgroup");
};

// The tile's source and destination pointers in the global address space.
// A pointer read from a device table (the batched kernels' argument blocks)
// is otherwise generic, and generic (flat) loads also count against the LDS
// wait counter: every wait for an LDS read would then wait for the window
// loads in flight as well.
template <class T>
using global_ptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ global_ptr<T> to_global(T* p) {
    return (global_ptr<T>)p;
}

// Compile-time loop: f(integral_constant<int, I>) for I = 0..N-1.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(std::integer_sequence<int, I...>, F&& f) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(std::make_integer_sequence<int, N>{}, f);
}

// Makes the compiler hold the N uniform values in scalar registers at this
// point (an empty asm use per value): loads of kernel arguments, which the
// compiler may otherwise move across the barrier asm, are complete here.
template <int N, int... I>
__device__ __forceinline__ void pin_sgprs_impl(const float (&v)[N], std::integer_sequence<int, I...>) {
    (..., [](float x) { asm volatile("" ::"s"(x)); }(v[I]));
}
template <int N>
__device__ __forceinline__ void pin_sgprs(const float (&v)[N]) {
    pin_sgprs_impl(v, std::make_integer_sequence<int, N>{});
}

// The z ring of column C.  Each column's ring is its own array so that the
// compiler keeps it in registers and reaches slots with an indexed move.
template <int C, class A>
__device__ __forceinline__ float* ring_of(A& a0, A& a1, A& a2, A& a3, A& a4, A& a5, A& a6, A& a7) {
    if constexpr (C == 0) return a0;
    else if constexpr (C == 1) return a1;
    else if constexpr (C == 2) return a2;
    else if constexpr (C == 3) return a3;
    else if constexpr (C == 4) return a4;
    else if constexpr (C == 5) return a5;
    else if constexpr (C == 6) return a6;
    else return a7;
}

// One output of a 1-D pass over window values v[0 .. 2r].
template <int R, int MODE>
__device__ __forceinline__ float tap_sum(const float* v, const float* w) {
    if constexpr (fused_is_exact(MODE)) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k <= 2 * R; ++k) {
            const float prod = v[k] * w[k];
            acc = acc + prod;
        }
        return acc;
    } else {
        float acc = v[0] * w[0];
#pragma unroll
        for (int k = 1; k <= 2 * R; ++k) acc = __builtin_fmaf(v[k], w[k], acc);
        return acc;
    }
}

// The 2r+1 weights of one axis, copied into scalar registers.
template <int N>
struct Weights {
    float w[N];
    __device__ __forceinline__ explicit Weights(const float* p) {
#pragma unroll
        for (int k = 0; k < N; ++k) w[k] = p[k];
    }
};
template <int R, int MODE, int N>
__device__ __forceinline__ float tap_sum(const float* v, const Weights<N>& w) {
    return tap_sum<R, MODE>(v, w.w);
}

// One destination tile (64 x TY x TZ) of one volume.  The caller's previous
// tile, if any, may still be using the LDS buffers (the tile starts with a
// barrier).  PERSIST: called from the persistent batch kernel, whose work
// loop holds more registers: there 64-row tiles keep two load sets, the
// adjoint keeps the plain step order and all outputs take plain stores.
template <int RXY, int RZ, int MODE, bool PERSIST, class ARGS = Fused3dArgs>
__device__ __forceinline__ void fused3d_tile(const ARGS& a, const int tile, float* sIn, float* sX,
                                             [[maybe_unused]] const double amount = 0.0,
                                             [[maybe_unused]] const double threshold = 0.0) {
    using C = Cfg<RXY, RZ, MODE>;
    constexpr int CPT = C::CPT, P = C::P, LJ = C::LJ;
    // The persistent kernel uses the register-hungry options only where its
    // registers allow them (no spills): the non-temporal dense stores and the
    // opaque load offsets in the FMA modes up to radius 6 and with 32-row
    // tiles up to radius 10; two load sets and the plain step order
    // elsewhere in 64-row tiles.
    constexpr bool FMA = !fused_is_exact(MODE);
    constexpr bool PFIT = FMA && ((RXY <= 6 && RZ <= 6) || (CPT == 2 && RXY <= 10 && RZ <= 10));
    constexpr int NSET = PERSIST && CPT == 4 ? 2 : C::NSET;
    constexpr bool EARLY = C::EARLY && !PERSIST && RZ > 0;
    static_assert(!EARLY || !fused_is_unsharp(MODE), "the unsharp loads follow the plain step order");
    constexpr bool PNT = !PERSIST || PFIT;   // non-temporal dense stores
    constexpr bool PASM = !PERSIST || PFIT;  // opaque load offsets (scalar-base loads)
    constexpr bool ADJOINT = fused_is_adjoint(MODE);
    // Radii from 16 up (forward): the x and y passes take their inputs one
    // LDS read at a time and add each into the outputs it belongs to, instead
    // of holding all 2r + Q (x) or 2r + CPT (y) inputs in registers first.
    // Every output still receives its taps k = 0..2r in order (the same FMA
    // chain as tap_sum).
    constexpr bool STREAM = !ADJOINT && FMA && (RXY >= 16 || (PERSIST && RXY >= 12));
    const int tx_i = tile % a.tiles_x;
    const int ty_i = (tile / a.tiles_x) % a.tiles_y;
    const int tz_i = tile / (a.tiles_x * a.tiles_y);
    const int x0 = tx_i * C::TX, y0 = ty_i * C::TY, z0 = tz_i * a.tz;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // this wave's 64-column half of the tile and its row group (y pass, z
    // ring and stores)
    const int half = C::HALVES > 1 ? wave % C::HALVES : 0;
    const int wrow = C::HALVES > 1 ? wave / C::HALVES : wave;

    const int nsx = a.n_src[0], nsy = a.n_src[1], nsz = a.n_src[2];
    const int ndx = a.n_dst[0], ndy = a.n_dst[1], ndz = a.n_dst[2];
    auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };

    // ---- window loads.  Window element e = row * WX + col (row-major, the
    // sIn layout) is loaded by lane e % 64 of wave (e / 64) % NW as its load
    // j = e / NTH.  The in-plane byte offsets are computed once per tile
    // (forward: clamped into the source; adjoint: zeros outside, kept as a
    // bit mask), so a plane's loads are a scalar base plus these offsets and
    // every step issues the same LJ loads.
    uint32_t gofs[LJ];
    uint32_t okmask = 0;
#pragma unroll
    for (int j = 0; j < LJ; ++j) {
        const int e = min((j * C::NW + wave) * 64 + lane, C::E - 1);
        const int row = e / C::WX, col = e - row * C::WX;
        const int py = y0 + a.shift[1] - RXY + row;
        const int px = x0 + a.shift[0] - RXY + col;
        if (ADJOINT && py >= 0 && py < nsy && px >= 0 && px < nsx) okmask |= 1u << j;
        gofs[j] = (uint32_t)(clampi(py, 0, nsy - 1) * a.src_stride[1] +
                             clampi(px, 0, nsx - 1) * a.src_stride[0]) *
                  (uint32_t)sizeof(float);
    }
    const int zv0 = z0 + a.shift[2] - RZ;  // source plane of step u = 0
    const global_ptr<const char> src_bytes = to_global(reinterpret_cast<const char*>(a.src));
    auto load_plane = [&](float* r, int u) {
        const int zr = clampi(zv0 + u, 0, nsz - 1);
        const global_ptr<const char> plane =
            src_bytes + (int64_t)zr * a.src_stride[2] * (int64_t)sizeof(float);
#pragma unroll
        for (int j = 0; j < LJ; ++j) {
            // the offset made opaque at every call, so that its widening is
            // not hoisted out of the loop: the load keeps the scalar-base +
            // 32-bit-offset form (no 64-bit address per lane and load)
            if constexpr (PASM) asm volatile("" : "+v"(gofs[j]));
            r[j] = *reinterpret_cast<const global_ptr<const float>>(plane + gofs[j]);
        }
    };
    float* const sIn_lane = sIn + wave * 64 + lane;
    auto store_plane = [&](const float* r, int half, int u) {
        float* d = sIn_lane + half * C::PLANE;
        if constexpr (ADJOINT) {
            const int zr = zv0 + u;
            const uint32_t m = (zr >= 0 && zr < nsz) ? okmask : 0u;
#pragma unroll
            for (int j = 0; j < LJ; ++j) d[j * C::NTH] = ((m >> j) & 1u) ? r[j] : 0.0f;
        } else {
#pragma unroll
            for (int j = 0; j < LJ; ++j) d[j * C::NTH] = r[j];
        }
    };

    const float* wx = a.wx;
    const float* wy = a.wy;
    // EARLY: the same weights in scalar registers, read ahead of the first
    // barrier (a scalar load still in flight at the loop head would make
    // every step's first LDS wait a full drain).
    [[maybe_unused]] const Weights<2 * RXY + 1> wxr(a.wx), wyr(a.wy);
    // The adjoint with 32-row tiles is short of scalar registers and, up to
    // radius 8, has vector registers to spare: its y weights live in vector
    // registers there (made lane-varying to the compiler by an empty asm).
    constexpr bool WYV = ADJOINT && CPT == 2 && !EARLY && RXY <= 8 && RZ <= 8;
    [[maybe_unused]] float wyv[2 * RXY + 1];
    if constexpr (WYV) {
#pragma unroll
        for (int k = 0; k <= 2 * RXY; ++k) {
            wyv[k] = a.wy[k];
            asm volatile("" : "+v"(wyv[k]));
        }
    }

    // Adjoint edge terms exist only in tiles that hold an image edge (tile-
    // uniform tests, so interior tiles skip them).
    const bool x_edge = x0 == 0 || x0 + C::TX >= ndx;

    // ---- x pass: items g = wave, wave+NW, ...; lane -> (row in item, Q-group)
    constexpr int Q = C::Q, LPR = C::TX / Q, RPI = C::RPI, QL = C::QL;
    static_assert(QL == 1 || (!STREAM && !ADJOINT), "two outputs per lane in the left-over rows");
    using vecq = std::conditional_t<Q == 4, float4, std::conditional_t<Q == 2, float2, float>>;
    const int xq = lane % LPR, xr = lane / LPR;
    auto x_pass = [&](const float* srcp, float* dstp) {
        constexpr int NG = C::XSPLIT ? C::NA : C::RG;  // Q-row items
#pragma unroll
        for (int g0 = 0; g0 < NG; g0 += C::NW) {
            const int g = g0 + wave;
            if (g < NG) {
                const int row = RPI * g + xr;
                const vecq* srcq = reinterpret_cast<const vecq*>(&srcp[row * C::WX + Q * xq]);
                float v[STREAM ? 1 : Q * C::XQ];
                float o[Q];
                if constexpr (STREAM) {
#pragma unroll
                    for (int i = 0; i < C::XQ; ++i) {
                        const vecq t = srcq[i];
                        float tv[Q];
                        if constexpr (Q == 4) {
                            tv[0] = t.x, tv[1] = t.y, tv[2] = t.z, tv[3] = t.w;
                        } else if constexpr (Q == 2) {
                            tv[0] = t.x, tv[1] = t.y;
                        } else {
                            tv[0] = t;
                        }
#pragma unroll
                        for (int q = 0; q < Q; ++q) {
#pragma unroll
                            for (int j = 0; j < Q; ++j) {
                                const int k = Q * i + q - j;  // tap of input Q*i+q in output j
                                if (k == 0) o[j] = tv[q] * wx[0];
                                else if (k > 0 && k <= 2 * RXY) o[j] = __builtin_fmaf(tv[q], wx[k], o[j]);
                            }
                        }
                    }
                } else {
#pragma unroll
                for (int i = 0; i < C::XQ; ++i) {
                    const vecq t = srcq[i];
                    if constexpr (Q == 4) {
                        v[4 * i + 0] = t.x;
                        v[4 * i + 1] = t.y;
                        v[4 * i + 2] = t.z;
                        v[4 * i + 3] = t.w;
                    } else if constexpr (Q == 2) {
                        v[2 * i + 0] = t.x;
                        v[2 * i + 1] = t.y;
                    } else {
                        v[i] = t;
                    }
                }
#pragma unroll
                for (int j = 0; j < Q; ++j) o[j] = tap_sum<RXY, MODE>(v + j, wx);
                }
                if constexpr (ADJOINT) {
                    const int xg = x0 + Q * xq;  // destination x of o[0]
                    const int jlo = -xg, jhi = ndx - 1 - xg;  // output index of x = 0 / ndx - 1
                    // one test per lane group; only the (at most two) lanes
                    // holding an edge output enter
                    if (x_edge && ((unsigned)jlo < (unsigned)Q || (unsigned)jhi < (unsigned)Q)) {
#pragma unroll
                        for (int j = 0; j < Q; ++j) {
                            if (j == jlo) {
#pragma unroll
                                for (int m = 1; m <= RXY; ++m)
                                    o[j] = __builtin_fmaf(a.px[m], v[j + 2 * RXY - m], o[j]);
                            }
                            if (j == jhi) {
#pragma unroll
                                for (int m = 1; m <= RXY; ++m)
                                    o[j] = __builtin_fmaf(a.px[m], v[j + m], o[j]);
                            }
                        }
                    }
                }
                vecq* dq = reinterpret_cast<vecq*>(&dstp[row * C::SX + Q * xq]);
                if constexpr (Q == 4) *dq = make_float4(o[0], o[1], o[2], o[3]);
                else if constexpr (Q == 2) *dq = make_float2(o[0], o[1]);
                else *dq = o[0];
            }
        }
        if constexpr (C::XSPLIT) {  // the rows left over: one output per lane
#pragma unroll
            for (int k0 = 0; k0 < C::RL; k0 += C::NW) {
                const int k = k0 + wave;
                if (k < C::RL) {
                    const int row = C::RA + k;
                    const float* rp = srcp + row * C::WX + QL * lane;
                    float o;
                    if constexpr (QL == 2) {
                        // 128-wide tiles: two outputs per lane, float2 reads
                        float v[2 * RXY + 2];
#pragma unroll
                        for (int i = 0; i <= RXY; ++i) {
                            const float2 t = reinterpret_cast<const float2*>(rp)[i];
                            v[2 * i] = t.x;
                            v[2 * i + 1] = t.y;
                        }
                        *reinterpret_cast<float2*>(&dstp[row * C::SX + 2 * lane]) =
                            make_float2(tap_sum<RXY, MODE>(v, wx), tap_sum<RXY, MODE>(v + 1, wx));
                        continue;
                    } else if constexpr (STREAM) {
                        o = rp[0] * wx[0];
#pragma unroll
                        for (int i = 1; i <= 2 * RXY; ++i) o = __builtin_fmaf(rp[i], wx[i], o);
                    } else {
                    float v[2 * RXY + 1];
#pragma unroll
                    for (int i = 0; i <= 2 * RXY; ++i) v[i] = rp[i];
                    o = tap_sum<RXY, MODE>(v, wx);
                    if constexpr (ADJOINT) {
                        // one test; only the (at most two) edge lanes enter
                        if (x_edge && (x0 + lane == 0 || x0 + lane == ndx - 1)) {
                            if (x0 + lane == 0) {
#pragma unroll
                                for (int m = 1; m <= RXY; ++m) o = __builtin_fmaf(a.px[m], v[2 * RXY - m], o);
                            }
                            if (x0 + lane == ndx - 1) {
#pragma unroll
                                for (int m = 1; m <= RXY; ++m) o = __builtin_fmaf(a.px[m], v[m], o);
                            }
                        }
                    }
                    }
                    dstp[row * C::SX + lane] = o;
                }
            }
        }
    };

    // ---- EARLY x pass, in two parts: x_read issues the LDS reads of the
    // first round's item (item `wave`) early in a step; x_finish computes and
    // writes it, then does the further rounds and the rows left over.
    constexpr int NGE = C::XSPLIT ? C::NA : C::RG;  // Q-row items
    const bool x_first = NGE >= C::NW || wave < NGE;  // wave-uniform; constant if every wave has one
    [[maybe_unused]] float xv[Q * C::XQ];
    auto x_read_item = [&](const float* srcp, const int g, float* v) {
        const vecq* srcq = reinterpret_cast<const vecq*>(&srcp[(RPI * g + xr) * C::WX + Q * xq]);
#pragma unroll
        for (int i = 0; i < C::XQ; ++i) {
            const vecq t = srcq[i];
            if constexpr (Q == 4) {
                v[4 * i + 0] = t.x;
                v[4 * i + 1] = t.y;
                v[4 * i + 2] = t.z;
                v[4 * i + 3] = t.w;
            } else if constexpr (Q == 2) {
                v[2 * i + 0] = t.x;
                v[2 * i + 1] = t.y;
            } else {
                v[i] = t;
            }
        }
    };
    auto x_item = [&](const float* v, float* dstp, const int g) {
        float o[Q];
#pragma unroll
        for (int j = 0; j < Q; ++j) o[j] = tap_sum<RXY, MODE>(v + j, wxr);
        if constexpr (ADJOINT) {
            const int xg = x0 + Q * xq;  // destination x of o[0]
            const int jlo = -xg, jhi = ndx - 1 - xg;  // output index of x = 0 / ndx - 1
            if (x_edge && ((unsigned)jlo < (unsigned)Q || (unsigned)jhi < (unsigned)Q)) {
#pragma unroll
                for (int j = 0; j < Q; ++j) {
                    if (j == jlo) {
#pragma unroll
                        for (int m = 1; m <= RXY; ++m)
                            o[j] = __builtin_fmaf(a.px[m], v[j + 2 * RXY - m], o[j]);
                    }
                    if (j == jhi) {
#pragma unroll
                        for (int m = 1; m <= RXY; ++m)
                            o[j] = __builtin_fmaf(a.px[m], v[j + m], o[j]);
                    }
                }
            }
        }
        vecq* dq = reinterpret_cast<vecq*>(&dstp[(RPI * g + xr) * C::SX + Q * xq]);
        if constexpr (Q == 4) *dq = make_float4(o[0], o[1], o[2], o[3]);
        else if constexpr (Q == 2) *dq = make_float2(o[0], o[1]);
        else *dq = o[0];
    };
    auto x_read = [&](const float* srcp) {
        if (x_first) x_read_item(srcp, wave, xv);
    };
    auto x_finish = [&](const float* srcp, float* dstp) {
        if (x_first) x_item(xv, dstp, wave);
#pragma unroll
        for (int g0 = C::NW; g0 < NGE; g0 += C::NW) {
            const int g = g0 + wave;
            if (g < NGE) {
                float v[Q * C::XQ];
                x_read_item(srcp, g, v);
                x_item(v, dstp, g);
            }
        }
        if constexpr (C::XSPLIT) {
#pragma unroll
            for (int k0 = 0; k0 < C::RL; k0 += C::NW) {
                const int k = k0 + wave;
                if (k < C::RL) {
                    const int row = C::RA + k;
                    const float* rp = srcp + row * C::WX + lane;
                    float v[2 * RXY + 1];
#pragma unroll
                    for (int i = 0; i <= 2 * RXY; ++i) v[i] = rp[i];
                    float o = tap_sum<RXY, MODE>(v, wxr);
                    if constexpr (ADJOINT) {
                        // one test; only the (at most two) edge lanes enter
                        if (x_edge && (x0 + lane == 0 || x0 + lane == ndx - 1)) {
                            if (x0 + lane == 0) {
#pragma unroll
                                for (int m = 1; m <= RXY; ++m) o = __builtin_fmaf(a.px[m], v[2 * RXY - m], o);
                            }
                            if (x0 + lane == ndx - 1) {
#pragma unroll
                                for (int m = 1; m <= RXY; ++m) o = __builtin_fmaf(a.px[m], v[m], o);
                            }
                        }
                    }
                    dstp[row * C::SX + lane] = o;
                }
            }
        }
    };

    // ---- y pass: CPT outputs down tile column xcol, rows wrow*CPT ...
    const int ybase = y0 + wrow * CPT;
    const int xcol = half * 64 + lane;  // this lane's tile column
    // wave-uniform: this wave's rows hold the first or the last output row
    [[maybe_unused]] const bool y_edge = ybase <= 0 || ybase + CPT > ndy - 1;
    float val[CPT];
    auto y_pass = [&](const float* xp) {
        if constexpr (STREAM) {
#pragma unroll
            for (int i = 0; i < CPT + 2 * RXY; ++i) {
                const float t = xp[(wrow * CPT + i) * C::SX + xcol];
#pragma unroll
                for (int c = 0; c < CPT; ++c) {
                    const int k = i - c;  // tap of input row i in output c
                    if (k == 0) val[c] = t * wy[0];
                    else if (k > 0 && k <= 2 * RXY) val[c] = __builtin_fmaf(t, wy[k], val[c]);
                }
            }
            return;
        }
        float v[CPT + 2 * RXY];
#pragma unroll
        for (int i = 0; i < CPT + 2 * RXY; ++i) v[i] = xp[(wrow * CPT + i) * C::SX + xcol];
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            if constexpr (EARLY) val[c] = tap_sum<RXY, MODE>(v + c, wyr);
            else if constexpr (WYV) val[c] = tap_sum<RXY, MODE>(v + c, wyv);
            else val[c] = tap_sum<RXY, MODE>(v + c, wy);
        }
        if constexpr (ADJOINT) {
            // y-edge terms: one wave-uniform test per step for the waves
            // away from the edges
            if (y_edge) {
#pragma unroll
                for (int c = 0; c < CPT; ++c) {
                    if (ybase + c == 0) {
#pragma unroll
                        for (int m = 1; m <= RXY; ++m)
                            val[c] = __builtin_fmaf(a.py[m], v[c + 2 * RXY - m], val[c]);
                    }
                    if (ybase + c == ndy - 1) {
#pragma unroll
                        for (int m = 1; m <= RXY; ++m)
                            val[c] = __builtin_fmaf(a.py[m], v[c + m], val[c]);
                    }
                }
            }
        }
    };

    // ---- output column of this lane.  Dense rows (x stride 1) are written
    // non-temporally at a wave-uniform base (scalar registers) plus the
    // lane's byte offset.
    const int ox = x0 + xcol;
    const bool x_ok = ox < ndx;
    const int64_t dxs = a.dst_stride[0], dys = a.dst_stride[1], dzs = a.dst_stride[2];
    const bool store_nt = PNT && dxs == 1 && !(ADJOINT && a.accumulate);  // tile-uniform
    const global_ptr<float> dwave =
        to_global(a.dst) + x0 + half * 64 + (int64_t)ybase * dys;  // used when dxs == 1
    uint32_t dlane = (uint32_t)lane * (uint32_t)sizeof(float);

    // ---- z ring: fixed slots, rotating weights (wz2 = wz repeated twice)
    static_assert(CPT <= 8, "at most 8 ring columns per lane");
    // A column's ring is one array of up to 32 slots (the widest vector the
    // compiler keeps in registers with indexed access) plus, for P > 32, a
    // second array with slots 32..P-1; the slot that opens and the one that
    // completes are each indexed within one of the two.
    constexpr int PL = P > 32 ? 32 : P, PH = P > 32 ? P - 32 : 1;
    static_assert(P <= 32 || !ADJOINT, "adjoint rings hold at most 32 slots");
    float r0[PL], r1[PL], r2[PL], r3[PL], r4[PL], r5[PL], r6[PL], r7[PL];
    [[maybe_unused]] float h0[PH], h1[PH], h2[PH], h3[PH], h4[PH], h5[PH], h6[PH], h7[PH];
#pragma unroll
    for (int p = 0; p < PL; ++p)
        r0[p] = r1[p] = r2[p] = r3[p] = r4[p] = r5[p] = r6[p] = r7[p] = 0.0f;
    if constexpr (P > 32) {
#pragma unroll
        for (int p = 0; p < PH; ++p)
            h0[p] = h1[p] = h2[p] = h3[p] = h4[p] = h5[p] = h6[p] = h7[p] = 0.0f;
    }
    const int tz_act = min(a.tz, ndz - z0);
    const int U = tz_act + 2 * RZ;
    const int t_hi = ndz - 1 - z0;  // step-relative index of the last destination plane
    const bool z_edge = z0 == 0 || z0 + tz_act >= ndz;
    // Adjoint z edges: the folded tail terms go straight into the ring slot
    // of the edge output (output t lives in slot t mod P from step t to
    // step t + 2r): slot 0 for the first plane, slot t_hi mod P for the last.
    const int s_hi = ((t_hi % P) + P) % P;

    // Step u, entered with plane u x-filtered in sX half u % 2, plane u + 1
    // in sIn half (u + 1) % 2 and plane u + 2 in register set `r`:
    //   y pass of plane u, one z step, store of the completed output plane;
    //   plane u + 2 into sIn half u % 2 (its plane u was read before the last
    //   barrier), `r` reloaded with plane u + 2 + NSET;
    //   x pass of plane u + 1 into sX half (u + 1) % 2 (read by step u - 1,
    //   before the last barrier);
    //   one barrier.
    // EARLY reads the first x item right after the y pass and runs the rest
    // of the x pass before the window store.
    int s = 0;
    // EARLY: this step's z weights wzs[j] = wz[(s - j) mod P], loop-carried
    // in scalar registers; the next step's are loaded before the barrier.
    [[maybe_unused]] float wzs[P];
    auto load_wz = [&]() {
#pragma unroll
        for (int j = 0; j < P; ++j) wzs[j] = a.wz2[s + P - j];
        pin_sgprs(wzs);
    };
    // Unsharp modes: the source pixels at the output pixels of step u (the
    // centre tap of each; indexes clamped into the tile, so that the loads
    // are unconditional and in bounds), loaded one step ahead and before
    // that step's window loads: the wait for them then never waits for the
    // newest window set.  At z radius 15 the registers held a step longer
    // would spill: there they are loaded where they are used.
    constexpr bool OG_AHEAD = fused_is_unsharp(MODE) && RZ < 15;
    [[maybe_unused]] float og[CPT];
    auto load_orig = [&](int u) {
        if constexpr (fused_is_unsharp(MODE)) {
            const int t = min(max(u - 2 * RZ, 0), tz_act - 1);
            const global_ptr<const float> scol =
                to_global(a.src) + (int64_t)(min(ox, ndx - 1) + a.shift[0]) * a.src_stride[0] +
                (int64_t)(z0 + t + a.shift[2]) * a.src_stride[2];
#pragma unroll
            for (int c = 0; c < CPT; ++c)
                og[c] = scol[(int64_t)(min(ybase + c, ndy - 1) + a.shift[1]) * a.src_stride[1]];
        }
    };
    // last: the tile's only step (a plane), which ends with its stores
    auto step = [&](int u, float* r, const int h, const bool last = false) {
        const int t = u - 2 * RZ;  // output completed at this step (if 0 <= t < tz_act)
        y_pass(sX + h * C::XPLANE);
        // EARLY: the first x-pass round's LDS reads go out now, so that the
        // z step and the stores cover their latency.
        if constexpr (EARLY) x_read(sIn + (1 - h) * C::PLANE);
        const int sn = (s + 1 == P) ? 0 : s + 1;
        const float* w = a.wz2 + s + P;  // w[-j] = wz[(s - j) mod P]
        // Adjoint z-edge weights of this plane (uniform scalar loads; zero
        // outside the first plane's and the last plane's windows).
        float plo = 0.0f, phi = 0.0f;
        if (ADJOINT && z_edge) {
            const int mlo = 2 * RZ - u, mhi = u - t_hi;
            if (z0 == 0 && mlo >= 1 && mlo <= RZ) plo = a.pz[mlo];
            if (mhi >= 1 && mhi <= RZ) phi = a.pz[mhi];
        }
        float o[CPT];
        static_for<CPT>([&](auto ci) {
            constexpr int c = decltype(ci)::value;
            float* acc = ring_of<c>(r0, r1, r2, r3, r4, r5, r6, r7);
            [[maybe_unused]] float* hac = ring_of<c>(h0, h1, h2, h3, h4, h5, h6, h7);
            // the output whose first tap is this plane
            if constexpr (P > 32) {
                // each array indexed within its bounds, whichever holds slot s
                // (a select between the two arrays' addresses would take both
                // out of registers)
                const int sl = min(s, PL - 1), sh = max(s - PL, 0);
                acc[sl] = s < PL ? 0.0f : acc[sl];
                hac[sh] = s >= PL ? 0.0f : hac[sh];
            } else {
                acc[s] = 0.0f;
            }
#pragma unroll
            for (int j = 0; j < P; ++j) {
                float& aj = j < PL ? acc[j] : hac[j - PL];
                if constexpr (fused_is_exact(MODE)) {
                    const float prod = val[c] * w[-j];
                    aj = aj + prod;
                } else if constexpr (EARLY) {
                    aj = __builtin_fmaf(val[c], wzs[j], aj);
                } else {
                    aj = __builtin_fmaf(val[c], w[-j], aj);
                }
            }
        });
        if constexpr (ADJOINT) {
            // z-edge terms, all columns behind one uniform test per step (a
            // test per column splits the step into blocks the scheduler
            // cannot interleave); plo / phi are zero outside the windows
            if (z_edge && (plo != 0.0f || phi != 0.0f)) {
                static_for<CPT>([&](auto ci) {
                    constexpr int c = decltype(ci)::value;
                    float* acc = ring_of<c>(r0, r1, r2, r3, r4, r5, r6, r7);
                    acc[0] = __builtin_fmaf(val[c], plo, acc[0]);
                    acc[s_hi] = __builtin_fmaf(val[c], phi, acc[s_hi]);
                });
            }
        }
        static_for<CPT>([&](auto ci) {
            constexpr int c = decltype(ci)::value;
            // tap 2r has just been added
            if constexpr (P > 32) {
                const float lo = ring_of<c>(r0, r1, r2, r3, r4, r5, r6, r7)[min(sn, PL - 1)];
                const float hi = ring_of<c>(h0, h1, h2, h3, h4, h5, h6, h7)[max(sn - PL, 0)];
                o[c] = sn < PL ? lo : hi;
            } else
                o[c] = ring_of<c>(r0, r1, r2, r3, r4, r5, r6, r7)[sn];
        });
        s = sn;
        if (t >= 0 && t < tz_act) {
            if constexpr (fused_is_unsharp(MODE)) {
                // UnsharpMaskFilter's step: the source pixel at the output
                // pixel's own index against its blurred value, in f64, each
                // product and sum rounded on its own
                if constexpr (!OG_AHEAD) load_orig(u);
#pragma unroll
                for (int c = 0; c < CPT; ++c) {
                    const double orig = (double)og[c];
                    const double edge = orig - (double)o[c];
                    const double scaled = amount * edge;
                    o[c] = (float)(fabs(edge) > threshold ? orig + scaled : orig);
                }
            }
            if (store_nt) {
                // dense output, non-temporal (keeps L2 for the input halos);
                // the lane offset is opaque here so that its widening stays
                // at the stores, which take the scalar-base + 32-bit-offset
                // form
                asm volatile("" : "+v"(dlane));
                const global_ptr<char> dplane =
                    reinterpret_cast<global_ptr<char>>(dwave + (int64_t)(z0 + t) * dzs);
#pragma unroll
                for (int c = 0; c < CPT; ++c)
                    if (x_ok && ybase + c < ndy)
                        __builtin_nontemporal_store(
                            o[c], reinterpret_cast<global_ptr<float>>(dplane + (int64_t)c * dys * 4 + dlane));
            } else {
                // one component of interleaved pixels (L2 merges the partial
                // lines), or the adjoint's accumulation; its own address
                // arithmetic, so that the compiler cannot merge these stores
                // with the non-temporal ones (a merge drops the hint)
                const global_ptr<float> dcol =
                    to_global(a.dst) + (int64_t)ox * dxs + (int64_t)ybase * dys + (int64_t)(z0 + t) * dzs;
#pragma unroll
                for (int c = 0; c < CPT; ++c) {
                    if (x_ok && ybase + c < ndy) {
                        if (ADJOINT && a.accumulate)
                            dcol[c * dys] = dcol[c * dys] + o[c];
                        else
                            dcol[c * dys] = o[c];
                    }
                }
            }
        }
        if (last) return;
        if constexpr (EARLY) {
            x_finish(sIn + (1 - h) * C::PLANE, sX + (1 - h) * C::XPLANE);
            store_plane(r, h, u + 2);
            load_plane(r, u + 2 + NSET);
            load_wz();
            lds_barrier_tracked();
        } else {
            store_plane(r, h, u + 2);
            if constexpr (NSET == 4) {
                // the x pass while `r` is free (the register peak), then the
                // reload of `r`
                x_pass(sIn + (1 - h) * C::PLANE, sX + (1 - h) * C::XPLANE);
                __builtin_amdgcn_sched_barrier(0);
                if constexpr (OG_AHEAD) load_orig(u + 1);
                load_plane(r, u + 2 + NSET);
            } else {
                if constexpr (OG_AHEAD) load_orig(u + 1);
                load_plane(r, u + 4);
                x_pass(sIn + (1 - h) * C::PLANE, sX + (1 - h) * C::XPLANE);
            }
            lds_barrier();
        }
    };

    float ra[LJ], rb[LJ], rc[LJ], rd[LJ];
    if constexpr (EARLY) {
        pin_sgprs(wxr.w);
        pin_sgprs(wyr.w);
    }
    if constexpr (RZ == 0) {
        // a plane (2-D image): one step and nothing to prefetch, where whole
        // rounds of NSET steps would repeat the loads, the x pass and the
        // y pass of the same clamped plane
        lds_barrier();
        load_plane(ra, 0);
        if constexpr (OG_AHEAD) load_orig(0);
        store_plane(ra, 0, 0);
        lds_barrier();
        x_pass(sIn, sX);
        lds_barrier();
        step(0, ra, 0, true);
        return;
    }
    lds_barrier();  // the previous tile of a persistent workgroup is done with LDS
    load_plane(ra, 0);
    load_plane(rb, 1);
    store_plane(ra, 0, 0);
    store_plane(rb, 1, 1);
    load_plane(ra, 2);
    if constexpr (NSET == 4) {
        // each set's loads kept together, in plane order, as the loop issues
        // them (the compiler's load waits in the loop assume that order)
        __builtin_amdgcn_sched_barrier(0);
        load_plane(rb, 3);
        __builtin_amdgcn_sched_barrier(0);
        load_plane(rc, 4);
        __builtin_amdgcn_sched_barrier(0);
        load_plane(rd, 5);
        __builtin_amdgcn_sched_barrier(0);
    } else {
        load_plane(rb, 3);
    }
    if constexpr (OG_AHEAD) load_orig(0);
    lds_barrier();
    if constexpr (EARLY) {
        x_read(sIn);
        x_finish(sIn, sX);
        load_wz();
        lds_barrier_tracked();
    } else {
        x_pass(sIn, sX);
        lds_barrier();
    }
    // Whole rounds of NSET steps: the extra last steps (if any) store nothing.
    if constexpr (NSET == 4) {
        for (int u = 0; u < U; u += 4) {
            step(u, ra, 0);
            step(u + 1, rb, 1);
            step(u + 2, rc, 0);
            step(u + 3, rd, 1);
        }
    } else {
        for (int u = 0; u < U; u += 2) {
            step(u, ra, 0);
            step(u + 1, rb, 1);
        }
    }
}

template <int RXY, int RZ, int MODE>
__global__ __launch_bounds__((Cfg<RXY, RZ, MODE>::NTH))
void fused3d_kernel([[maybe_unused]] const FusedMulti args) {
    using C = Cfg<RXY, RZ, MODE>;
    // The argument blocks are read in place from the kernel-argument segment
    // (where `args` lies, at offset 0), so that runtime-indexed weights are
    // scalar loads and no block is copied to private memory.
    const FusedMulti& m = *(const FusedMulti*)__builtin_amdgcn_kernarg_segment_ptr();
    __shared__ __attribute__((aligned(16))) float sIn[2 * C::PLANE];
    __shared__ __attribute__((aligned(16))) float sX[2 * C::XPLANE];
    // XCD-aware order of the (tile, component) items: hardware deals
    // workgroups to the 8 XCDs round robin, so consecutive items (the
    // components of one tile, then neighbouring tiles, which share halos and
    // cache lines) go to one XCD's L2.
    const int nc = m.nc;
    const Fused3dArgs& a0 = m.a[0];
    const int nt = a0.tiles_x * a0.tiles_y * a0.tiles_z * nc;
    const int per = (nt + 7) / 8;
    const int item = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (item >= nt) return;
    const int comp = item % nc;
    fused3d_tile<RXY, RZ, MODE, false>(m.a[comp], item / nc, sIn, sX, m.amount, m.threshold);
}

// Persistent form over a packed work list of (volume, tile) items for many
// volumes of one radius: workgroup b takes items b, b + G, b + 2G, ...
// The host interleaves the list so that items b and b + G share an XCD.
template <int RXY, int RZ, int MODE>
__global__ __launch_bounds__((Cfg<RXY, RZ, MODE>::NTH))
void fused3d_batch_kernel(const Fused3dArgs* __restrict__ vols, const int2* __restrict__ work,
                          const int nwork) {
    using C = Cfg<RXY, RZ, MODE>;
    __shared__ __attribute__((aligned(16))) float sIn[2 * C::PLANE];
    __shared__ __attribute__((aligned(16))) float sX[2 * C::XPLANE];
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work[w];
        if (it.x < 0) continue;  // padding of the interleaved list
        fused3d_tile<RXY, RZ, MODE, true>(vols[it.x], it.y, sIn, sX);
    }
}

// ---- the single persistent launch of a batched plan ---------------------
// One kernel serves every fused volume of a plan: each work item names a
// volume, a tile of it and a class -- the radius r (1..kMegaMaxR for FAST,
// 1..15 for the adjoint) on the family's own tile shape (64 rows up to
// radius 7, 32 rows beyond), or kMegaSmall + r for radius r (<= 7) on
// 32-row tiles -- and the workgroup runs that class's tile.  Items sit in 8
// queues, one per XCD (workgroup b runs on XCD b mod 8), each in decreasing
// order of modelled cost; a workgroup takes the next item of its own queue
// with one atomic add and, once that queue is empty, takes items from the
// others.  The last workgroup to finish resets the 8 queue heads and the
// finish count, so the next launch (or graph replay) starts from zero with
// no memset.
constexpr int kMegaSmall = 32;
constexpr int kMegaMaxR = kFusedMaxRBatch;
constexpr int kMegaQueues = 8;

template <int FAMILY>
constexpr int mega_mode(int cls) {
    return cls > kMegaSmall ? (FAMILY == FUSED_FWD_FAST ? FUSED_FWD_FAST_S : FUSED_ADJ_S) : FAMILY;
}
template <int FAMILY>
constexpr int mega_max_r() {
    return FAMILY == FUSED_FWD_FAST ? kMegaMaxR : 15;
}

template <int FAMILY, int CLS>
__device__ __forceinline__ void mega_case(const Fused3dArgsB& a, int tile, float* lds) {
    constexpr int R = CLS > kMegaSmall ? CLS - kMegaSmall : CLS;
    constexpr int MODE = mega_mode<FAMILY>(CLS);
    fused3d_tile<R, R, MODE, true, Fused3dArgsB>(a, tile, lds, lds + 2 * Cfg<R, R, MODE>::PLANE);
}

template <int FAMILY, int... I>
__device__ __forceinline__ void mega_dispatch(std::integer_sequence<int, I...>, int cls,
                                              const Fused3dArgsB& a, int tile, float* lds) {
    // classes 1..maxR, then kMegaSmall + 1..7
    (void)((cls == (I < mega_max_r<FAMILY>() ? I + 1 : kMegaSmall + I - mega_max_r<FAMILY>() + 1)
                ? (mega_case<FAMILY, (I < mega_max_r<FAMILY>() ? I + 1
                                                               : kMegaSmall + I - mega_max_r<FAMILY>() + 1)>(
                       a, tile, lds),
                   true)
                : false) ||
           ...);
}

// `trace` (null unless the caller asks for a breakdown) receives, per work
// item w, the start and end wall-clock ticks and the workgroup.
template <int FAMILY>
__global__ __launch_bounds__(1024) void fused3d_mega_kernel(const Fused3dArgsB* __restrict__ vols,
                                                            const int4* __restrict__ work,
                                                            const int* __restrict__ qstart,
                                                            int* __restrict__ heads,
                                                            unsigned long long* __restrict__ trace) {
    extern __shared__ __attribute__((aligned(16))) float mega_lds[];
    __shared__ int s_item;
    const int home = (int)(blockIdx.x % kMegaQueues);
    int tries = 0;  // queues found empty so far, from `home` on (they stay empty)
    for (;;) {
        if (threadIdx.x == 0) {
            int w = -1;
            while (tries < kMegaQueues) {
                const int q = (home + tries) % kMegaQueues;
                const int n = qstart[q + 1] - qstart[q];
                if (__hip_atomic_load(&heads[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < n) {
                    const int k = atomicAdd(&heads[q], 1);
                    if (k < n) {
                        w = qstart[q] + k;
                        break;
                    }
                }
                ++tries;
            }
            s_item = w;
        }
        __syncthreads();
        // uniform: the item's table entries then come in with scalar loads
        const int w = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();  // every wave has read s_item before lane 0 writes the next
        if (w < 0) break;
        const int4 it = work[w];
        const unsigned long long t0 = trace ? wall_clock64() : 0ull;
        mega_dispatch<FAMILY>(std::make_integer_sequence<int, mega_max_r<FAMILY>() + 7>{}, it.z,
                              vols[it.x], it.y, mega_lds);
        if (trace) {
            __syncthreads();
            if (threadIdx.x == 0) {
                trace[3 * w] = t0;
                trace[3 * w + 1] = wall_clock64();
                trace[3 * w + 2] = blockIdx.x;
            }
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const int done = atomicAdd(&heads[kMegaQueues], 1);
        if (done == (int)gridDim.x - 1) {
            // every other workgroup has made its last queue access
            for (int q = 0; q <= kMegaQueues; ++q) heads[q] = 0;
            __threadfence();
        }
    }
}

// LDS bytes and tile rows of a class (host and device).
template <int FAMILY, int CLS>
constexpr int mega_lds_of() {
    constexpr int R = CLS > kMegaSmall ? CLS - kMegaSmall : CLS;
    using C = Cfg<R, R, mega_mode<FAMILY>(CLS)>;
    return (2 * C::PLANE + 2 * C::XPLANE) * (int)sizeof(float);
}
template <int FAMILY, int CLS>
constexpr int mega_ty_of() {
    constexpr int R = CLS > kMegaSmall ? CLS - kMegaSmall : CLS;
    return Cfg<R, R, mega_mode<FAMILY>(CLS)>::TY;
}
template <int FAMILY, int... I>
void mega_table(std::integer_sequence<int, I...>, int* lds, int* ty) {
    constexpr int M = mega_max_r<FAMILY>();
    ((lds[(I < M ? I + 1 : kMegaSmall + I - M + 1)] =
          mega_lds_of<FAMILY, (I < M ? I + 1 : kMegaSmall + I - M + 1)>(),
      ty[(I < M ? I + 1 : kMegaSmall + I - M + 1)] =
          mega_ty_of<FAMILY, (I < M ? I + 1 : kMegaSmall + I - M + 1)>()),
     ...);
}

// The 128-wide tiles (plain launches) are built for 3-D isotropic radii
// where they measured at least as fast as 64 x 64 tiles: 1..5 in FAST, 1..4
// in EXACT.
template <int RXY, int RZ>
constexpr bool wide_built() { return RXY >= 1 && RXY <= 5 && RZ == RXY; }
template <int RXY, int RZ>
constexpr bool wide_exact_built() { return RXY >= 1 && RXY <= 4 && RZ == RXY; }
// Tile rows of the 128-wide shape, 0 where it is not built.
template <int RXY, int RZ>
constexpr int wide_ty() {
    if constexpr (wide_built<RXY, RZ>()) return Cfg<RXY, RZ, FUSED_FWD_FAST_W>::TY;
    else return 0;
}
template <int RXY, int RZ>
constexpr int wide_exact_ty() {
    if constexpr (wide_exact_built<RXY, RZ>()) return Cfg<RXY, RZ, FUSED_FWD_EXACT_W>::TY;
    else return 0;
}
// The unsharp modes: the radii UnsharpMaskFilter's index-space kernel
// gives, the same on every axis (3-D) or on x and y (2-D, z radius 0).
// EXACT at radius 15 would spill (12 bytes a lane): that radius keeps the
// blur and the step as two launches.
template <int RXY, int RZ, int MODE>
constexpr bool unsharp_built() {
    return (RZ == RXY || RZ == 0) && !(MODE == FUSED_FWD_EXACT_U && RZ == 15);
}
template <int RXY, int RZ, int MODE>
constexpr int unsharp_ty() {
    if constexpr (unsharp_built<RXY, RZ, MODE>()) return Cfg<RXY, RZ, MODE>::TY;
    else return 0;
}
// Workgroup size of mode `mode` for (RXY, RZ).
template <int RXY, int RZ>
int block_of(int mode) {
    switch (mode) {
        case FUSED_FWD_EXACT: return Cfg<RXY, RZ, FUSED_FWD_EXACT>::NTH;
        case FUSED_FWD_FAST: return Cfg<RXY, RZ, FUSED_FWD_FAST>::NTH;
        case FUSED_ADJ: return Cfg<RXY, RZ, FUSED_ADJ>::NTH;
        case FUSED_FWD_FAST_S: return Cfg<RXY, RZ, FUSED_FWD_FAST_S>::NTH;
        case FUSED_FWD_FAST_W: return 1024;
        case FUSED_FWD_EXACT_W: return 1024;
        case FUSED_FWD_EXACT_U: return Cfg<RXY, RZ, FUSED_FWD_EXACT_U>::NTH;
        case FUSED_FWD_FAST_U: return Cfg<RXY, RZ, FUSED_FWD_FAST_U>::NTH;
        default: return Cfg<RXY, RZ, FUSED_ADJ_S>::NTH;
    }
}

template <int RXY, int RZ>
hipError_t launch_r(const FusedMulti& a, int mode, hipStream_t st) {
    const int nt = a.a[0].tiles_x * a.a[0].tiles_y * a.a[0].tiles_z * a.nc;
    const dim3 grid(((nt + 7) / 8) * 8), block(block_of<RXY, RZ>(mode));
    switch (mode) {
        case FUSED_FWD_EXACT:
            hipLaunchKernelGGL((fused3d_kernel<RXY, RZ, FUSED_FWD_EXACT>), grid, block, 0, st, a);
            break;
        case FUSED_FWD_FAST:
            hipLaunchKernelGGL((fused3d_kernel<RXY, RZ, FUSED_FWD_FAST>), grid, block, 0, st, a);
            break;
        case FUSED_ADJ:
            hipLaunchKernelGGL((fused3d_kernel<RXY, RZ, FUSED_ADJ>), grid, block, 0, st, a);
            break;
        case FUSED_FWD_FAST_S:
            hipLaunchKernelGGL((fused3d_kernel<RXY, RZ, FUSED_FWD_FAST_S>), grid, block, 0, st, a);
            break;
        case FUSED_ADJ_S:
            hipLaunchKernelGGL((fused3d_kernel<RXY, RZ, FUSED_ADJ_S>), grid, block, 0, st, a);
            break;
        case FUSED_FWD_FAST_W:
            if constexpr (wide_built<RXY, RZ>()) {
                hipLaunchKernelGGL((fused3d_kernel<RXY, RZ, FUSED_FWD_FAST_W>), grid, block, 0, st, a);
                break;
            }
            return hipErrorInvalidValue;
        case FUSED_FWD_EXACT_W:
            if constexpr (wide_exact_built<RXY, RZ>()) {
                hipLaunchKernelGGL((fused3d_kernel<RXY, RZ, FUSED_FWD_EXACT_W>), grid, block, 0, st, a);
                break;
            }
            return hipErrorInvalidValue;
        case FUSED_FWD_EXACT_U:
            if constexpr (unsharp_built<RXY, RZ, FUSED_FWD_EXACT_U>()) {
                hipLaunchKernelGGL((fused3d_kernel<RXY, RZ, FUSED_FWD_EXACT_U>), grid, block, 0, st, a);
                break;
            }
            return hipErrorInvalidValue;
        case FUSED_FWD_FAST_U:
            if constexpr (unsharp_built<RXY, RZ, FUSED_FWD_FAST_U>()) {
                hipLaunchKernelGGL((fused3d_kernel<RXY, RZ, FUSED_FWD_FAST_U>), grid, block, 0, st, a);
                break;
            }
            return hipErrorInvalidValue;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int RXY, int RZ>
hipError_t launch_batch_r(const Fused3dArgs* vols, const int2* work, int nwork, int grid_blocks,
                          int mode, hipStream_t st) {
    const dim3 grid(grid_blocks), block(block_of<RXY, RZ>(mode));
    switch (mode) {
        case FUSED_FWD_EXACT:
            hipLaunchKernelGGL((fused3d_batch_kernel<RXY, RZ, FUSED_FWD_EXACT>), grid, block, 0, st,
                               vols, work, nwork);
            break;
        case FUSED_FWD_FAST:
            hipLaunchKernelGGL((fused3d_batch_kernel<RXY, RZ, FUSED_FWD_FAST>), grid, block, 0, st,
                               vols, work, nwork);
            break;
        case FUSED_ADJ:
            hipLaunchKernelGGL((fused3d_batch_kernel<RXY, RZ, FUSED_ADJ>), grid, block, 0, st, vols,
                               work, nwork);
            break;
        case FUSED_FWD_FAST_S:
            hipLaunchKernelGGL((fused3d_batch_kernel<RXY, RZ, FUSED_FWD_FAST_S>), grid, block, 0, st,
                               vols, work, nwork);
            break;
        case FUSED_ADJ_S:
            hipLaunchKernelGGL((fused3d_batch_kernel<RXY, RZ, FUSED_ADJ_S>), grid, block, 0, st, vols,
                               work, nwork);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Resident workgroups per CU of the persistent batch kernel.
template <int RXY, int RZ>
int occupancy_r(int mode) {
    int blocks = 0;
    hipError_t e = hipErrorInvalidValue;
    auto query = [&](auto kern, int nth) {
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kern, nth, 0);
    };
    if (mode == FUSED_FWD_EXACT)
        e = query(fused3d_batch_kernel<RXY, RZ, FUSED_FWD_EXACT>, Cfg<RXY, RZ, FUSED_FWD_EXACT>::NTH);
    else if (mode == FUSED_FWD_FAST)
        e = query(fused3d_batch_kernel<RXY, RZ, FUSED_FWD_FAST>, Cfg<RXY, RZ, FUSED_FWD_FAST>::NTH);
    else if (mode == FUSED_ADJ)
        e = query(fused3d_batch_kernel<RXY, RZ, FUSED_ADJ>, Cfg<RXY, RZ, FUSED_ADJ>::NTH);
    else if (mode == FUSED_FWD_FAST_S)
        e = query(fused3d_batch_kernel<RXY, RZ, FUSED_FWD_FAST_S>, Cfg<RXY, RZ, FUSED_FWD_FAST_S>::NTH);
    else if (mode == FUSED_ADJ_S)
        e = query(fused3d_batch_kernel<RXY, RZ, FUSED_ADJ_S>, Cfg<RXY, RZ, FUSED_ADJ_S>::NTH);
    return e == hipSuccess && blocks > 0 ? blocks : 1;
}

// Registry entry of one (rxy, rz) instantiation.
template <int RXY, int RZ>
constexpr FusedEntry make_entry() {
    return FusedEntry{RXY, RZ,
                      {Cfg<RXY, RZ, FUSED_FWD_EXACT>::TY, Cfg<RXY, RZ, FUSED_FWD_FAST>::TY,
                       Cfg<RXY, RZ, FUSED_ADJ>::TY, Cfg<RXY, RZ, FUSED_FWD_FAST_S>::TY,
                       Cfg<RXY, RZ, FUSED_ADJ_S>::TY, wide_ty<RXY, RZ>(), wide_exact_ty<RXY, RZ>(),
                       unsharp_ty<RXY, RZ, FUSED_FWD_EXACT_U>(), unsharp_ty<RXY, RZ, FUSED_FWD_FAST_U>()},
                      &launch_r<RXY, RZ>, &launch_batch_r<RXY, RZ>, &occupancy_r<RXY, RZ>};
}
template <int RXY, int RZ>
constexpr FusedEntry make_entry_single() {
    return FusedEntry{RXY, RZ,
                      {Cfg<RXY, RZ, FUSED_FWD_EXACT>::TY, Cfg<RXY, RZ, FUSED_FWD_FAST>::TY,
                       Cfg<RXY, RZ, FUSED_ADJ>::TY, Cfg<RXY, RZ, FUSED_FWD_FAST_S>::TY,
                       Cfg<RXY, RZ, FUSED_ADJ_S>::TY, wide_ty<RXY, RZ>(), wide_exact_ty<RXY, RZ>(),
                       unsharp_ty<RXY, RZ, FUSED_FWD_EXACT_U>(), unsharp_ty<RXY, RZ, FUSED_FWD_FAST_U>()},
                      &launch_r<RXY, RZ>, nullptr, nullptr};
}

}  // namespace

}  // namespace itkg
