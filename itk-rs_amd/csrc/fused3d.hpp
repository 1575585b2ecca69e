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
// Workgroup: NW waves.  It owns a 64 x TY destination tile in x/y and a run
// of TZ destination planes, and walks the source planes the run needs in z
// order, one step per plane (planes clamped at the image edge are filtered
// again; adjoint planes outside the source are zeros, so every step is the
// same).  Per step:
//   - the x pass of the step's plane from one half of a double-buffered LDS
//     window ((TY+2r) x (64+2r), filled by coalesced row loads; forward:
//     clamped at load time, adjoint: zero outside);
//   - the next plane's window, loaded into registers two steps earlier, is
//     stored into the other half, and the loads of the plane two steps ahead
//     are issued (two register sets alternate, the step loop is unrolled by
//     two, so the loads have two steps to arrive);
//   - the y pass (CPT+2r LDS values per CPT outputs down a column);
//   - one z step: every in-flight output of the lane's columns receives its
//     next tap (scatter form), so each output sums taps k = 0..2r in order --
//     EXACT: fl(v * w_k) added to an accumulator that starts at +0;
//     FAST/adjoint: one FMA per tap.  The ring of 2r+1 accumulators has fixed
//     slots: the slot that opens (zeroed) and the one that completes are
//     selected by a uniform index (one indexed register move each), and the
//     weights rotate instead of the data;
//   - the completed output plane is stored.
#pragma once

#include <type_traits>
#include <utility>

#include "itkg_impl.hpp"

// Minimum resident waves per SIMD the register allocator must keep (1: no
// constraint).
#ifndef ITKG_WAVES_PER_EU
#define ITKG_WAVES_PER_EU 1
#endif

namespace itkg {

namespace {

// LDS-only workgroup barrier: waits for this wave's LDS traffic, not for
// global memory, so the prefetch loads stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Waves and output columns per lane, by z radius: the z ring holds
// CPT * (2rz+1) accumulators per lane; the y tile is TY = NW * CPT rows.
template <int RXY, int RZ>
struct Cfg {
    static constexpr int NW = RZ <= 3 ? 4 : 8;
    static constexpr int NTH = 64 * NW;
    static constexpr int TX = 64;
    static constexpr int CPT = RZ <= 3 ? 8 : (RZ <= 10 ? 4 : 2);
    static constexpr int TY = NW * CPT;
    static constexpr int WX = TX + 2 * RXY;       // loaded window width
    static constexpr int HY = TY + 2 * RXY;       // loaded window rows
    static constexpr int RG = (HY + 3) / 4;       // x-pass row groups of 4
    static constexpr int ROWS = 4 * RG;           // LDS rows allocated
    static constexpr int SIN = (WX + 3) / 4 * 4;  // sIn row stride (floats)
    static constexpr int PLANE = ROWS * SIN;      // one half of the sIn double buffer
    static constexpr int SX = TX;                 // sX row stride (floats)
    static constexpr int XQ = (4 + 2 * RXY + 3) / 4;  // float4 reads per x item
    static constexpr int LROWS = (HY + NW - 1) / NW;  // rows loaded per wave
    static constexpr int P = 2 * RZ + 1;
    static_assert(WX - 64 <= 64, "halo wider than one wave");
};

// Compile-time loop: f(integral_constant<int, I>) for I = 0..N-1.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(std::integer_sequence<int, I...>, F&& f) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(std::make_integer_sequence<int, N>{}, f);
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
    if constexpr (MODE == FUSED_FWD_EXACT) {
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

// One destination tile (64 x TY x TZ) of one volume.  The caller's previous
// tile, if any, may still be using the LDS tiles (the tile starts with a
// barrier).
template <int RXY, int RZ, int MODE>
__device__ __forceinline__ void fused3d_tile(const Fused3dArgs& a, const int tile, float* sIn,
                                             float* sX) {
    using C = Cfg<RXY, RZ>;
    constexpr int CPT = C::CPT, P = C::P;
    constexpr bool ADJOINT = MODE == FUSED_ADJ;
    const int tx_i = tile % a.tiles_x;
    const int ty_i = (tile / a.tiles_x) % a.tiles_y;
    const int tz_i = tile / (a.tiles_x * a.tiles_y);
    const int x0 = tx_i * C::TX, y0 = ty_i * C::TY, z0 = tz_i * a.tz;

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    const int nsx = a.n_src[0], nsy = a.n_src[1], nsz = a.n_src[2];
    const int ndx = a.n_dst[0], ndy = a.n_dst[1], ndz = a.n_dst[2];
    auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };

    // ---- loads of one source plane (rows wave, wave+NW, ...; two columns per
    // lane, the second only meaningful for the 2r halo lanes).  The loads are
    // unconditional, at clamped addresses, so every step issues the same
    // number and the wait before a register set is used covers that set only;
    // zeroing (adjoint) and the lanes without a second column are applied
    // when the set is written to LDS.
    const int px0 = x0 + a.shift[0] - RXY + lane;
    const int px1 = px0 + 64;
    const bool has1 = 64 + lane < C::WX;
    const bool ok0 = !ADJOINT || (px0 >= 0 && px0 < nsx);
    const bool ok1 = !ADJOINT || (px1 >= 0 && px1 < nsx);
    const int64_t sxs = a.src_stride[0], sys = a.src_stride[1], szs = a.src_stride[2];
    const int64_t xo0 = clampi(px0, 0, nsx - 1) * sxs;
    const int64_t xo1 = has1 ? clampi(px1, 0, nsx - 1) * sxs : xo0;
    const int zv0 = z0 + a.shift[2] - RZ;  // source plane of step u = 0

    // Two register sets of one plane each.  Source plane of step u: forward
    // clamps it into the source, adjoint reads zeros outside.
    float pa0[C::LROWS], pa1[C::LROWS], pb0[C::LROWS], pb1[C::LROWS];
    auto load_plane = [&](float* r0p, float* r1p, int u) {
        const int zr = clampi(zv0 + u, 0, nsz - 1);
        const float* plane = a.src + (int64_t)zr * szs;
#pragma unroll
        for (int j = 0; j < C::LROWS; ++j) {
            const int row = min(wave + j * C::NW, C::HY - 1);
            const int py = clampi(y0 + a.shift[1] - RXY + row, 0, nsy - 1);
            const float* rp = plane + (int64_t)py * sys;
            r0p[j] = rp[xo0];
            r1p[j] = rp[xo1];
        }
    };
    auto store_plane = [&](const float* r0p, const float* r1p, float* dstp, int u) {
        const int zr = zv0 + u;
        const bool zok = !ADJOINT || (zr >= 0 && zr < nsz);
#pragma unroll
        for (int j = 0; j < C::LROWS; ++j) {
            const int row = wave + j * C::NW;
            if (row < C::HY) {
                const int py = y0 + a.shift[1] - RXY + row;
                const bool rok = zok && (!ADJOINT || (py >= 0 && py < nsy));
                dstp[row * C::SIN + lane] = (rok && ok0) ? r0p[j] : 0.0f;
                if (has1) dstp[row * C::SIN + 64 + lane] = (rok && ok1) ? r1p[j] : 0.0f;
            }
        }
    };

    const float* wx = a.wx;
    const float* wy = a.wy;

    // ---- x pass: row groups g = wave, wave+NW, ...; lane -> (row in group, quad)
    const int xq = lane & 15, xr = lane >> 4;
    auto x_pass = [&](const float* srcp) {
#pragma unroll
        for (int g0 = 0; g0 < C::RG; g0 += C::NW) {
            const int g = g0 + wave;
            if (g < C::RG) {
                const int row = 4 * g + xr;
                const float4* src4 = reinterpret_cast<const float4*>(&srcp[row * C::SIN + 4 * xq]);
                float v[4 * C::XQ];
#pragma unroll
                for (int i = 0; i < C::XQ; ++i) {
                    const float4 t = src4[i];
                    v[4 * i + 0] = t.x;
                    v[4 * i + 1] = t.y;
                    v[4 * i + 2] = t.z;
                    v[4 * i + 3] = t.w;
                }
                float o[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) o[j] = tap_sum<RXY, MODE>(v + j, wx);
                if constexpr (ADJOINT) {
                    const int xg = x0 + 4 * xq;  // destination x of o[0]
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (xg + j == 0) {
#pragma unroll
                            for (int m = 1; m <= RXY; ++m)
                                o[j] = __builtin_fmaf(a.px[m], v[j + 2 * RXY - m], o[j]);
                        }
                        if (xg + j == ndx - 1) {
#pragma unroll
                            for (int m = 1; m <= RXY; ++m)
                                o[j] = __builtin_fmaf(a.px[m], v[j + m], o[j]);
                        }
                    }
                }
                *reinterpret_cast<float4*>(&sX[row * C::SX + 4 * xq]) =
                    make_float4(o[0], o[1], o[2], o[3]);
            }
        }
    };

    // ---- y pass: CPT outputs down column `lane`, rows wave*CPT ...
    const int ybase = y0 + wave * CPT;
    float val[CPT];
    auto y_pass = [&]() {
        float v[CPT + 2 * RXY];
#pragma unroll
        for (int i = 0; i < CPT + 2 * RXY; ++i) v[i] = sX[(wave * CPT + i) * C::SX + lane];
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            val[c] = tap_sum<RXY, MODE>(v + c, wy);
            if constexpr (ADJOINT) {
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
    };

    // ---- output column of this lane
    const int ox = x0 + lane;
    const bool x_ok = ox < ndx;
    const int64_t dxs = a.dst_stride[0], dys = a.dst_stride[1], dzs = a.dst_stride[2];
    float* dcol = a.dst + (int64_t)ox * dxs + (int64_t)ybase * dys;

    // ---- z ring: fixed slots, rotating weights (wz2 = wz repeated twice)
    static_assert(CPT <= 8, "at most 8 ring columns per lane");
    float r0[P], r1[P], r2[P], r3[P], r4[P], r5[P], r6[P], r7[P];
#pragma unroll
    for (int p = 0; p < P; ++p)
        r0[p] = r1[p] = r2[p] = r3[p] = r4[p] = r5[p] = r6[p] = r7[p] = 0.0f;
    float tlo[CPT], thi[CPT];  // adjoint: z-edge sums of the first / last plane
#pragma unroll
    for (int c = 0; c < CPT; ++c) tlo[c] = thi[c] = 0.0f;

    const int tz_act = min(a.tz, ndz - z0);
    const int U = tz_act + 2 * RZ;
    const int t_hi = ndz - 1 - z0;  // step-relative index of the last destination plane

    // One step: filter plane u from LDS half `cur`; store plane u+1 (in the
    // register set `r`) into the other half and reload `r` with plane u+3.
    int s = 0;
    auto step = [&](int u, float* ra, float* rb, float* cur, float* nxt) {
        const int t = u - 2 * RZ;  // output completed at this step (if 0 <= t < tz_act)
        x_pass(cur);
        lds_barrier();  // x-filtered tile complete, previous step's readers of nxt done
        store_plane(ra, rb, nxt, u + 1);
        load_plane(ra, rb, u + 3);
        y_pass();
        const int sn = (s + 1 == P) ? 0 : s + 1;
        const float* w = a.wz2 + s + P;  // w[-j] = wz[(s - j) mod P]
        float o[CPT];
        static_for<CPT>([&](auto ci) {
            constexpr int c = decltype(ci)::value;
            float* acc = ring_of<c>(r0, r1, r2, r3, r4, r5, r6, r7);
            acc[s] = 0.0f;  // the output whose first tap is this plane
#pragma unroll
            for (int j = 0; j < P; ++j) {
                if constexpr (MODE == FUSED_FWD_EXACT) {
                    const float prod = val[c] * w[-j];
                    acc[j] = acc[j] + prod;
                } else {
                    acc[j] = __builtin_fmaf(val[c], w[-j], acc[j]);
                }
            }
            o[c] = acc[sn];  // tap 2r has just been added
            if constexpr (ADJOINT) {
                const float plo = (z0 == 0 && u >= RZ && u < 2 * RZ) ? a.pz[2 * RZ - u] : 0.0f;
                const float phi = (u > t_hi && u <= t_hi + RZ) ? a.pz[u - t_hi] : 0.0f;
                tlo[c] = __builtin_fmaf(val[c], plo, tlo[c]);
                thi[c] = __builtin_fmaf(val[c], phi, thi[c]);
                if (z0 + t == 0) o[c] = o[c] + tlo[c];
                if (t == t_hi) o[c] = o[c] + thi[c];
            }
        });
        s = sn;
        if (t >= 0 && t < tz_act) {
            float* dst = dcol + (int64_t)(z0 + t) * dzs;
#pragma unroll
            for (int c = 0; c < CPT; ++c) {
                if (x_ok && ybase + c < ndy) {
                    float q = o[c];
                    if constexpr (ADJOINT) {
                        if (a.accumulate) q = dst[c * dys] + q;
                    }
                    dst[c * dys] = q;
                }
            }
        }
        lds_barrier();  // next plane stored, x-filtered tile consumed
    };

    lds_barrier();  // the previous tile of a persistent workgroup is done with LDS
    load_plane(pa0, pa1, 0);
    store_plane(pa0, pa1, sIn, 0);
    load_plane(pa0, pa1, 1);
    load_plane(pb0, pb1, 2);
    lds_barrier();
    // An even number of steps: the extra last step (if any) stores nothing.
    for (int u = 0; u < U; u += 2) {
        step(u, pa0, pa1, sIn, sIn + C::PLANE);
        step(u + 1, pb0, pb1, sIn + C::PLANE, sIn);
    }
}

template <int RXY, int RZ, int MODE>
__global__ __launch_bounds__((Cfg<RXY, RZ>::NTH)) __attribute__((amdgpu_waves_per_eu(ITKG_WAVES_PER_EU)))
void fused3d_kernel(const Fused3dArgs a) {
    using C = Cfg<RXY, RZ>;
    __shared__ __attribute__((aligned(16))) float sIn[2 * C::PLANE];
    __shared__ __attribute__((aligned(16))) float sX[C::ROWS * C::SX];
    // XCD-aware tile order: hardware deals workgroups to the 8 XCDs round
    // robin, so consecutive tiles (which share halos) go to one XCD's L2.
    const int nt = a.tiles_x * a.tiles_y * a.tiles_z;
    const int per = (nt + 7) / 8;
    const int tile = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (tile >= nt) return;
    fused3d_tile<RXY, RZ, MODE>(a, tile, sIn, sX);
}

// Persistent form over a packed work list of (volume, tile) items for many
// volumes of one radius: workgroup b takes items b, b + G, b + 2G, ...
// The host interleaves the list so that items b and b + G share an XCD.
template <int RXY, int RZ, int MODE>
__global__ __launch_bounds__((Cfg<RXY, RZ>::NTH)) __attribute__((amdgpu_waves_per_eu(ITKG_WAVES_PER_EU)))
void fused3d_batch_kernel(const Fused3dArgs* __restrict__ vols, const int2* __restrict__ work,
                          const int nwork) {
    using C = Cfg<RXY, RZ>;
    __shared__ __attribute__((aligned(16))) float sIn[2 * C::PLANE];
    __shared__ __attribute__((aligned(16))) float sX[C::ROWS * C::SX];
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work[w];
        if (it.x < 0) continue;  // padding of the interleaved list
        fused3d_tile<RXY, RZ, MODE>(vols[it.x], it.y, sIn, sX);
    }
}

template <int RXY, int RZ>
hipError_t launch_r(const Fused3dArgs& a, int mode, hipStream_t st) {
    const int nt = a.tiles_x * a.tiles_y * a.tiles_z;
    const dim3 grid(((nt + 7) / 8) * 8), block(Cfg<RXY, RZ>::NTH);
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
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int RXY, int RZ>
hipError_t launch_batch_r(const Fused3dArgs* vols, const int2* work, int nwork, int grid_blocks,
                          int mode, hipStream_t st) {
    const dim3 grid(grid_blocks), block(Cfg<RXY, RZ>::NTH);
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
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int RXY, int RZ>
int batch_occupancy_r(int mode) {
    int blocks = 0;
    hipError_t e = hipErrorInvalidValue;
    const int nth = Cfg<RXY, RZ>::NTH;
    if (mode == FUSED_FWD_EXACT)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &blocks, fused3d_batch_kernel<RXY, RZ, FUSED_FWD_EXACT>, nth, 0);
    else if (mode == FUSED_FWD_FAST)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &blocks, fused3d_batch_kernel<RXY, RZ, FUSED_FWD_FAST>, nth, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &blocks, fused3d_batch_kernel<RXY, RZ, FUSED_ADJ>, nth, 0);
    return e == hipSuccess ? blocks : 1;
}

// Registry entry of one (rxy, rz) instantiation.
template <int RXY, int RZ>
constexpr FusedEntry make_entry() {
    return FusedEntry{RXY, RZ, Cfg<RXY, RZ>::TY, &launch_r<RXY, RZ>, &launch_batch_r<RXY, RZ>,
                      &batch_occupancy_r<RXY, RZ>};
}
template <int RXY, int RZ>
constexpr FusedEntry make_entry_single() {
    return FusedEntry{RXY, RZ, Cfg<RXY, RZ>::TY, &launch_r<RXY, RZ>, nullptr, nullptr};
}

}  // namespace

}  // namespace itkg
