// Fused one-round-trip 3-D Gaussian (f32 scalar pixels): forward and adjoint.
//
// Forward: GaussianFilter::generate_region (src/filters/gaussian.rs:59-90)
// for D = 3 in one kernel.  Every pixel of the requested region gets exactly
// the three convolve_axis passes of src/filters/conv.rs:11-41 (x, then y,
// then z), each rounded to f32, so the result is bit-identical to the
// per-axis kernel and to the reference (EXACT), or follows the same FMA chain
// as the per-axis FAST kernel.
//
// Adjoint: the transpose of that operator.  Per axis it is a zero-padded
// correlation with the (symmetric) weights plus, at the two clamp edges, the
// clamped taps folded onto the edge voxel: with P[m] = sum_{k<m} w_k,
//   g_in[0]   += sum_{m=1..r} P[m] * g_out[r - off - m]
//   g_in[n-1] += sum_{m=1..r} P[m] * g_out[n - 1 - off - r + m]
// (terms outside g_out are zero).  The transposed passes commute in exact
// arithmetic, so they run in the same x, y, z order as the forward.
//
// Data flow of one workgroup (4 or 8 waves):
//   - it owns a 64 x TY destination tile in x/y and a run of TZ destination
//     planes, and walks the source planes the run needs in z order;
//   - per source plane: the (TY+2r) x (64+2r) window is loaded with
//     coalesced row loads (forward: clamped at load time; adjoint: zero
//     outside) into LDS while the next plane's loads are already in flight;
//     the x pass reads 4+2r LDS values per 4 outputs and writes an
//     x-filtered tile to LDS; the y pass reads CPT+2r values per CPT outputs
//     down a column;
//   - the z pass is a register ring per output column advanced with a
//     compile-time slot (the plane counter modulo 2r+1 selects one of 2r+1
//     unrolled steps), so no register ever moves.
//
// EXACT ring (scatter form): when plane u arrives every in-flight output
// gets its next tap, fl(v * w_k) added in tap order, so each output still
// sums taps k = 0..2r left to right from +0; each plane's products are
// formed once per distinct weight (the weights are symmetric).
// FAST / adjoint ring (gather form): the ring holds the last 2r+1 plane
// values and an output is one FMA chain over them.
#include <array>
#include <type_traits>
#include <utility>

#include "itkg_impl.hpp"

namespace itkg {

namespace {

// LDS-only workgroup barrier: waits for this wave's LDS traffic, not for
// global loads, so the next plane's prefetch stays in flight across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Waves per workgroup and output columns per lane, by z radius: the z ring
// holds CPT * (2rz+1) registers per lane and is kept near 64; the y tile is
// TY = NW * CPT rows (32, or 16 beyond rz = 10).
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
    static constexpr int SX = TX;                 // sX row stride (floats)
    static constexpr int XQ = (4 + 2 * RXY + 3) / 4;  // float4 reads per x item
    static constexpr int LROWS = (HY + NW - 1) / NW;  // rows loaded per wave
    static constexpr int P = 2 * RZ + 1;
    static_assert(WX - 64 <= 64, "halo wider than one wave");
};

template <int S, int N, class F>
__device__ __forceinline__ void dispatch_slot(int s, F&& f) {
    if constexpr (S < N) {
        if (s == S) {
            f(std::integral_constant<int, S>{});
        } else {
            dispatch_slot<S + 1, N>(s, f);
        }
    }
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

// One destination tile (64 x TY x TZ) of one volume.  sIn and sX are the
// workgroup's LDS tiles; the caller's previous tile, if any, may still be
// reading them (the first plane starts with a barrier).
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
    auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
    // source x of this lane's two load columns
    const int px0 = x0 + a.shift[0] - RXY + lane;
    const int px1 = px0 + 64;
    const bool ok0 = !ADJOINT || (px0 >= 0 && px0 < nsx);
    const bool ok1 = !ADJOINT || (px1 >= 0 && px1 < nsx);
    const int64_t sxs = a.src_stride[0], sys = a.src_stride[1], szs = a.src_stride[2];
    const int64_t xo0 = clampi(px0, 0, nsx - 1) * sxs, xo1 = clampi(px1, 0, nsx - 1) * sxs;
    const bool has1 = 64 + lane < C::WX;

    float pre0[C::LROWS], pre1[C::LROWS];
    auto load_plane = [&](int cz) {
        const float* plane = a.src + (int64_t)cz * szs;
#pragma unroll
        for (int j = 0; j < C::LROWS; ++j) {
            const int row = wave + j * C::NW;
            if (row < C::HY) {
                const int py = y0 + a.shift[1] - RXY + row;
                if (ADJOINT && (py < 0 || py >= nsy)) {
                    pre0[j] = 0.0f;
                    pre1[j] = 0.0f;
                } else {
                    const float* rp = plane + (int64_t)clampi(py, 0, nsy - 1) * sys;
                    pre0[j] = ok0 ? rp[xo0] : 0.0f;
                    if (has1) pre1[j] = ok1 ? rp[xo1] : 0.0f;
                }
            }
        }
    };
    auto store_plane = [&]() {
#pragma unroll
        for (int j = 0; j < C::LROWS; ++j) {
            const int row = wave + j * C::NW;
            if (row < C::HY) {
                sIn[row * C::SIN + lane] = pre0[j];
                if (has1) sIn[row * C::SIN + 64 + lane] = pre1[j];
            }
        }
    };

    const float* wx = a.wx;
    const float* wy = a.wy;
    const float* wz = a.wz;
    const int ndx = a.n_dst[0], ndy = a.n_dst[1], ndz = a.n_dst[2];

    // x pass: wave handles row groups g = wave, wave+NW, ...; lane -> (row in
    // group, quad)
    const int xq = lane & 15, xr = lane >> 4;
    auto x_pass = [&]() {
#pragma unroll
        for (int g0 = 0; g0 < C::RG; g0 += C::NW) {
            const int g = g0 + wave;
            if (g < C::RG) {
                const int row = 4 * g + xr;
                const float4* src4 =
                    reinterpret_cast<const float4*>(&sIn[row * C::SIN + 4 * xq]);
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

    // output addressing
    const int ox = x0 + lane;
    const bool x_ok = ox < ndx;
    const int64_t dxs = a.dst_stride[0], dys = a.dst_stride[1], dzs = a.dst_stride[2];
    float* dcol = a.dst + (int64_t)ox * dxs + (int64_t)ybase * dys;

    float ring[CPT][P];
#pragma unroll
    for (int c = 0; c < CPT; ++c)
#pragma unroll
        for (int p = 0; p < P; ++p) ring[c][p] = 0.0f;

    const int tz_act = min(a.tz, ndz - z0);
    const int U = tz_act + 2 * RZ;
    const int zv0 = z0 + a.shift[2] - RZ;  // source plane of step u = 0
    // source planes this run reads: [p_first, p_last] (empty if p_first > p_last)
    const int p_first = ADJOINT ? max(zv0, 0) : clampi(zv0, 0, nsz - 1);
    const int p_last = ADJOINT ? min(zv0 + U - 1, nsz - 1) : clampi(zv0 + U - 1, 0, nsz - 1);
    int cz_prev = -1;

    if (p_first <= p_last) load_plane(p_first);
    int s = 0;
    for (int u = 0; u < U; ++u) {
        const int zraw = zv0 + u;
        const int cz = ADJOINT ? zraw : clampi(zraw, 0, nsz - 1);
        if (ADJOINT && (cz < 0 || cz >= nsz)) {
#pragma unroll
            for (int c = 0; c < CPT; ++c) val[c] = 0.0f;
        } else if (cz != cz_prev) {
            lds_barrier();
            store_plane();
            lds_barrier();
            if (cz + 1 <= p_last) load_plane(cz + 1);
            x_pass();
            lds_barrier();
            y_pass();
            cz_prev = cz;
        }
        const int t = u - 2 * RZ;  // output completed at this step (if >= 0)
        const int zo = z0 + t;
        float* dst = dcol + (int64_t)zo * dzs;
        dispatch_slot<0, P>(s, [&](auto sc) {
            constexpr int S = decltype(sc)::value;
            if constexpr (MODE == FUSED_FWD_EXACT) {
#pragma unroll
                for (int c = 0; c < CPT; ++c) {
                    float prod[RZ + 1];
#pragma unroll
                    for (int m = 0; m <= RZ; ++m) prod[m] = val[c] * wz[m];
                    // tap 0 opens the output whose window starts at this plane
                    ring[c][S] = 0.0f + prod[0];
#pragma unroll
                    for (int k = 1; k < P; ++k) {
                        const int m = k <= RZ ? k : 2 * RZ - k;
                        ring[c][(S - k + P) % P] = ring[c][(S - k + P) % P] + prod[m];
                    }
                    if (t >= 0 && x_ok && ybase + c < ndy) dst[c * dys] = ring[c][(S + 1) % P];
                }
            } else {
#pragma unroll
                for (int c = 0; c < CPT; ++c) ring[c][S] = val[c];
                if (t >= 0) {
#pragma unroll
                    for (int c = 0; c < CPT; ++c) {
                        float acc = ring[c][(S + 1) % P] * wz[0];
#pragma unroll
                        for (int k = 1; k < P; ++k)
                            acc = __builtin_fmaf(ring[c][(S + 1 + k) % P], wz[k], acc);
                        if constexpr (ADJOINT) {
                            if (zo == 0) {
#pragma unroll
                                for (int m = 1; m <= RZ; ++m)
                                    acc = __builtin_fmaf(a.pz[m], ring[c][(S - m + P) % P], acc);
                            }
                            if (zo == ndz - 1) {
#pragma unroll
                                for (int m = 1; m <= RZ; ++m)
                                    acc = __builtin_fmaf(a.pz[m], ring[c][(S + 1 + m) % P], acc);
                            }
                        }
                        if (x_ok && ybase + c < ndy) {
                            if constexpr (ADJOINT) {
                                if (a.accumulate) acc = dst[c * dys] + acc;
                            }
                            dst[c * dys] = acc;
                        }
                    }
                }
            }
        });
        s = (s + 1 == P) ? 0 : s + 1;
    }
}

template <int RXY, int RZ, int MODE>
__global__ __launch_bounds__((Cfg<RXY, RZ>::NTH)) void fused3d_kernel(const Fused3dArgs a) {
    using C = Cfg<RXY, RZ>;
    __shared__ __attribute__((aligned(16))) float sIn[C::ROWS * C::SIN];
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
__global__ __launch_bounds__((Cfg<RXY, RZ>::NTH)) void fused3d_batch_kernel(
    const Fused3dArgs* __restrict__ vols, const int2* __restrict__ work, const int nwork) {
    using C = Cfg<RXY, RZ>;
    __shared__ __attribute__((aligned(16))) float sIn[C::ROWS * C::SIN];
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

using LaunchFn = hipError_t (*)(const Fused3dArgs&, int, hipStream_t);
using BatchFn = hipError_t (*)(const Fused3dArgs*, const int2*, int, int, int, hipStream_t);
using OccFn = int (*)(int);

template <int... R>
constexpr auto make_iso_table(std::integer_sequence<int, R...>) {
    return std::array<LaunchFn, sizeof...(R)>{&launch_r<R + 1, R + 1>...};
}
template <int... R>
constexpr auto make_batch_table(std::integer_sequence<int, R...>) {
    return std::array<BatchFn, sizeof...(R)>{&launch_batch_r<R + 1, R + 1>...};
}

template <int... R>
constexpr auto make_occ_table(std::integer_sequence<int, R...>) {
    return std::array<OccFn, sizeof...(R)>{&batch_occupancy_r<R + 1, R + 1>...};
}

constexpr int kIsoMax = kFusedMaxR;
const auto kIsoTable = make_iso_table(std::make_integer_sequence<int, kIsoMax>{});
const auto kBatchTable = make_batch_table(std::make_integer_sequence<int, kIsoMax>{});
const auto kOccTable = make_occ_table(std::make_integer_sequence<int, kIsoMax>{});

}  // namespace

bool fused3d_supported(int rx, int ry, int rz) {
    return rx == ry && ry == rz && rx >= 1 && rx <= kIsoMax;
}

int fused3d_tile_x() { return Cfg<1, 1>::TX; }
int fused3d_tile_y(int rz) {
    return rz <= 3 ? Cfg<1, 3>::TY : (rz <= 10 ? Cfg<1, 10>::TY : Cfg<1, 11>::TY);
}

hipError_t launch_fused3d(const Fused3dArgs& a, int rx, int ry, int rz, int mode,
                          hipStream_t s) {
    if (!fused3d_supported(rx, ry, rz)) return hipErrorInvalidValue;
    return kIsoTable[rx - 1](a, mode, s);
}

hipError_t launch_fused3d_batch(const Fused3dArgs* vols, const int2* work, int nwork,
                                int grid_blocks, int r, int mode, hipStream_t s) {
    if (!fused3d_supported(r, r, r) || nwork <= 0 || grid_blocks <= 0) return hipErrorInvalidValue;
    return kBatchTable[r - 1](vols, work, nwork, grid_blocks, mode, s);
}

int fused3d_batch_occupancy(int r, int mode) {
    if (!fused3d_supported(r, r, r)) return 1;
    return kOccTable[r - 1](mode);
}

}  // namespace itkg
