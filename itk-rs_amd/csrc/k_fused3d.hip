// Fused one-round-trip 3-D Gaussian forward (f32 scalar pixels).
//
// Computes GaussianFilter::generate_region (src/filters/gaussian.rs:59-90)
// for D = 3 in one kernel: every pixel of the requested region gets exactly
// the three convolve_axis passes of src/filters/conv.rs:11-41 (x, then y,
// then z), each rounded to f32, so the result is bit-identical to the
// per-axis path and to the reference.
//
// Data flow of one workgroup (4 to 16 waves):
//   - it owns a TX x TY output tile in x/y and a run of TZ output planes;
//   - it walks the input planes the run needs in z order (planes outside the
//     image clamp to the edge plane, whose x/y result is reused);
//   - per input plane: the (TY+2ry) x (TX+2rx) input window is loaded from
//     HBM with coalesced row loads (x clamped at load time) into LDS while
//     the next plane's loads are already in flight; the x pass reads LDS
//     windows of 4+2rx values per 4 outputs and writes an x-filtered tile to
//     LDS; the y pass reads CPT+2ry values per CPT outputs down a column;
//   - the z pass is a register ring per output column, advanced with a
//     compile-time slot index (the plane counter modulo 2rz+1 selects one of
//     2rz+1 unrolled ring steps), so no register ever moves.
//
// EXACT ring (scatter form): when plane u arrives every in-flight output
// gets its next tap, fl(v * w_k) added in tap order, so each output still
// sums taps k = 0..2rz left to right from +0.  Each plane's products are
// formed once per distinct weight (the weights are symmetric).
// FAST ring (gather form): the ring holds the last 2rz+1 plane values and an
// output is one FMA chain over them.
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

template <int RXY, int RZ, bool FAST>
__global__ __launch_bounds__((Cfg<RXY, RZ>::NTH)) void fused3d_kernel(const Fused3dArgs a) {
    using C = Cfg<RXY, RZ>;
    constexpr int CPT = C::CPT, P = C::P;
    __shared__ __attribute__((aligned(16))) float sIn[C::ROWS * C::SIN];
    __shared__ __attribute__((aligned(16))) float sX[C::ROWS * C::SX];

    // XCD-aware tile order: hardware deals workgroups to the 8 XCDs round
    // robin, so consecutive tiles (which share halos) go to one XCD's L2.
    const int nt = a.tiles_x * a.tiles_y * a.tiles_z;
    const int per = (nt + 7) / 8;
    const int tile = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (tile >= nt) return;
    const int tx_i = tile % a.tiles_x;
    const int ty_i = (tile / a.tiles_x) % a.tiles_y;
    const int tz_i = tile / (a.tiles_x * a.tiles_y);
    const int x0 = tx_i * C::TX, y0 = ty_i * C::TY, z0 = tz_i * a.tz;

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    const int nix = a.n_in[0], niy = a.n_in[1], niz = a.n_in[2];
    // x offsets of this lane's two load columns (clamped to the input)
    auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
    const int gx0 = clampi(x0 + a.off[0] - RXY + lane, 0, nix - 1);
    const int gx1 = clampi(x0 + a.off[0] - RXY + 64 + lane, 0, nix - 1);
    const int64_t sxs = a.src_stride[0], sys = a.src_stride[1], szs = a.src_stride[2];
    const int64_t xo0 = gx0 * sxs, xo1 = gx1 * sxs;
    const bool has1 = 64 + lane < C::WX;

    float pre0[C::LROWS], pre1[C::LROWS];
    auto load_plane = [&](int cz) {
        const float* plane = a.src + (int64_t)cz * szs;
#pragma unroll
        for (int j = 0; j < C::LROWS; ++j) {
            const int row = wave + j * C::NW;
            if (row < C::HY) {
                const int gy = clampi(y0 + a.off[1] - RXY + row, 0, niy - 1);
                const float* rp = plane + (int64_t)gy * sys;
                pre0[j] = rp[xo0];
                if (has1) pre1[j] = rp[xo1];
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
                for (int j = 0; j < 4; ++j) {
                    if constexpr (FAST) {
                        float acc = v[j] * wx[0];
#pragma unroll
                        for (int k = 1; k <= 2 * RXY; ++k) acc = __builtin_fmaf(v[j + k], wx[k], acc);
                        o[j] = acc;
                    } else {
                        float acc = 0.0f;
#pragma unroll
                        for (int k = 0; k <= 2 * RXY; ++k) {
                            const float prod = v[j + k] * wx[k];
                            acc = acc + prod;
                        }
                        o[j] = acc;
                    }
                }
                *reinterpret_cast<float4*>(&sX[row * C::SX + 4 * xq]) =
                    make_float4(o[0], o[1], o[2], o[3]);
            }
        }
    };

    float val[CPT];
    auto y_pass = [&]() {
        float v[CPT + 2 * RXY];
#pragma unroll
        for (int i = 0; i < CPT + 2 * RXY; ++i) v[i] = sX[(wave * CPT + i) * C::SX + lane];
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            if constexpr (FAST) {
                float acc = v[c] * wy[0];
#pragma unroll
                for (int k = 1; k <= 2 * RXY; ++k) acc = __builtin_fmaf(v[c + k], wy[k], acc);
                val[c] = acc;
            } else {
                float acc = 0.0f;
#pragma unroll
                for (int k = 0; k <= 2 * RXY; ++k) {
                    const float prod = v[c + k] * wy[k];
                    acc = acc + prod;
                }
                val[c] = acc;
            }
        }
    };

    // output addressing
    const int ox = x0 + lane;
    const bool x_ok = ox < a.n_out[0];
    const int64_t dxs = a.dst_stride[0], dys = a.dst_stride[1], dzs = a.dst_stride[2];
    float* dcol = a.dst + (int64_t)ox * dxs + (int64_t)(y0 + wave * CPT) * dys;
    const int ybase = y0 + wave * CPT;

    float ring[CPT][P];
#pragma unroll
    for (int c = 0; c < CPT; ++c)
#pragma unroll
        for (int p = 0; p < P; ++p) ring[c][p] = 0.0f;

    const int tz_act = min(a.tz, a.n_out[2] - z0);
    const int U = tz_act + 2 * RZ;
    const int zv0 = z0 + a.off[2] - RZ;  // virtual input plane of u = 0
    const int cz_last = clampi(zv0 + U - 1, 0, niz - 1);
    int cz_prev = -1;

    load_plane(clampi(zv0, 0, niz - 1));
    int s = 0;
    for (int u = 0; u < U; ++u) {
        const int cz = clampi(zv0 + u, 0, niz - 1);
        if (cz != cz_prev) {
            lds_barrier();
            store_plane();
            lds_barrier();
            if (cz + 1 <= cz_last) load_plane(cz + 1);
            x_pass();
            lds_barrier();
            y_pass();
            cz_prev = cz;
        }
        const int t = u - 2 * RZ;  // output completed at this step (if >= 0)
        float* dst = dcol + (int64_t)(z0 + t) * dzs;
        dispatch_slot<0, P>(s, [&](auto sc) {
            constexpr int S = decltype(sc)::value;
            if constexpr (FAST) {
#pragma unroll
                for (int c = 0; c < CPT; ++c) ring[c][S] = val[c];
                if (t >= 0) {
#pragma unroll
                    for (int c = 0; c < CPT; ++c) {
                        float acc = ring[c][(S + 1) % P] * wz[0];
#pragma unroll
                        for (int k = 1; k < P; ++k)
                            acc = __builtin_fmaf(ring[c][(S + 1 + k) % P], wz[k], acc);
                        if (x_ok && ybase + c < a.n_out[1]) dst[c * dys] = acc;
                    }
                }
            } else {
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
                    if (t >= 0 && x_ok && ybase + c < a.n_out[1])
                        dst[c * dys] = ring[c][(S + 1) % P];
                }
            }
        });
        s = (s + 1 == P) ? 0 : s + 1;
    }
}

template <int RXY, int RZ>
hipError_t launch_r(const Fused3dArgs& a, itkg_arith arith, hipStream_t st) {
    const int nt = a.tiles_x * a.tiles_y * a.tiles_z;
    const dim3 grid(((nt + 7) / 8) * 8), block(Cfg<RXY, RZ>::NTH);
    if (arith == ITKG_FAST)
        hipLaunchKernelGGL((fused3d_kernel<RXY, RZ, true>), grid, block, 0, st, a);
    else
        hipLaunchKernelGGL((fused3d_kernel<RXY, RZ, false>), grid, block, 0, st, a);
    return hipGetLastError();
}

using LaunchFn = hipError_t (*)(const Fused3dArgs&, itkg_arith, hipStream_t);

template <int... R>
constexpr auto make_iso_table(std::integer_sequence<int, R...>) {
    return std::array<LaunchFn, sizeof...(R)>{&launch_r<R + 1, R + 1>...};
}

constexpr int kIsoMax = 16;
const auto kIsoTable = make_iso_table(std::make_integer_sequence<int, kIsoMax>{});

}  // namespace

bool fused3d_supported(int rx, int ry, int rz) {
    return rx == ry && ry == rz && rx >= 1 && rx <= kIsoMax;
}

int fused3d_tile_x() { return Cfg<1, 1>::TX; }
int fused3d_tile_y(int rz) {
    return rz <= 3 ? Cfg<1, 3>::TY : (rz <= 10 ? Cfg<1, 10>::TY : Cfg<1, 11>::TY);
}

hipError_t launch_fused3d(const Fused3dArgs& a, int rx, int ry, int rz, itkg_arith arith,
                          hipStream_t s) {
    if (!fused3d_supported(rx, ry, rz)) return hipErrorInvalidValue;
    return kIsoTable[rx - 1](a, arith, s);
}

}  // namespace itkg
