// The steps of CannyEdgeDetectionFilter::generate_region after smoothing
// (src/filters/edges.rs:588-664), 2-D, whole image of nx x ny (axis 0
// fastest): f64 gradient by clamped central differences, magnitude
// sqrt(gx*gx + gy*gy), direction atan2(gy, gx) in degrees modulo 180 binned
// at 22.5 / 67.5 / 112.5 / 157.5; non-maximum suppression against the two
// clamped neighbours along the bin; hysteresis (strong = nms >= upper, then
// every pixel with nms >= lower 8-connected to an edge).  Compiled with
// -ffp-contract=off: the products and sums round as the reference's do.
#include <algorithm>
#include <mutex>

#include "itkg_impl.hpp"

namespace itkg {

namespace {

#include "canny_edges.inc"

// The direction bin (0: 0, 1: 45, 2: 90, 3: 135 degrees) of a gradient as
// the reference computes it with the host libm's atan2 (edges.rs:602,
// 612-620).  The device atan2 may differ from the host's in the last bit,
// which only matters next to a bin edge: within 1e-9 degrees of one, the bin
// is decided by where the true angle lies against the edge's exact
// threshold m (the midpoint below the smallest double direction that
// to_degrees and rem_euclid take onto the edge; a correctly rounded atan2
// lands on the edge side exactly when the angle is above m), from the sign
// of gy cos(m) - gx sin(m) in double-double arithmetic.  A sign too close to
// zero to trust (below 2^-100 of the gradient, never seen for double
// inputs) keeps the ordinary bin.
__device__ __forceinline__ int canny_bin(const double gx, const double gy) {
    // to_degrees multiplies by the f64 nearest 180/pi; rem_euclid(180)
    const double deg = atan2(gy, gx) * (180.0 / 3.14159265358979323846);
    double angle = fmod(deg, 180.0);
    if (angle < 0.0) angle += 180.0;
    int b = angle < 22.5 || angle >= 157.5 ? 0 : (angle < 67.5 ? 1 : (angle < 112.5 ? 2 : 3));
    const double k = rint((angle - 22.5) * (1.0 / 45.0));
    if (k >= 0.0 && k <= 3.0 && fabs(angle - (22.5 + 45.0 * k)) < 1e-9) {
        const int e = (int)k;
        const double* t = kCannyEdgeTrig[deg < 0.0 ? 1 : 0][e];
        const double p1 = gy * t[0], e1 = fma(gy, t[0], -p1);
        const double p2 = gx * t[2], e2 = fma(gx, t[2], -p2);
        const double side = (p1 - p2) + ((e1 - e2) + (gy * t[1] - gx * t[3]));
        if (fabs(side) > 0x1p-100 * (fabs(gx) + fabs(gy)))
            b = side > 0.0 ? ((e + 1) & 3) : e;  // above edge e: the next bin
    }
    return b;
}

// canny_bin, with the common case decided without atan2: on the upper
// half-plane (the gradient negated when gy < 0, which keeps its angle
// modulo 180 degrees) the angle is below 22.5 degrees or at least 157.5
// exactly when |gy| < |gx| tan(22.5), between 67.5 and 112.5 when
// |gy| > |gx| tan(67.5), and otherwise in bin 1 or 3 by the sign of gx.
// Both differences are within a few ulps of (|gx| + |gy|); where either
// lies within 1e-12 of it (the angle within about 6e-11 degrees of an
// edge, far wider than the rounding of the reference's atan2 and degree
// conversion), and for NaN, infinite or huge gradients, canny_bin decides
// (one call, not inlined).  gy = 0 is bin 0 (atan2 gives 0 or +-180).
__device__ __noinline__ int canny_bin_call(const double gx, const double gy) {
    return canny_bin(gx, gy);
}

__device__ __forceinline__ int canny_bin_fast(const double gx, const double gy) {
    const double ax = fabs(gx), ay = fabs(gy), ag = ax + ay;
    if (!(ag <= 0x1p+1000)) return canny_bin_call(gx, gy);  // NaN, infinite or huge
    if (gy == 0.0) return 0;
    const double d1 = ay - ax * 0x1.a827999fcef32p-2;  // tan(22.5 degrees)
    const double d2 = ay - ax * 0x1.3504f333f9de6p+1;  // tan(67.5 degrees)
    const double tol = 1e-12 * ag;
    if (!(fabs(d1) > tol && fabs(d2) > tol)) return canny_bin_call(gx, gy);
    if (d1 < 0.0) return 0;
    if (d2 > 0.0) return 2;
    return (gy > 0.0) == (gx > 0.0) ? 1 : 3;
}

__global__ __launch_bounds__(256) void canny_bins_kernel(const double* gx, const double* gy,
                                                         uint8_t* bins, const int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256)
        bins[k] = (uint8_t)canny_bin_fast(gx[k], gy[k]);
}

// Gradient, direction bin and non-maximum suppression of a TW x TH tile in
// one pass.  The smoothed tile and a two-pixel ring go to LDS first (one
// round of independent loads, each at the image coordinate clamped as the
// central differences clamp it); the f64 magnitudes of the tile and of its
// one-pixel ring (the suppression's neighbours) follow in LDS, so the
// magnitude never reaches memory.  The outputs are the strong-edge map
// (nms >= upper) and the weak map (nms >= lower, suppressed pixels counting
// as 0.0) that hysteresis reads.
constexpr int kGradW = 64, kGradH = 16;

template <typename T>
__global__ __launch_bounds__(256) void canny_grad_nms_kernel(const T* sm, uint8_t* edges,
                                                             uint8_t* weak, const int64_t nx,
                                                             const int64_t ny, const double lower,
                                                             const double upper) {
    constexpr int SW = kGradW + 4, SH = kGradH + 4;  // smoothed tile + 2-pixel ring
    constexpr int MW = kGradW + 2, MH = kGradH + 2;  // magnitudes: tile + 1-pixel ring
    __shared__ double s_sm[SH][SW];
    __shared__ double s_mag[MH][MW];
    const int64_t x0 = (int64_t)blockIdx.x * kGradW, y0 = (int64_t)blockIdx.y * kGradH;
    auto clampx = [&](int64_t i) { return i < 0 ? 0 : (i > nx - 1 ? nx - 1 : i); };
    auto clampy = [&](int64_t j) { return j < 0 ? 0 : (j > ny - 1 ? ny - 1 : j); };
    // cell (ty, tx) holds the pixel at the clamped coordinate of
    // (x0 - 2 + tx, y0 - 2 + ty): the gradient at an image pixel reads its
    // clamped neighbours from the cells next to its own
    constexpr int SN = SW * SH, SPER = (SN + 255) / 256;
    T v[SPER];
#pragma unroll
    for (int m = 0; m < SPER; ++m) {
        const int k = min((int)threadIdx.x + 256 * m, SN - 1);
        const int ty = k / SW, tx = k - ty * SW;
        v[m] = sm[clampx(x0 - 2 + tx) + clampy(y0 - 2 + ty) * nx];
    }
#pragma unroll
    for (int m = 0; m < SPER; ++m) {
        const int k = (int)threadIdx.x + 256 * m;
        if (k < SN) s_sm[k / SW][k % SW] = (double)v[m];
    }
    __syncthreads();
    // the gradient of tile-relative pixel (px, py), px in -1 .. TW, py in
    // -1 .. TH
    auto grad = [&](int px, int py, double& gx, double& gy) {
        gx = s_sm[py + 2][px + 3] - s_sm[py + 2][px + 1];
        gy = s_sm[py + 3][px + 2] - s_sm[py + 1][px + 2];
    };
    // the tile's magnitudes and bins (each thread keeps the bins of the
    // pixels it suppresses below) ...
    constexpr int PER = kGradW * kGradH / 256;
    int bins[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int k = threadIdx.x + 256 * m;
        const int tx = k % kGradW, ty = k / kGradW;
        double gx, gy;
        grad(tx, ty, gx, gy);
        s_mag[ty + 1][tx + 1] = sqrt(gx * gx + gy * gy);
        bins[m] = canny_bin_fast(gx, gy);
    }
    // ... and the magnitudes of its one-pixel ring: the top and bottom rows,
    // then the left and right columns between them (cells outside the image
    // hold values the suppression never reads: it clamps its neighbours
    // into the image)
    constexpr int RING = 2 * MW + 2 * kGradH;
    if (threadIdx.x < RING) {
        int r = threadIdx.x, tx, ty;
        if (r < 2 * MW) {
            ty = r < MW ? 0 : MH - 1;
            tx = r < MW ? r : r - MW;
        } else {
            r -= 2 * MW;
            ty = 1 + (r >> 1);
            tx = (r & 1) ? MW - 1 : 0;
        }
        double gx, gy;
        grad(tx - 1, ty - 1, gx, gy);
        s_mag[ty][tx] = sqrt(gx * gx + gy * gy);
    }
    __syncthreads();
    // the suppression's neighbours, clamped into the image, as tile-relative
    // coordinates: lo (-1, or 0 on the image's first tile) .. hi (TW, or the
    // image's last pixel)
    const int lox = x0 == 0 ? 0 : -1, loy = y0 == 0 ? 0 : -1;
    const int hix = (int)min<int64_t>(kGradW, nx - 1 - x0), hiy = (int)min<int64_t>(kGradH, ny - 1 - y0);
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int k = threadIdx.x + 256 * m;
        const int tx = k % kGradW, ty = k / kGradW;
        if (tx > hix || ty > hiy) continue;  // outside the image
        const int b = bins[m];
        const int di = b == 0 ? 1 : (b == 1 ? 1 : (b == 2 ? 0 : -1));
        const int dj = b == 0 ? 0 : 1;
        auto at = [&](int a, int c) {
            return s_mag[min(max(c, loy), hiy) + 1][min(max(a, lox), hix) + 1];
        };
        const double mm = s_mag[ty + 1][tx + 1];
        const double val = (mm >= at(tx + di, ty + dj) && mm >= at(tx - di, ty - dj)) ? mm : 0.0;
        const int64_t o = (x0 + tx) + (y0 + ty) * nx;
        edges[o] = val >= upper ? 1 : 0;
        weak[o] = val >= lower ? 1 : 0;
    }
}

// One hysteresis pass: each block grows the edges inside its 32 x 32 tile
// (in LDS, with a one-pixel ring read from the global edge map) until the
// tile is stable, writes them back and raises *changed if it added any.
constexpr int kTile = 32;

__global__ __launch_bounds__(256) void canny_hysteresis_kernel(const uint8_t* weakmap, uint8_t* edges,
                                                               const int64_t nx, const int64_t ny,
                                                               int* changed) {
    __shared__ uint8_t e[kTile + 2][kTile + 2];
    __shared__ uint8_t weak[kTile][kTile];
    __shared__ int grew;
    const int64_t x0 = (int64_t)blockIdx.x * kTile, y0 = (int64_t)blockIdx.y * kTile;
    // the tile's own pixels first: a tile with no weak pixel that is not an
    // edge yet (most of an image) has nothing to grow and stops here
    static_assert(kTile * kTile == 4 * 256, "four pixels a thread");
    bool cand = false;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int k = threadIdx.x + 256 * m;
        const int ty = k / kTile, tx = k % kTile;
        const int64_t gx = x0 + tx, gy = y0 + ty;
        uint8_t w = 0, ev = 0;
        if (gx < nx && gy < ny) {
            w = weakmap[gx + gy * nx];
            ev = edges[gx + gy * nx];
        }
        weak[ty][tx] = w;
        e[ty + 1][tx + 1] = ev;
        cand = cand || (w && !ev);
    }
    if (!__syncthreads_or(cand)) return;
    // the one-pixel ring: the top and bottom rows, then the left and right
    // columns between them
    constexpr int RING = 4 * kTile + 4;
    if (threadIdx.x < RING) {
        int r = threadIdx.x, tx, ty;
        if (r < 2 * (kTile + 2)) {
            ty = r < kTile + 2 ? 0 : kTile + 1;
            tx = r < kTile + 2 ? r : r - (kTile + 2);
        } else {
            r -= 2 * (kTile + 2);
            ty = 1 + (r >> 1);
            tx = (r & 1) ? kTile + 1 : 0;
        }
        const int64_t gx = x0 + tx - 1, gy = y0 + ty - 1;
        e[ty][tx] = (gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? edges[gx + gy * nx] : 0;
    }
    bool any = false;
    for (;;) {
        if (threadIdx.x == 0) grew = 0;
        __syncthreads();
        bool mine = false;
        for (int k = threadIdx.x; k < kTile * kTile; k += 256) {
            const int ty = k / kTile, tx = k % kTile;
            if (weak[ty][tx] && !e[ty + 1][tx + 1]) {
                if (e[ty][tx] | e[ty][tx + 1] | e[ty][tx + 2] | e[ty + 1][tx] | e[ty + 1][tx + 2] |
                    e[ty + 2][tx] | e[ty + 2][tx + 1] | e[ty + 2][tx + 2]) {
                    e[ty + 1][tx + 1] = 1;
                    mine = true;
                }
            }
        }
        if (mine) grew = 1;
        __syncthreads();
        if (!grew) break;
        any = true;
        __syncthreads();
    }
    if (!any) return;
    for (int k = threadIdx.x; k < kTile * kTile; k += 256) {
        const int ty = k / kTile, tx = k % kTile;
        const int64_t gx = x0 + tx, gy = y0 + ty;
        if (gx < nx && gy < ny && e[ty + 1][tx + 1]) edges[gx + gy * nx] = 1;
    }
    if (threadIdx.x == 0) *changed = 1;
}

// The request as 1 / 0: kOutPer pixels of a row per thread, 256 apart.
constexpr int kOutPer = 4;

template <typename T>
__global__ __launch_bounds__(256) void canny_out_kernel(const uint8_t* edges, T* out,
                                                        const int64_t nx, const int64_t ox,
                                                        const int64_t oy, const int64_t mx,
                                                        const int64_t s0, const int64_t s1) {
    const int64_t j = blockIdx.y;
#pragma unroll
    for (int k = 0; k < kOutPer; ++k) {
        const int64_t i = (int64_t)blockIdx.x * (256 * kOutPer) + 256 * k + threadIdx.x;
        if (i < mx) out[i * s0 + j * s1] = edges[(i + ox) + (j + oy) * nx] ? (T)1.0 : (T)0.0;
    }
}

}  // namespace

hipError_t launch_canny_bins(const double* gx, const double* gy, uint8_t* bins, int64_t n,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(canny_bins_kernel, dim3(grid), dim3(256), 0, s, gx, gy, bins, n);
    return hipGetLastError();
}

// The flag the last pass of each hysteresis batch raises, in pinned
// coherent host memory (one per device, allocated on first use and kept
// until the process ends): the host clears it before the batch and reads it
// after the stream synchronisation, with no copy or fill on the stream.
// Calls on one device take turns at it.
constexpr int kFlagDevices = 64;
struct HostFlag {
    std::mutex mu;
    int* host = nullptr;  // host address
    int* dev = nullptr;   // the same memory as the kernels address it
};
HostFlag g_host_flag[kFlagDevices];

template <typename T>
hipError_t launch_canny(const CannyArgs<T>& a, hipStream_t s) {
    const int64_t nx = a.nx, ny = a.ny;
    if (a.my > 65535 || (ny + kTile - 1) / kTile > 65535 || (ny + kGradH - 1) / kGradH > 65535)
        return hipErrorInvalidValue;
    const dim3 g((unsigned)((nx + kGradW - 1) / kGradW), (unsigned)((ny + kGradH - 1) / kGradH));
    hipLaunchKernelGGL(canny_grad_nms_kernel<T>, g, dim3(256), 0, s, a.smoothed, a.edges, a.weak, nx,
                       ny, a.lower, a.upper);
    hipError_t e = hipGetLastError();
    const dim3 gt((unsigned)((nx + kTile - 1) / kTile), (unsigned)((ny + kTile - 1) / kTile));
    // passes until one adds nothing (each pass grows every tile to its own
    // fixed point; a chain crossing k tiles needs about k passes).  Passes
    // go in batches of 1, 2, 4, ... 64, and only a batch's last pass raises
    // the flag that is read back: a pass after the fixed point changes
    // nothing, so a chain across k tiles costs about log2(k) waits on the
    // host instead of k, for at most twice the passes.  The batch's other
    // passes raise a second flag nobody reads.
    int dev = 0;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kFlagDevices) return hipErrorInvalidDevice;
    HostFlag& hf = g_host_flag[dev];
    std::lock_guard<std::mutex> lock(hf.mu);
    if (!hf.host) {
        void* p = nullptr;
        e = hipHostMalloc(&p, sizeof(int), hipHostMallocCoherent);
        if (e != hipSuccess) return e;
        void* d = nullptr;
        e = hipHostGetDevicePointer(&d, p, 0);
        if (e != hipSuccess) return e;  // p stays allocated; the next call retries the mapping
        hf.host = static_cast<int*>(p);
        hf.dev = static_cast<int*>(d);
    }
    volatile int* const flag = hf.host;
    int batch = 1;
    while (e == hipSuccess) {
        for (int i = 0; i + 1 < batch; ++i)
            hipLaunchKernelGGL(canny_hysteresis_kernel, gt, dim3(256), 0, s, a.weak, a.edges, nx,
                               ny, a.changed + 1);
        *flag = 0;  // the last batch's passes have completed (synchronised below)
        hipLaunchKernelGGL(canny_hysteresis_kernel, gt, dim3(256), 0, s, a.weak, a.edges, nx, ny,
                           hf.dev);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess || !*flag) break;
        batch = batch < 64 ? 2 * batch : 64;
    }
    if (e != hipSuccess) return e;
    if (a.mx > 0 && a.my > 0)
        hipLaunchKernelGGL(canny_out_kernel<T>,
                           dim3((unsigned)((a.mx + 256 * kOutPer - 1) / (256 * kOutPer)), (unsigned)a.my),
                           dim3(256), 0, s, a.edges, a.out, nx, a.ox, a.oy, a.mx, a.out_stride[0],
                           a.out_stride[1]);
    return hipGetLastError();
}

template hipError_t launch_canny<float>(const CannyArgs<float>&, hipStream_t);
template hipError_t launch_canny<double>(const CannyArgs<double>&, hipStream_t);

}  // namespace itkg
