// The steps of CannyEdgeDetectionFilter::generate_region after smoothing
// (src/filters/edges.rs:588-664), 2-D, whole image of nx x ny (axis 0
// fastest): f64 gradient by clamped central differences, magnitude
// sqrt(gx*gx + gy*gy), direction atan2(gy, gx) in degrees modulo 180 binned
// at 22.5 / 67.5 / 112.5 / 157.5; non-maximum suppression against the two
// clamped neighbours along the bin; hysteresis (strong = nms >= upper, then
// every pixel with nms >= lower 8-connected to an edge).  Compiled with
// -ffp-contract=off: the products and sums round as the reference's do.
#include <algorithm>

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

__global__ __launch_bounds__(256) void canny_bins_kernel(const double* gx, const double* gy,
                                                         uint8_t* bins, const int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256)
        bins[k] = (uint8_t)canny_bin(gx[k], gy[k]);
}

template <typename T>
__global__ __launch_bounds__(256) void canny_grad_kernel(const T* sm, double* mag, uint8_t* bin,
                                                         const int64_t nx, const int64_t ny) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= nx) return;
    const int64_t ip = i + 1 > nx - 1 ? nx - 1 : i + 1, im = i - 1 < 0 ? 0 : i - 1;
    const int64_t jp = j + 1 > ny - 1 ? ny - 1 : j + 1, jm = j - 1 < 0 ? 0 : j - 1;
    const double gx = (double)sm[ip + j * nx] - (double)sm[im + j * nx];
    const double gy = (double)sm[i + jp * nx] - (double)sm[i + jm * nx];
    mag[i + j * nx] = sqrt(gx * gx + gy * gy);
    bin[i + j * nx] = (uint8_t)canny_bin(gx, gy);
}

__global__ __launch_bounds__(256) void canny_nms_kernel(const double* mag, const uint8_t* bin,
                                                        double* nms, uint8_t* edges, const int64_t nx,
                                                        const int64_t ny, const double upper) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= nx) return;
    const int b = bin[i + j * nx];
    const int64_t di = b == 0 ? 1 : (b == 1 ? 1 : (b == 2 ? 0 : -1));
    const int64_t dj = b == 0 ? 0 : 1;
    auto at = [&](int64_t a, int64_t c) {
        a = a < 0 ? 0 : (a > nx - 1 ? nx - 1 : a);
        c = c < 0 ? 0 : (c > ny - 1 ? ny - 1 : c);
        return mag[a + c * nx];
    };
    const double m = mag[i + j * nx];
    const double v = (m >= at(i + di, j + dj) && m >= at(i - di, j - dj)) ? m : 0.0;
    nms[i + j * nx] = v;
    edges[i + j * nx] = v >= upper ? 1 : 0;
}

// One hysteresis pass: each block grows the edges inside its 32 x 32 tile
// (in LDS, with a one-pixel ring read from the global edge map) until the
// tile is stable, writes them back and raises *changed if it added any.
constexpr int kTile = 32;

__global__ __launch_bounds__(256) void canny_hysteresis_kernel(const double* nms, uint8_t* edges,
                                                               const int64_t nx, const int64_t ny,
                                                               const double lower, int* changed) {
    __shared__ uint8_t e[kTile + 2][kTile + 2];
    __shared__ uint8_t weak[kTile][kTile];
    __shared__ int grew;
    const int64_t x0 = (int64_t)blockIdx.x * kTile, y0 = (int64_t)blockIdx.y * kTile;
    for (int k = threadIdx.x; k < (kTile + 2) * (kTile + 2); k += 256) {
        const int ty = k / (kTile + 2), tx = k % (kTile + 2);
        const int64_t gx = x0 + tx - 1, gy = y0 + ty - 1;
        e[ty][tx] = (gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? edges[gx + gy * nx] : 0;
    }
    for (int k = threadIdx.x; k < kTile * kTile; k += 256) {
        const int ty = k / kTile, tx = k % kTile;
        const int64_t gx = x0 + tx, gy = y0 + ty;
        weak[ty][tx] = (gx < nx && gy < ny) ? (nms[gx + gy * nx] >= lower ? 1 : 0) : 0;
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

template <typename T>
__global__ __launch_bounds__(256) void canny_out_kernel(const uint8_t* edges, T* out,
                                                        const int64_t nx, const int64_t ox,
                                                        const int64_t oy, const int64_t mx,
                                                        const int64_t s0, const int64_t s1) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= mx) return;
    out[i * s0 + j * s1] = edges[(i + ox) + (j + oy) * nx] ? (T)1.0 : (T)0.0;
}

}  // namespace

hipError_t launch_canny_bins(const double* gx, const double* gy, uint8_t* bins, int64_t n,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(canny_bins_kernel, dim3(grid), dim3(256), 0, s, gx, gy, bins, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_canny(const CannyArgs<T>& a, hipStream_t s) {
    const int64_t nx = a.nx, ny = a.ny;
    if (ny > 65535 || a.my > 65535 || (ny + kTile - 1) / kTile > 65535) return hipErrorInvalidValue;
    const dim3 g((unsigned)((nx + 255) / 256), (unsigned)ny);
    hipLaunchKernelGGL(canny_grad_kernel<T>, g, dim3(256), 0, s, a.smoothed, a.mag, a.bin, nx, ny);
    hipLaunchKernelGGL(canny_nms_kernel, g, dim3(256), 0, s, a.mag, a.bin, a.nms, a.edges, nx, ny,
                       a.upper);
    hipError_t e = hipGetLastError();
    const dim3 gt((unsigned)((nx + kTile - 1) / kTile), (unsigned)((ny + kTile - 1) / kTile));
    // passes until one adds nothing (each pass grows every tile to its own
    // fixed point; a chain crossing k tiles needs about k passes).  Passes
    // go in batches of 1, 2, 4, ... 64, and only a batch's last pass raises
    // the flag that is read back: a pass after the fixed point changes
    // nothing, so a chain across k tiles costs about log2(k) waits on the
    // host instead of k, for at most twice the passes.  The batch's other
    // passes raise a second flag nobody reads.
    int batch = 1;
    while (e == hipSuccess) {
        for (int i = 0; i + 1 < batch; ++i)
            hipLaunchKernelGGL(canny_hysteresis_kernel, gt, dim3(256), 0, s, a.nms, a.edges, nx,
                               ny, a.lower, a.changed + 1);
        int host = 0;
        e = hipMemsetAsync(a.changed, 0, sizeof(int), s);
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(canny_hysteresis_kernel, gt, dim3(256), 0, s, a.nms, a.edges, nx, ny,
                           a.lower, a.changed);
        e = hipMemcpyAsync(&host, a.changed, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess || !host) break;
        batch = batch < 64 ? 2 * batch : 64;
    }
    if (e != hipSuccess) return e;
    if (a.mx > 0 && a.my > 0)
        hipLaunchKernelGGL(canny_out_kernel<T>, dim3((unsigned)((a.mx + 255) / 256), (unsigned)a.my),
                           dim3(256), 0, s, a.edges, a.out, nx, a.ox, a.oy, a.mx, a.out_stride[0],
                           a.out_stride[1]);
    return hipGetLastError();
}

template hipError_t launch_canny<float>(const CannyArgs<float>&, hipStream_t);
template hipError_t launch_canny<double>(const CannyArgs<double>&, hipStream_t);

}  // namespace itkg
