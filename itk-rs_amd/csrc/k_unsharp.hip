// The per-pixel step of UnsharpMaskFilter::generate_region
// (src/filters/edges.rs:434-449): with orig and blur widened to f64,
// edge = orig - blur, and the output is orig + amount * edge where
// |edge| > threshold, else orig, narrowed back to the pixel type.  This file
// is compiled with -ffp-contract=off: the multiply and the add round
// separately, as the reference computes them.  f32 3-D and 2-D images with
// a fused kernel for their radius apply the same step in the blur's own
// launch (fused3d.hpp, the unsharp modes); this kernel serves the rest
// (f64, 1-D and 4-D images, the per-axis path).
#include "itkg_impl.hpp"

namespace itkg {

namespace {

template <typename T>
__device__ __forceinline__ void unsharp_pixel(const UnsharpArgs<T>& a, int64_t x, int64_t i1,
                                              int64_t i2, int64_t i3) {
    const T* ip = a.in + x * a.in_stride[0] + i1 * a.in_stride[1] + i2 * a.in_stride[2] +
                  i3 * a.in_stride[3];
    const T* bp = a.blur + x * a.blur_stride[0] + i1 * a.blur_stride[1] + i2 * a.blur_stride[2] +
                  i3 * a.blur_stride[3];
    T* op = a.out + x * a.out_stride[0] + i1 * a.out_stride[1] + i2 * a.out_stride[2] +
            i3 * a.out_stride[3];
    const double orig = (double)*ip;
    const double blur = (double)*bp;
    const double edge = orig - blur;
    const double scaled = a.amount * edge;
    const double enhanced = fabs(edge) > a.threshold ? orig + scaled : orig;
    *op = (T)enhanced;
}

// Grid (x blocks of kUnsharpX pixels, rows of axis 1, planes of axes 2 and
// 3): the block's row and plane come from its indices, with no division
// when the image has at most three axes.  Each thread takes kUnsharpX / 256
// pixels, 256 apart (coalesced).
constexpr int kUnsharpX = 1024;

template <typename T>
__global__ __launch_bounds__(256) void unsharp_kernel(const UnsharpArgs<T> a) {
    const int64_t x0 = (int64_t)blockIdx.x * kUnsharpX + threadIdx.x;
    const int64_t i1 = blockIdx.y;
    int64_t i2 = blockIdx.z, i3 = 0;
    if (a.n[3] > 1) {
        const uint32_t n2 = (uint32_t)a.n[2];
        i2 = (uint32_t)blockIdx.z % n2;
        i3 = (uint32_t)blockIdx.z / n2;
    }
#pragma unroll
    for (int j = 0; j < kUnsharpX / 256; ++j) {
        const int64_t x = x0 + 256 * j;
        if (x < a.n[0]) unsharp_pixel(a, x, i1, i2, i3);
    }
}

// Any extents: blocks walk (row, x block) pairs, one division chain per pair.
template <typename T>
__global__ __launch_bounds__(256) void unsharp_kernel_rows(const UnsharpArgs<T> a) {
    const int64_t bx = (a.n[0] + 255) / 256;
    const int64_t rows = a.n[1] * a.n[2] * a.n[3];
    for (int64_t b = blockIdx.x; b < rows * bx; b += gridDim.x) {
        const int64_t row = b / bx;
        const int64_t x = (b - row * bx) * 256 + threadIdx.x;
        if (x >= a.n[0]) continue;
        const int64_t i1 = row % a.n[1];
        const int64_t r2 = row / a.n[1];
        unsharp_pixel(a, x, i1, r2 % a.n[2], r2 / a.n[2]);
    }
}

}  // namespace

template <typename T>
hipError_t launch_unsharp(const UnsharpArgs<T>& a, hipStream_t s) {
    const int64_t bx = (a.n[0] + kUnsharpX - 1) / kUnsharpX, planes = a.n[2] * a.n[3];
    const int64_t total = ((a.n[0] + 255) / 256) * a.n[1] * planes;
    if (total == 0) return hipSuccess;
    if (bx <= INT32_MAX && a.n[1] <= 65535 && planes <= 65535) {
        hipLaunchKernelGGL(unsharp_kernel<T>, dim3((unsigned)bx, (unsigned)a.n[1], (unsigned)planes),
                           dim3(256), 0, s, a);
    } else {
        const unsigned grid = (unsigned)std::min<int64_t>(total, 1 << 20);
        hipLaunchKernelGGL(unsharp_kernel_rows<T>, dim3(grid), dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

template hipError_t launch_unsharp<float>(const UnsharpArgs<float>&, hipStream_t);
template hipError_t launch_unsharp<double>(const UnsharpArgs<double>&, hipStream_t);

}  // namespace itkg
