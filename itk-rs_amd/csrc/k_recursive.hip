// Recursive (IIR) Gaussian of src/filters/recursive_gaussian.rs: the whole
// image widened to f64; per filtered axis, over every line along it, a
// causal then an anticausal third-order pass
//   y[n] = B x[n] + a1 y[n-1] + a2 y[n-2] + a3 y[n-3]
// (recursive_gaussian.rs:55-86), each started from the line's first
// (causal) or last (anticausal) value; then the requested region narrowed to
// the pixel type.  One thread runs one line; consecutive threads take
// neighbouring lines, which are adjacent in memory (coalesced) for every
// axis but axis 0.  Compiled with -ffp-contract=off: each product and sum
// rounds on its own, in the reference's order.
#include "itkg_impl.hpp"

namespace itkg {

namespace {

// f64 copy of the whole (largest) region; `n` sizes, axis 0 fastest.
template <typename T>
__global__ __launch_bounds__(256) void widen_kernel(const RecursiveArgs<T> a) {
    const int64_t total = a.n[0] * a.n[1] * a.n[2] * a.n[3];
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * 256) {
        int64_t r = i;
        const int64_t x = r % a.n[0];
        r /= a.n[0];
        const int64_t y = r % a.n[1];
        r /= a.n[1];
        const int64_t z = r % a.n[2];
        const int64_t w = r / a.n[2];
        a.work[i] = (double)a.in[x * a.in_stride[0] + y * a.in_stride[1] + z * a.in_stride[2] +
                                 w * a.in_stride[3]];
    }
}

// One line per thread along `axis`: lines enumerate the other axes, the
// lowest of them fastest.
__global__ __launch_bounds__(256) void line_kernel(double* work, const int64_t n0, const int64_t n1,
                                                   const int64_t n2, const int64_t n3, const int axis,
                                                   const double B, const double a1, const double a2,
                                                   const double a3) {
    const int64_t n[4] = {n0, n1, n2, n3};
    int64_t pitch[4];
    pitch[0] = 1;
    for (int d = 1; d < 4; ++d) pitch[d] = pitch[d - 1] * n[d - 1];
    const int64_t len = n[axis], stride = pitch[axis];
    const int64_t lines = n0 * n1 * n2 * n3 / len;
    const int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (l >= lines) return;
    int64_t rem = l, base = 0;
    for (int d = 0; d < 4; ++d) {
        if (d == axis) continue;
        base += (rem % n[d]) * pitch[d];
        rem /= n[d];
    }
    double* p = work + base;
    double init = p[0];
    double ym1 = init, ym2 = init, ym3 = init;
    for (int64_t i = 0; i < len; ++i) {
        const double v = B * p[i * stride] + a1 * ym1 + a2 * ym2 + a3 * ym3;
        ym3 = ym2;
        ym2 = ym1;
        ym1 = v;
        p[i * stride] = v;
    }
    init = ym1;  // the line's last value after the causal pass
    ym1 = ym2 = ym3 = init;
    for (int64_t i = len - 1; i >= 0; --i) {
        const double v = B * p[i * stride] + a1 * ym1 + a2 * ym2 + a3 * ym3;
        ym3 = ym2;
        ym2 = ym1;
        ym1 = v;
        p[i * stride] = v;
    }
}

// requested region of the f64 image, narrowed to T
template <typename T>
__global__ __launch_bounds__(256) void narrow_kernel(const RecursiveArgs<T> a) {
    const int64_t total = a.m[0] * a.m[1] * a.m[2] * a.m[3];
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * 256) {
        int64_t r = i;
        const int64_t x = r % a.m[0];
        r /= a.m[0];
        const int64_t y = r % a.m[1];
        r /= a.m[1];
        const int64_t z = r % a.m[2];
        const int64_t w = r / a.m[2];
        const int64_t src = (x + a.off[0]) + a.n[0] * ((y + a.off[1]) + a.n[1] * ((z + a.off[2]) +
                                                                                a.n[2] * (w + a.off[3])));
        a.out[x * a.out_stride[0] + y * a.out_stride[1] + z * a.out_stride[2] + w * a.out_stride[3]] =
            (T)a.work[src];
    }
}

unsigned grid_for(int64_t items) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((items + 255) / 256, 1 << 20));
}

}  // namespace

template <typename T>
hipError_t launch_recursive(const RecursiveArgs<T>& a, const int* axes, int naxes,
                            const double (*coef)[4], hipStream_t s) {
    const int64_t total = a.n[0] * a.n[1] * a.n[2] * a.n[3];
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(widen_kernel<T>, dim3(grid_for(total)), dim3(256), 0, s, a);
    for (int k = 0; k < naxes; ++k) {
        const int ax = axes[k];
        if (a.n[ax] <= 1) continue;  // a line of one sample is left as it is
        const int64_t lines = total / a.n[ax];
        hipLaunchKernelGGL(line_kernel, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, s,
                           a.work, a.n[0], a.n[1], a.n[2], a.n[3], ax, coef[k][0], coef[k][1],
                           coef[k][2], coef[k][3]);
    }
    const int64_t m = a.m[0] * a.m[1] * a.m[2] * a.m[3];
    if (m > 0) hipLaunchKernelGGL(narrow_kernel<T>, dim3(grid_for(m)), dim3(256), 0, s, a);
    return hipGetLastError();
}

template hipError_t launch_recursive<float>(const RecursiveArgs<float>&, const int*, int,
                                            const double (*)[4], hipStream_t);
template hipError_t launch_recursive<double>(const RecursiveArgs<double>&, const int*, int,
                                             const double (*)[4], hipStream_t);

}  // namespace itkg
