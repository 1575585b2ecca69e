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

// Grids here are (x blocks, rows of axis 1, planes of axes 2 and 3):
// the block's row and plane come from its indices (no per-pixel division).
__device__ __forceinline__ void plane_of(const int64_t n2, const int64_t n3, int64_t& i2,
                                         int64_t& i3) {
    i2 = blockIdx.z;
    i3 = 0;
    if (n3 > 1) {
        i2 = (uint32_t)blockIdx.z % (uint32_t)n2;
        i3 = (uint32_t)blockIdx.z / (uint32_t)n2;
    }
}

// f64 copy of the whole (largest) region; `n` sizes, axis 0 fastest.
constexpr int kRowBlock = 1024;  // pixels of a row per block (four per thread, 256 apart)

template <typename T>
__global__ __launch_bounds__(256) void widen_kernel(const RecursiveArgs<T> a) {
    const int64_t y = blockIdx.y;
    int64_t z, w;
    plane_of(a.n[2], a.n[3], z, w);
    double* wrow = a.work + a.n[0] * (y + a.n[1] * (z + a.n[2] * w));
    const T* irow = a.in + y * a.in_stride[1] + z * a.in_stride[2] + w * a.in_stride[3];
#pragma unroll
    for (int j = 0; j < kRowBlock / 256; ++j) {
        const int64_t x = (int64_t)blockIdx.x * kRowBlock + j * 256 + threadIdx.x;
        if (x < a.n[0]) wrow[x] = (double)irow[x * a.in_stride[0]];
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
    auto step = [&](double x) {
        const double v = B * x + a1 * ym1 + a2 * ym2 + a3 * ym3;
        ym3 = ym2;
        ym2 = ym1;
        ym1 = v;
        return v;
    };
    // Rows (stride 1) that start 16-byte aligned go in chunks of eight
    // samples with 16-byte loads and stores, each lane through its own
    // contiguous run; the samples still pass through `step` one by one.
    const bool rows = stride == 1 && (reinterpret_cast<uintptr_t>(p) & 15) == 0;
    const int64_t body = rows ? len / 8 * 8 : 0;
    for (int64_t i = 0; i < body; i += 8) {
        double2* q = reinterpret_cast<double2*>(p + i);
        double2 c[4] = {q[0], q[1], q[2], q[3]};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c[j].x = step(c[j].x);
            c[j].y = step(c[j].y);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) q[j] = c[j];
    }
    for (int64_t i = body; i < len; ++i) p[i * stride] = step(p[i * stride]);
    init = ym1;  // the line's last value after the causal pass
    ym1 = ym2 = ym3 = init;
    // backwards: the samples past the last whole chunk first, then the chunks
    for (int64_t i = len - 1; i >= body; --i) p[i * stride] = step(p[i * stride]);
    for (int64_t i = body - 8; i >= 0; i -= 8) {
        double2* q = reinterpret_cast<double2*>(p + i);
        double2 c[4] = {q[0], q[1], q[2], q[3]};
#pragma unroll
        for (int j = 3; j >= 0; --j) {
            c[j].y = step(c[j].y);
            c[j].x = step(c[j].x);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) q[j] = c[j];
    }
}

// requested region of the f64 image, narrowed to T
template <typename T>
__global__ __launch_bounds__(256) void narrow_kernel(const RecursiveArgs<T> a) {
    const int64_t y = blockIdx.y;
    int64_t z, w;
    plane_of(a.m[2], a.m[3], z, w);
    const double* wrow =
        a.work + a.off[0] + a.n[0] * ((y + a.off[1]) + a.n[1] * ((z + a.off[2]) + a.n[2] * (w + a.off[3])));
    T* orow = a.out + y * a.out_stride[1] + z * a.out_stride[2] + w * a.out_stride[3];
#pragma unroll
    for (int j = 0; j < kRowBlock / 256; ++j) {
        const int64_t x = (int64_t)blockIdx.x * kRowBlock + j * 256 + threadIdx.x;
        if (x < a.m[0]) orow[x * a.out_stride[0]] = (T)wrow[x];
    }
}

}  // namespace

template <typename T>
hipError_t launch_recursive(const RecursiveArgs<T>& a, const int* axes, int naxes,
                            const double (*coef)[4], hipStream_t s) {
    const int64_t total = a.n[0] * a.n[1] * a.n[2] * a.n[3];
    if (total == 0) return hipSuccess;
    if (a.n[1] > 65535 || a.n[2] * a.n[3] > 65535 || a.m[1] > 65535 || a.m[2] * a.m[3] > 65535)
        return hipErrorInvalidValue;  // grid rows / planes beyond the launch limit
    hipLaunchKernelGGL(widen_kernel<T>,
                       dim3((unsigned)((a.n[0] + kRowBlock - 1) / kRowBlock), (unsigned)a.n[1],
                            (unsigned)(a.n[2] * a.n[3])),
                       dim3(256), 0, s, a);
    for (int k = 0; k < naxes; ++k) {
        const int ax = axes[k];
        if (a.n[ax] <= 1) continue;  // a line of one sample is left as it is
        const int64_t lines = total / a.n[ax];
        hipLaunchKernelGGL(line_kernel, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, s,
                           a.work, a.n[0], a.n[1], a.n[2], a.n[3], ax, coef[k][0], coef[k][1],
                           coef[k][2], coef[k][3]);
    }
    const int64_t m = a.m[0] * a.m[1] * a.m[2] * a.m[3];
    if (m > 0)
        hipLaunchKernelGGL(narrow_kernel<T>,
                           dim3((unsigned)((a.m[0] + kRowBlock - 1) / kRowBlock), (unsigned)a.m[1],
                                (unsigned)(a.m[2] * a.m[3])),
                           dim3(256), 0, s, a);
    return hipGetLastError();
}

template hipError_t launch_recursive<float>(const RecursiveArgs<float>&, const int*, int,
                                            const double (*)[4], hipStream_t);
template hipError_t launch_recursive<double>(const RecursiveArgs<double>&, const int*, int,
                                             const double (*)[4], hipStream_t);

}  // namespace itkg
