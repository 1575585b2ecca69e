// Recursive (IIR) Gaussian of src/filters/recursive_gaussian.rs: the whole
// image widened to f64; per filtered axis, over every line along it, a
// causal then an anticausal third-order pass
//   y[n] = B x[n] + a1 y[n-1] + a2 y[n-2] + a3 y[n-3]
// (recursive_gaussian.rs:55-86), each started from the line's first
// (causal) or last (anticausal) value; then the requested region narrowed to
// the pixel type.  The widening happens as the first pass reads the input
// and the narrowing as the last pass writes the output.  One thread runs one
// line (two neighbouring ones on the in-place passes of the strided axes);
// consecutive threads take neighbouring lines, which are adjacent in memory
// (coalesced) for every axis but axis 0, whose rows a wave moves through
// LDS.  The causal values are not stored: the anticausal pass
// recomputes them chunk by chunk from checkpointed states.  Compiled with
// -ffp-contract=off: each product and sum rounds on its own, in the
// reference's order.
#include <type_traits>

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

// Where a line pass reads its samples (bit 0: the input image, widened to
// f64, instead of the f64 work image) and writes its results (bit 1: the
// requested part of each line, narrowed to T, into the output instead of
// the work image).  The first filtered axis reads the input and the last
// writes the output, so no pass of its own widens or narrows.
constexpr int kFromIn = 1, kToOut = 2;

// A line goes in chunks of K samples (the last one possibly shorter):
// kRecRowChunk on rows (axis 0; 16-byte loads and stores, each lane through
// its own contiguous row) and kRecLineChunk on the other axes (one load per
// sample, coalesced across the lanes' neighbouring lines).
static_assert(kRecRowChunk % 4 == 0 && kRecLineChunk % 4 == 0,
              "chunks are whole float4 / double2 groups");

template <typename T, int K>
__device__ __forceinline__ void load_vec(const T* q, double* c) {
    if constexpr (std::is_same_v<T, float>) {
#pragma unroll
        for (int j = 0; j < K / 4; ++j) {
            const float4 u = reinterpret_cast<const float4*>(q)[j];
            c[4 * j] = u.x, c[4 * j + 1] = u.y, c[4 * j + 2] = u.z, c[4 * j + 3] = u.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < K / 2; ++j) {
            const double2 t = reinterpret_cast<const double2*>(q)[j];
            c[2 * j] = t.x, c[2 * j + 1] = t.y;
        }
    }
}

template <int K>
__device__ __forceinline__ void store_vec(double* q, const double* c) {
#pragma unroll
    for (int j = 0; j < K / 2; ++j)
        reinterpret_cast<double2*>(q)[j] = double2{c[2 * j], c[2 * j + 1]};
}

// Rows (axis 0, the first filtered axis: the input is read) with the
// chunks of a wave's 64 rows moved through LDS: the wave loads each chunk
// of its rows cooperatively, 16 bytes per lane (an instruction covers a few
// whole row segments instead of 64 scattered pieces), each lane then reads
// its own row's chunk from a padded tile, and the results for the f64 work
// image go back the same way.  The recursion, the checkpoints and the
// order of every operation are those of line_kernel.  Returns false (and
// does nothing) where a wave's rows are not 16-byte aligned; every wave of
// the launch decides alike (the test depends on the strides and the base
// alone), so a launch runs one path or the other.
template <typename T, int IO>
__device__ __forceinline__ bool rows_cooperative(const RecursiveArgs<T>& a, const int64_t len,
                                                 const int64_t lines, const double B,
                                                 const double a1, const double a2,
                                                 const double a3) {
    constexpr bool to_out = (IO & kToOut) != 0;
    constexpr int K = kRecRowChunk;
    constexpr int S = K + 1;  // tile row stride (doubles)
    constexpr int VPR = std::is_same_v<T, float> ? K / 4 : K / 2;  // 16-byte vectors per row chunk
    constexpr int RPI = 64 / VPR;                                   // rows per instruction
    static_assert((K == 16 || K == 32) && 64 % VPR == 0, "16- or 32-sample row chunks");
    // every row of the input and of the work image starts 16-byte aligned
    // (the row pitch and the bases), the input's samples are contiguous
    const bool aligned = a.in_stride[0] == 1 && (reinterpret_cast<uintptr_t>(a.in) & 15) == 0 &&
                         (a.in_stride[1] * (int64_t)sizeof(T)) % 16 == 0 &&
                         (a.in_stride[2] * (int64_t)sizeof(T)) % 16 == 0 &&
                         (a.in_stride[3] * (int64_t)sizeof(T)) % 16 == 0 &&
                         (to_out || ((reinterpret_cast<uintptr_t>(a.work) & 15) == 0 && a.n[0] % 2 == 0));
    if (!aligned) return false;
    __shared__ double tiles[4][64 * S];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double* const tile = tiles[wave];
    const int64_t lw = (int64_t)blockIdx.x * 256 + wave * 64;  // the wave's first line
    if (lw >= lines) return true;
    const int nrow = (int)(lines - lw < 64 ? lines - lw : 64);
    // this lane's row: the line index's coordinates on axes 1..3 (axis 1
    // fastest), its input row and, with kToOut, its output row
    const int64_t l = lw + lane;
    int64_t rem = l < lines ? l : lines - 1, ibase = 0, obase = 0;
    bool active = l < lines;
    for (int d = 1; d < 4; ++d) {
        const int64_t c = rem % a.n[d];
        rem /= a.n[d];
        ibase += c * a.in_stride[d];
        if constexpr (to_out) {
            const int64_t o = c - a.off[d];
            active = active && o >= 0 && o < a.m[d];
            obase += (o < 0 ? 0 : o) * a.out_stride[d];
        }
    }
    const T* q = a.in + ibase;
    T* o = a.out + obase;
    double* p = a.work + (l < lines ? l : lines - 1) * a.n[0];
    const int64_t os = a.out_stride[0];
    const int64_t lo = a.off[0], hi = lo + a.m[0];
    auto step = [&](double x, double& m1, double& m2, double& m3) {
        const double v = B * x + a1 * m1 + a2 * m2 + a3 * m3;
        m3 = m2;
        m2 = m1;
        m1 = v;
        return v;
    };
    // another lane's row start (a 64-bit pointer moved as two 32-bit halves)
    auto row_ptr = [](auto* ptr, int src_lane) {
        const uint64_t u = reinterpret_cast<uint64_t>(ptr);
        const uint32_t lo32 = __shfl((uint32_t)u, src_lane), hi32 = __shfl((uint32_t)(u >> 32), src_lane);
        return reinterpret_cast<decltype(ptr)>(((uint64_t)hi32 << 32) | lo32);
    };
    auto sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    // cooperative load of chunk c: vector v of row r = RPI i + lane / VPR,
    // part lane % VPR; rows past the wave's lines and samples past the line
    // end are not loaded (the short last chunk loads whole vectors while
    // they lie inside the row)
    const int vpart = lane % VPR, vrow = lane / VPR;
    const int64_t nch = (len + K - 1) / K;
    using V = std::conditional_t<std::is_same_v<T, float>, float4, double2>;
    constexpr int EPV = std::is_same_v<T, float> ? 4 : 2;  // samples per vector
    auto vload = [&](int64_t c, V (&v)[64 / RPI]) {
#pragma unroll
        for (int i = 0; i < 64 / RPI; ++i) {
            const int r = RPI * i + vrow;
            const int64_t s0 = c * K + (int64_t)EPV * vpart;
            const T* qr = row_ptr(q, r < nrow ? r : nrow - 1);
            v[i] = V{};
            if (r < nrow && s0 + EPV <= len) v[i] = *reinterpret_cast<const V*>(qr + s0);
        }
    };
    auto vput = [&](const V (&v)[64 / RPI]) {
#pragma unroll
        for (int i = 0; i < 64 / RPI; ++i) {
            double* t = tile + (RPI * i + vrow) * S + EPV * vpart;
            if constexpr (std::is_same_v<T, float>) {
                t[0] = v[i].x, t[1] = v[i].y, t[2] = v[i].z, t[3] = v[i].w;
            } else {
                t[0] = v[i].x, t[1] = v[i].y;
            }
        }
    };
    // a short last chunk's samples that no whole vector covers, one by one
    auto tail_fix = [&](int64_t c) {
        const int64_t i0 = c * K;
        for (int64_t s0 = (len - i0) / EPV * EPV; s0 < len - i0; ++s0)
            if (l < lines) tile[lane * S + s0] = (double)q[i0 + s0];
    };
    double x[K];
    auto read_row = [&] {
#pragma unroll
        for (int j = 0; j < K; ++j) x[j] = tile[lane * S + j];
    };
    V nxt[64 / RPI];
    const double init = (double)q[0];
    // causal pass: only the checkpoints are stored
    double ym1 = init, ym2 = init, ym3 = init;
    double* const ck = a.ckpt + (l < lines ? l : lines - 1);
    vload(0, nxt);
    for (int64_t c = 0; c < nch; ++c) {
        sync();  // every lane is done reading the tile
        vput(nxt);
        if (c * K + K > len) {
            sync();
            tail_fix(c);
        }
        if (c + 1 < nch) vload(c + 1, nxt);
        sync();
        read_row();
        if (c > 0 && l < lines) {
            ck[((c - 1) * 3 + 0) * lines] = ym1;
            ck[((c - 1) * 3 + 1) * lines] = ym2;
            ck[((c - 1) * 3 + 2) * lines] = ym3;
        }
        const int64_t cnt = len - c * K;
        if (cnt >= K) {
#pragma unroll
            for (int j = 0; j < K; ++j) step(x[j], ym1, ym2, ym3);
        } else {
#pragma unroll
            for (int j = 0; j < K; ++j)
                if (j < cnt) step(x[j], ym1, ym2, ym3);
        }
    }
    // anticausal pass
    double am1 = ym1, am2 = ym1, am3 = ym1;
    double nck[3] = {init, init, init};
    if (nch > 1 && l < lines)
#pragma unroll
        for (int k = 0; k < 3; ++k) nck[k] = ck[((nch - 2) * 3 + k) * lines];
    vload(nch - 1, nxt);
    for (int64_t c = nch - 1; c >= 0; --c) {
        sync();
        vput(nxt);
        if (c * K + K > len) {
            sync();
            tail_fix(c);
        }
        double s1 = nck[0], s2 = nck[1], s3 = nck[2];  // causal state at the chunk's start
        if (c > 0) {
            vload(c - 1, nxt);
            if (c > 1 && l < lines) {
#pragma unroll
                for (int k = 0; k < 3; ++k) nck[k] = ck[((c - 2) * 3 + k) * lines];
            } else {
                nck[0] = nck[1] = nck[2] = init;
            }
        }
        sync();
        read_row();
        const int64_t i0 = c * K, cnt = len - i0;
        if (cnt >= K) {
#pragma unroll
            for (int j = 0; j < K; ++j) x[j] = step(x[j], s1, s2, s3);
#pragma unroll
            for (int j = K - 1; j >= 0; --j) x[j] = step(x[j], am1, am2, am3);
        } else {
#pragma unroll
            for (int j = 0; j < K; ++j)
                if (j < cnt) x[j] = step(x[j], s1, s2, s3);
#pragma unroll
            for (int j = K - 1; j >= 0; --j)
                if (j < cnt) x[j] = step(x[j], am1, am2, am3);
        }
        if constexpr (to_out) {
            if (active) {
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const int64_t i = i0 + j;
                    if (j < cnt && i >= lo && i < hi) o[(i - lo) * os] = (T)x[j];
                }
            }
        } else {
            // the results back through the tile, stored 16 bytes per lane
            sync();  // every lane is done reading the tile
#pragma unroll
            for (int j = 0; j < K; ++j) tile[lane * S + j] = x[j];
            sync();
            // K / 2 lanes per row (16 bytes = 2 doubles each), 128 / K rows
            // per instruction
            constexpr int LPR = K / 2, RPS = 64 / LPR;
#pragma unroll
            for (int i = 0; i < 64 / RPS; ++i) {
                const int r = RPS * i + lane / LPR, part = lane % LPR;
                const int64_t s0 = i0 + 2 * part;
                double* const pr = row_ptr(p, r < nrow ? r : 0);
                const double* t = tile + r * S + 2 * part;
                if (r < nrow && s0 + 2 <= len) {
                    *reinterpret_cast<double2*>(pr + s0) = double2{t[0], t[1]};
                } else if (r < nrow && s0 < len) {
                    pr[s0] = t[0];
                }
            }
        }
    }
    return true;
}

// One line per thread along `axis`: lines enumerate the other axes, the
// lowest of them fastest.  With kToOut, lines outside the request (in the
// other axes) are not needed and not run.
//
// The causal pass stores nothing but, at the start of every chunk after the
// first, its state (y[n-1], y[n-2], y[n-3]) in the checkpoint table
// a.ckpt[((c - 1) * 3 + k) * lines + line] (neighbouring lines adjacent).
// The anticausal pass walks the chunks backwards: it loads the chunk's
// input again, recomputes the chunk's causal values from the checkpoint
// (the same sequence of f64 operations from the same state, so the same
// values the causal pass produced), runs the anticausal recursion over them
// and stores the results.  Per sample that is two reads of the input and one
// write, instead of one read, a causal write, its read and a write.  Each
// chunk's successor (and its checkpoint) is loaded before the chunk's steps.
template <typename T, int IO, bool ROWS>
__global__ __launch_bounds__(256) void line_kernel(const RecursiveArgs<T> a, const int axis,
                                                   const double B, const double a1, const double a2,
                                                   const double a3) {
    constexpr bool from_in = (IO & kFromIn) != 0, to_out = (IO & kToOut) != 0;
    constexpr int K = ROWS ? kRecRowChunk : kRecLineChunk;
    int64_t pitch[4];
    pitch[0] = 1;
    for (int d = 1; d < 4; ++d) pitch[d] = pitch[d - 1] * a.n[d - 1];
    const int64_t len = a.n[axis], stride = pitch[axis];
    const int64_t lines = a.n[0] * a.n[1] * a.n[2] * a.n[3] / len;
    if constexpr (ROWS && from_in) {
        if (rows_cooperative<T, IO>(a, len, lines, B, a1, a2, a3)) return;
    }
    const int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (l >= lines) return;
    int64_t rem = l, base = 0, ibase = 0, obase = 0;
    for (int d = 0; d < 4; ++d) {
        if (d == axis) continue;
        const int64_t c = rem % a.n[d];
        rem /= a.n[d];
        base += c * pitch[d];
        ibase += c * a.in_stride[d];
        if constexpr (to_out) {
            const int64_t o = c - a.off[d];
            if (o < 0 || o >= a.m[d]) return;
            obase += o * a.out_stride[d];
        }
    }
    double* p = a.work + base;
    const T* q = a.in + ibase;
    T* o = a.out + obase;
    const int64_t is = a.in_stride[axis], os = a.out_stride[axis];
    const int64_t lo = a.off[axis], hi = lo + a.m[axis];
    auto src = [&](int64_t i) -> double {
        if constexpr (from_in) return (double)q[i * is];
        else return p[i * stride];
    };
    auto put = [&](int64_t i, double v) {
        if constexpr (to_out) {
            if (i >= lo && i < hi) o[(i - lo) * os] = (T)v;
        } else {
            p[i * stride] = v;
        }
    };
    auto step = [&](double x, double& m1, double& m2, double& m3) {
        const double v = B * x + a1 * m1 + a2 * m2 + a3 * m3;
        m3 = m2;
        m2 = m1;
        m1 = v;
        return v;
    };
    // rows (stride 1) that start 16-byte aligned load and store whole chunks
    // as 16-byte vectors
    bool vec = ROWS && stride == 1 && (reinterpret_cast<uintptr_t>(p) & 15) == 0;
    if constexpr (from_in) vec = vec && is == 1 && (reinterpret_cast<uintptr_t>(q) & 15) == 0;
    const int64_t nch = (len + K - 1) / K;
    double* const ck = a.ckpt + l;
    auto load = [&](int64_t c, double (&x)[K]) {
        const int64_t i0 = c * K;
        if (ROWS && vec && i0 + K <= len) {
            if constexpr (from_in) load_vec<T, K>(q + i0, x);
            else load_vec<double, K>(p + i0, x);
            return;
        }
        // the last chunk may be short: its missing samples read the last
        // sample again (in bounds) and are not used
#pragma unroll
        for (int j = 0; j < K; ++j) x[j] = src(i0 + j < len ? i0 + j : len - 1);
    };
    const double init = src(0);
    // causal pass
    double ym1 = init, ym2 = init, ym3 = init;
    double nxt[K];
    load(0, nxt);
    for (int64_t c = 0; c < nch; ++c) {
        double x[K];
#pragma unroll
        for (int j = 0; j < K; ++j) x[j] = nxt[j];
        if (c + 1 < nch) load(c + 1, nxt);
        if (c > 0) {
            ck[((c - 1) * 3 + 0) * lines] = ym1;
            ck[((c - 1) * 3 + 1) * lines] = ym2;
            ck[((c - 1) * 3 + 2) * lines] = ym3;
        }
        const int64_t cnt = len - c * K;
        if (cnt >= K) {
#pragma unroll
            for (int j = 0; j < K; ++j) step(x[j], ym1, ym2, ym3);
        } else {
#pragma unroll
            for (int j = 0; j < K; ++j)
                if (j < cnt) step(x[j], ym1, ym2, ym3);
        }
    }
    // anticausal pass, from the line's last causal value
    double am1 = ym1, am2 = ym1, am3 = ym1;
    double nck[3] = {init, init, init};
    load(nch - 1, nxt);
    if (nch > 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) nck[k] = ck[((nch - 2) * 3 + k) * lines];
    for (int64_t c = nch - 1; c >= 0; --c) {
        double x[K];
#pragma unroll
        for (int j = 0; j < K; ++j) x[j] = nxt[j];
        double s1 = nck[0], s2 = nck[1], s3 = nck[2];  // causal state at the chunk's start
        if (c > 0) {
            load(c - 1, nxt);
            if (c > 1) {
#pragma unroll
                for (int k = 0; k < 3; ++k) nck[k] = ck[((c - 2) * 3 + k) * lines];
            } else {
                nck[0] = nck[1] = nck[2] = init;
            }
        }
        const int64_t i0 = c * K, cnt = len - i0;
        if (cnt >= K) {
#pragma unroll
            for (int j = 0; j < K; ++j) x[j] = step(x[j], s1, s2, s3);
#pragma unroll
            for (int j = K - 1; j >= 0; --j) x[j] = step(x[j], am1, am2, am3);
            if (!to_out && ROWS && vec) {
                store_vec<K>(p + i0, x);
            } else {
#pragma unroll
                for (int j = 0; j < K; ++j) put(i0 + j, x[j]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < K; ++j)
                if (j < cnt) x[j] = step(x[j], s1, s2, s3);
#pragma unroll
            for (int j = K - 1; j >= 0; --j)
                if (j < cnt) x[j] = step(x[j], am1, am2, am3);
#pragma unroll
            for (int j = 0; j < K; ++j)
                if (j < cnt) put(i0 + j, x[j]);
        }
    }
}

// The strided axes between the first and the last filtered one (the f64
// work image in place) with n0 even: two neighbouring lines per thread,
// adjacent in memory and 16-byte aligned, so every load and store of the
// work image and of the checkpoints moves 16 bytes a lane (an instruction
// covers 1 KB of a row instead of 512 B) and each thread runs two
// independent recursions.  Each line's operations, in their order, are
// those of line_kernel.  512^3, axis 1: 647 us against 685 with one line a
// thread; the last pass (f32 output) measured slower this way (589 against
// 564 us) and keeps line_kernel.
template <typename T>
__global__ __launch_bounds__(256) void line_kernel2(const RecursiveArgs<T> a, const int axis,
                                                    const double B, const double a1, const double a2,
                                                    const double a3) {
    constexpr int K = kRecLineChunk;
    int64_t pitch[4];
    pitch[0] = 1;
    for (int d = 1; d < 4; ++d) pitch[d] = pitch[d - 1] * a.n[d - 1];
    const int64_t len = a.n[axis], stride = pitch[axis];
    const int64_t lines = a.n[0] * a.n[1] * a.n[2] * a.n[3] / len;
    const int64_t l = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 2;  // lines l, l + 1
    if (l >= lines) return;
    int64_t rem = l, base = 0;
    for (int d = 0; d < 4; ++d) {
        if (d == axis) continue;
        const int64_t c = rem % a.n[d];
        rem /= a.n[d];
        base += c * pitch[d];
    }
    double* p = a.work + base;
    auto ld = [&](int64_t i) { return *reinterpret_cast<const double2*>(p + i * stride); };
    auto st = [&](int64_t i, double v0, double v1) {
        *reinterpret_cast<double2*>(p + i * stride) = double2{v0, v1};
    };
    auto step = [&](double x, double& m1, double& m2, double& m3) {
        const double v = B * x + a1 * m1 + a2 * m2 + a3 * m3;
        m3 = m2;
        m2 = m1;
        m1 = v;
        return v;
    };
    const int64_t nch = (len + K - 1) / K;
    double2* const ck = reinterpret_cast<double2*>(a.ckpt + l);
    const int64_t ck_row = lines / 2;  // double2 per checkpoint row
    auto load = [&](int64_t c, double2 (&x)[K]) {
        const int64_t i0 = c * K;
        // the last chunk may be short: its missing samples read the last
        // sample again (in bounds) and are not used
#pragma unroll
        for (int j = 0; j < K; ++j) x[j] = ld(i0 + j < len ? i0 + j : len - 1);
    };
    const double2 init = ld(0);
    // causal pass
    double y0[3] = {init.x, init.x, init.x}, y1[3] = {init.y, init.y, init.y};
    double2 nxt[K];
    load(0, nxt);
    for (int64_t c = 0; c < nch; ++c) {
        double2 x[K];
#pragma unroll
        for (int j = 0; j < K; ++j) x[j] = nxt[j];
        if (c + 1 < nch) load(c + 1, nxt);
        if (c > 0) {
#pragma unroll
            for (int k = 0; k < 3; ++k) ck[((c - 1) * 3 + k) * ck_row] = double2{y0[k], y1[k]};
        }
        const int64_t cnt = len - c * K;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (cnt >= K || j < cnt) {
                step(x[j].x, y0[0], y0[1], y0[2]);
                step(x[j].y, y1[0], y1[1], y1[2]);
            }
        }
    }
    // anticausal pass, from each line's last causal value
    double b0[3] = {y0[0], y0[0], y0[0]}, b1[3] = {y1[0], y1[0], y1[0]};
    double2 nck[3] = {init, init, init};
    load(nch - 1, nxt);
    if (nch > 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) nck[k] = ck[((nch - 2) * 3 + k) * ck_row];
    for (int64_t c = nch - 1; c >= 0; --c) {
        double2 x[K];
#pragma unroll
        for (int j = 0; j < K; ++j) x[j] = nxt[j];
        // each line's causal state at the chunk's start
        double s0[3] = {nck[0].x, nck[1].x, nck[2].x}, s1[3] = {nck[0].y, nck[1].y, nck[2].y};
        if (c > 0) {
            load(c - 1, nxt);
            if (c > 1) {
#pragma unroll
                for (int k = 0; k < 3; ++k) nck[k] = ck[((c - 2) * 3 + k) * ck_row];
            } else {
                nck[0] = nck[1] = nck[2] = init;
            }
        }
        const int64_t i0 = c * K, cnt = len - i0;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (cnt >= K || j < cnt) {
                x[j].x = step(x[j].x, s0[0], s0[1], s0[2]);
                x[j].y = step(x[j].y, s1[0], s1[1], s1[2]);
            }
        }
#pragma unroll
        for (int j = K - 1; j >= 0; --j) {
            if (cnt >= K || j < cnt) {
                x[j].x = step(x[j].x, b0[0], b0[1], b0[2]);
                x[j].y = step(x[j].y, b1[0], b1[1], b1[2]);
            }
        }
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (cnt >= K || j < cnt) st(i0 + j, x[j].x, x[j].y);
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

template <typename T, int IO>
void launch_line_io(bool rows, dim3 grid, hipStream_t s, const RecursiveArgs<T>& a, int ax,
                    const double* c) {
    if (rows)
        hipLaunchKernelGGL((line_kernel<T, IO, true>), grid, dim3(256), 0, s, a, ax, c[0], c[1], c[2],
                           c[3]);
    else
        hipLaunchKernelGGL((line_kernel<T, IO, false>), grid, dim3(256), 0, s, a, ax, c[0], c[1], c[2],
                           c[3]);
}

template <typename T>
void launch_line(int io, bool rows, dim3 grid, hipStream_t s, const RecursiveArgs<T>& a, int ax,
                 const double* c) {
    switch (io) {
        case 0: launch_line_io<T, 0>(rows, grid, s, a, ax, c); break;
        case kFromIn: launch_line_io<T, kFromIn>(rows, grid, s, a, ax, c); break;
        case kToOut: launch_line_io<T, kToOut>(rows, grid, s, a, ax, c); break;
        default: launch_line_io<T, kFromIn | kToOut>(rows, grid, s, a, ax, c); break;
    }
}

}  // namespace

template <typename T>
hipError_t launch_recursive(const RecursiveArgs<T>& a, const int* axes, int naxes,
                            const double (*coef)[4], hipStream_t s) {
    const int64_t total = a.n[0] * a.n[1] * a.n[2] * a.n[3];
    if (total == 0) return hipSuccess;
    int eff[4], ne = 0;  // the listed axes with lines of more than one sample
    for (int k = 0; k < naxes; ++k)
        if (a.n[axes[k]] > 1) eff[ne++] = k;  // a line of one sample is left as it is
    const int64_t m = a.m[0] * a.m[1] * a.m[2] * a.m[3];
    if (ne == 0) {
        // the copy grids: rows of axis 1 and planes of axes 2-3 as grid y / z
        if (a.n[1] > 65535 || a.n[2] * a.n[3] > 65535 || a.m[1] > 65535 ||
            a.m[2] * a.m[3] > 65535)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL(widen_kernel<T>,
                           dim3((unsigned)((a.n[0] + kRowBlock - 1) / kRowBlock), (unsigned)a.n[1],
                                (unsigned)(a.n[2] * a.n[3])),
                           dim3(256), 0, s, a);
        if (m > 0)
            hipLaunchKernelGGL(narrow_kernel<T>,
                               dim3((unsigned)((a.m[0] + kRowBlock - 1) / kRowBlock),
                                    (unsigned)a.m[1], (unsigned)(a.m[2] * a.m[3])),
                               dim3(256), 0, s, a);
        return hipGetLastError();
    }
    if (m == 0) return hipSuccess;
    for (int e = 0; e < ne; ++e) {
        const int k = eff[e], ax = axes[k];
        const int64_t lines = total / a.n[ax];
        const double* c = coef[k];
        const int io = (e == 0 ? kFromIn : 0) | (e == ne - 1 ? kToOut : 0);
        if (ax != 0 && io == 0 && a.n[0] % 2 == 0) {
            // the work image in place, two lines a thread (the work image is
            // 256-byte aligned and its rows, like the checkpoint rows, hold
            // an even number of doubles)
            const dim3 grid((unsigned)((lines / 2 + 255) / 256));
            hipLaunchKernelGGL((line_kernel2<T>), grid, dim3(256), 0, s, a, ax, c[0], c[1], c[2], c[3]);
            continue;
        }
        const dim3 grid((unsigned)((lines + 255) / 256));
        launch_line<T>(io, ax == 0, grid, s, a, ax, c);
    }
    return hipGetLastError();
}

template hipError_t launch_recursive<float>(const RecursiveArgs<float>&, const int*, int,
                                            const double (*)[4], hipStream_t);
template hipError_t launch_recursive<double>(const RecursiveArgs<double>&, const int*, int,
                                             const double (*)[4], hipStream_t);

}  // namespace itkg
