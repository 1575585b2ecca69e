/*
 * gauss_oracle.c -- CPU ORACLE FOR TESTS ONLY.
 *
 * This file is test infrastructure.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it, and
 * only as the checker (or as the timed CPU baseline), never as a product code
 * path.  The product library (libitkg.so) does not link or call it.
 *
 * It restates, in plain C, the itk-rs separable Gaussian path:
 *
 *   kernel_radius          src/filters/gaussian.rs:24-27
 *   build_kernel           src/filters/gaussian.rs:29-40
 *   generate_region        src/filters/gaussian.rs:59-90
 *   convolve_axis          src/filters/conv.rs:11-41
 *   Region::padded_per_axis / clipped_to   src/image.rs:28-37, 52-64
 *   Image::flat_index      src/image.rs:111-120   (axis 0 fastest)
 *   f32 / f64 / VecPixel scale+add         src/pixel.rs:19-31, 71-99
 *   leaf Image::generate_region (clipped copy)   src/source.rs:42-58
 *
 * Arithmetic contract (what makes the restatement bit-exact):
 *   - weights are computed in f64 with glibc exp (the symbol Rust's f64::exp
 *     calls on linux-gnu), normalised by a sequential f64 sum;
 *   - each tap is fl32(x * (float)w_k) and is added to an f32 accumulator
 *     that starts at +0.0, taps in order k = 0..2r, with no FMA
 *     (build with -ffp-contract=off);
 *   - each pass writes f32; passes run in axis order 0..D-1 and clamp to the
 *     pass input's region on the pass axis only.
 *
 * Pinning (each function is checked by tests/test_oracle.py):
 *   kernel_radius / build_kernel : closed form, symmetry, unit sum, radius
 *                                  boundary cases (3*(s/sp) evaluation order);
 *   generate_region (f32)        : bit-exact against an independent numpy
 *                                  float32 restatement, within 1e-6 of
 *                                  scipy.ndimage.correlate1d(mode='nearest'),
 *                                  interior-impulse closed form, the
 *                                  reference's own property tests, brute-force
 *                                  dense Kronecker operator in f64;
 *   generate_region (f64)        : dense Kronecker operator (1e-12);
 *   adjoint (f64 scatter)        : dense transpose of the Kronecker operator
 *                                  and <Ax, y> = <x, A^T y>;
 *   separable_opt (bench's optimised CPU comparison): bit-exact against
 *                                  generate_region;
 *   ref_faithful (threaded)      : bit-exact against generate_region.
 *
 * The other convolve_axis filters, restated the same way:
 *
 *   box_kernel                      src/filters/mean.rs:15-18
 *   MeanFilter / BoxMeanFilterN     src/filters/mean.rs:59-83, 126-144
 *     ::generate_region
 *   BinomialBlurFilter              src/filters/binomial_blur.rs:57-89
 *     ::generate_region (its own per-pass padding rule, loop for loop)
 *   UnsharpMaskFilter               src/filters/edges.rs:406-452
 *     ::generate_region (whole-buffer passes, the f64 per-pixel step)
 *   RecursiveGaussianFilter /       src/filters/recursive_gaussian.rs:33-240
 *     RecursiveGaussianAxisFilter (yvv_coeffs, causal / anticausal lines)
 *   CannyEdgeDetectionFilter        src/filters/edges.rs:548-664 (2-D)
 *
 *   mean / binomial (f32)          : interior-impulse closed forms (box
 *                                  products; binomial coefficients / 4^n,
 *                                  exact in f32), the reference's own tests
 *                                  (constant, ramp interior, impulse
 *                                  spread), brute-force dense operator in
 *                                  f64 (tests/test_oracle_filters.py).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_DIM 4

/* ---- region algebra: src/image.rs:17-77 -------------------------------- */

typedef struct {
    int64_t index[ORC_MAX_DIM];
    int64_t size[ORC_MAX_DIM];
} orc_region;

static orc_region region_make(int D, const int64_t* idx, const int64_t* sz) {
    orc_region r;
    memset(&r, 0, sizeof r);
    for (int d = 0; d < D; ++d) { r.index[d] = idx[d]; r.size[d] = sz[d]; }
    return r;
}

/* src/image.rs:28-37 */
static orc_region region_padded(int D, orc_region r, const uint64_t* radii) {
    for (int d = 0; d < D; ++d) {
        r.index[d] -= (int64_t)radii[d];
        r.size[d] += 2 * (int64_t)radii[d];
    }
    return r;
}

/* src/image.rs:52-64 */
static orc_region region_clipped(int D, orc_region r, orc_region b) {
    for (int d = 0; d < D; ++d) {
        int64_t self_end = r.index[d] + r.size[d];
        int64_t b_end = b.index[d] + b.size[d];
        int64_t s = r.index[d] > b.index[d] ? r.index[d] : b.index[d];
        int64_t e = self_end < b_end ? self_end : b_end;
        r.index[d] = s;
        r.size[d] = e - s > 0 ? e - s : 0;
    }
    return r;
}

static int64_t region_len(int D, orc_region r) {
    int64_t n = 1;
    for (int d = 0; d < D; ++d) n *= r.size[d];
    return n;
}

/* src/image.rs:111-120, axis 0 fastest */
static int64_t flat_index(int D, orc_region r, const int64_t* idx) {
    int64_t off = 0, stride = 1;
    for (int d = 0; d < D; ++d) {
        off += (idx[d] - r.index[d]) * stride;
        stride *= r.size[d];
    }
    return off;
}

/* ---- kernel: src/filters/gaussian.rs:24-40 ----------------------------- */

/* `(3.0 * (sigma / spacing)).ceil() as usize`; Rust's float->usize cast
 * saturates: NaN and negatives become 0. */
uint64_t orc_kernel_radius(double sigma, double spacing) {
    double sigma_px = sigma / spacing;
    double v = ceil(3.0 * sigma_px);
    if (!(v > 0.0)) return 0;
    if (v >= 18446744073709551616.0) return UINT64_MAX;
    return (uint64_t)v;
}

/* w_i = exp(-0.5 * ((i - r) / sigma_px)^2), normalised by the sequential
 * f64 sum.  powi(2) is a single multiply. */
void orc_build_kernel(double sigma_px, uint64_t radius, double* k) {
    uint64_t len = 2 * radius + 1;
    for (uint64_t i = 0; i < len; ++i) {
        double x = (double)i - (double)radius;
        double t = x / sigma_px;
        k[i] = exp(-0.5 * (t * t));
    }
    double sum = 0.0;
    for (uint64_t i = 0; i < len; ++i) sum += k[i];
    for (uint64_t i = 0; i < len; ++i) k[i] /= sum;
}

/* ---- convolve_axis: src/filters/conv.rs:11-41 --------------------------- */

#define DEFINE_CONV(NAME, T, WCAST)                                              \
static void NAME(int D, int ncomp, const T* src, orc_region sreg, int axis,      \
                 const double* kern, uint64_t klen, orc_region req, T* out) {    \
    int64_t n = region_len(D, req);                                              \
    if (n == 0) return;                                                          \
    int64_t radius = (int64_t)(klen / 2);                                        \
    int64_t lo = sreg.index[axis];                                               \
    int64_t hi = sreg.index[axis] + sreg.size[axis] - 1;                         \
    int64_t idx[ORC_MAX_DIM];                                                    \
    for (int d = 0; d < D; ++d) idx[d] = req.index[d];                           \
    for (int64_t o = 0; o < n; ++o) {                                            \
        for (int c = 0; c < ncomp; ++c) {                                        \
            T acc = (T)0.0;                                                      \
            for (uint64_t k = 0; k < klen; ++k) {                                \
                int64_t s[ORC_MAX_DIM];                                          \
                for (int d = 0; d < D; ++d) s[d] = idx[d];                       \
                int64_t v = idx[axis] + (int64_t)k - radius;                     \
                if (v < lo) v = lo;                                              \
                if (v > hi) v = hi;                                              \
                s[axis] = v;                                                     \
                T x = src[flat_index(D, sreg, s) * ncomp + c];                   \
                acc = acc + x * (T)WCAST(kern[k]);                               \
            }                                                                    \
            out[o * ncomp + c] = acc;                                            \
        }                                                                        \
        for (int d = 0; d < D; ++d) {                                            \
            if (++idx[d] < req.index[d] + req.size[d]) break;                    \
            idx[d] = req.index[d];                                               \
        }                                                                        \
    }                                                                            \
}

#define AS_F32(w) ((float)(w))
#define AS_F64(w) (w)
DEFINE_CONV(conv_axis_f32, float, AS_F32)
DEFINE_CONV(conv_axis_f64, double, AS_F64)

/* Leaf source (src/source.rs:42-58): clipped copy of `img` to `want`. */
#define DEFINE_LEAF(NAME, T)                                                     \
static T* NAME(int D, int ncomp, const T* img, orc_region ireg, orc_region want, \
               orc_region* got) {                                                \
    orc_region c = region_clipped(D, want, ireg);                                \
    int64_t n = region_len(D, c);                                                \
    T* out = (T*)malloc((size_t)(n > 0 ? n : 1) * ncomp * sizeof(T));           \
    int64_t idx[ORC_MAX_DIM];                                                    \
    for (int d = 0; d < D; ++d) idx[d] = c.index[d];                             \
    for (int64_t o = 0; o < n; ++o) {                                            \
        int64_t si = flat_index(D, ireg, idx);                                   \
        for (int k = 0; k < ncomp; ++k) out[o * ncomp + k] = img[si * ncomp + k];\
        for (int d = 0; d < D; ++d) {                                            \
            if (++idx[d] < c.index[d] + c.size[d]) break;                        \
            idx[d] = c.index[d];                                                 \
        }                                                                        \
    }                                                                            \
    *got = c;                                                                    \
    return out;                                                                  \
}
DEFINE_LEAF(leaf_copy_f32, float)
DEFINE_LEAF(leaf_copy_f64, double)

/* ---- generate_region: src/filters/gaussian.rs:59-90 -------------------- */
/* The upstream is a leaf Image whose region is the largest region (bounds).
 * `out` receives linear_len(requested) * ncomp values.  Returns 0 on success,
 * -1 on a bad argument. */
#define DEFINE_GEN(NAME, T, LEAF, CONV)                                          \
int NAME(int D, int ncomp, const T* img, const int64_t* img_index,               \
         const int64_t* img_size, const double* spacing, double sigma,           \
         const int64_t* req_index, const int64_t* req_size, T* out) {            \
    if (D < 1 || D > ORC_MAX_DIM || ncomp < 1) return -1;                        \
    orc_region bounds = region_make(D, img_index, img_size);                     \
    orc_region req = region_make(D, req_index, req_size);                        \
    uint64_t radii[ORC_MAX_DIM] = {0};                                           \
    for (int d = 0; d < D; ++d) radii[d] = orc_kernel_radius(sigma, spacing[d]); \
    orc_region in_reg = region_clipped(D, region_padded(D, req, radii), bounds); \
    orc_region cur_reg;                                                          \
    T* cur = LEAF(D, ncomp, img, bounds, in_reg, &cur_reg);                      \
    for (int axis = 0; axis < D; ++axis) {                                       \
        double sigma_px = sigma / spacing[axis];                                 \
        uint64_t klen = 2 * radii[axis] + 1;                                     \
        double* kern = (double*)malloc(klen * sizeof(double));                   \
        orc_build_kernel(sigma_px, radii[axis], kern);                           \
        orc_region out_reg;                                                      \
        if (axis + 1 < D) {                                                      \
            uint64_t rem[ORC_MAX_DIM] = {0};                                     \
            for (int d = axis + 1; d < D; ++d) rem[d] = radii[d];                \
            out_reg = region_clipped(D, region_padded(D, req, rem), bounds);     \
        } else {                                                                 \
            out_reg = req;                                                       \
        }                                                                        \
        int64_t n = region_len(D, out_reg);                                      \
        T* nxt = (T*)malloc((size_t)(n > 0 ? n : 1) * ncomp * sizeof(T));       \
        CONV(D, ncomp, cur, cur_reg, axis, kern, klen, out_reg, nxt);            \
        free(kern);                                                              \
        free(cur);                                                               \
        cur = nxt;                                                               \
        cur_reg = out_reg;                                                       \
    }                                                                            \
    memcpy(out, cur, (size_t)region_len(D, req) * ncomp * sizeof(T));            \
    free(cur);                                                                   \
    return 0;                                                                    \
}
DEFINE_GEN(orc_generate_region_f32, float, leaf_copy_f32, conv_axis_f32)
DEFINE_GEN(orc_generate_region_f64, double, leaf_copy_f64, conv_axis_f64)

/* ---- MeanFilter / BoxMeanFilterN: src/filters/mean.rs:15-18, 59-83,
 * 126-144.  Per-axis radii (MeanFilter: all equal); the same pass regions as
 * generate_region, each pass convolve_axis with box_kernel(radii[axis]). */
void orc_box_kernel(uint64_t radius, double* k) {
    uint64_t len = 2 * radius + 1;
    for (uint64_t i = 0; i < len; ++i) k[i] = 1.0 / (double)len;
}

#define DEFINE_MEAN(NAME, T, LEAF, CONV)                                         \
int NAME(int D, int ncomp, const T* img, const int64_t* img_index,               \
         const int64_t* img_size, const uint64_t* radii,                         \
         const int64_t* req_index, const int64_t* req_size, T* out) {            \
    if (D < 1 || D > ORC_MAX_DIM || ncomp < 1) return -1;                        \
    orc_region bounds = region_make(D, img_index, img_size);                     \
    orc_region req = region_make(D, req_index, req_size);                        \
    orc_region in_reg = region_clipped(D, region_padded(D, req, radii), bounds); \
    orc_region cur_reg;                                                          \
    T* cur = LEAF(D, ncomp, img, bounds, in_reg, &cur_reg);                      \
    for (int axis = 0; axis < D; ++axis) {                                       \
        uint64_t klen = 2 * radii[axis] + 1;                                     \
        double* kern = (double*)malloc(klen * sizeof(double));                   \
        orc_box_kernel(radii[axis], kern);                                       \
        orc_region out_reg;                                                      \
        if (axis + 1 < D) {                                                      \
            uint64_t rem[ORC_MAX_DIM] = {0};                                     \
            for (int d = axis + 1; d < D; ++d) rem[d] = radii[d];                \
            out_reg = region_clipped(D, region_padded(D, req, rem), bounds);     \
        } else {                                                                 \
            out_reg = req;                                                       \
        }                                                                        \
        int64_t n = region_len(D, out_reg);                                      \
        T* nxt = (T*)malloc((size_t)(n > 0 ? n : 1) * ncomp * sizeof(T));       \
        CONV(D, ncomp, cur, cur_reg, axis, kern, klen, out_reg, nxt);            \
        free(kern);                                                              \
        free(cur);                                                               \
        cur = nxt;                                                               \
        cur_reg = out_reg;                                                       \
    }                                                                            \
    memcpy(out, cur, (size_t)region_len(D, req) * ncomp * sizeof(T));            \
    free(cur);                                                                   \
    return 0;                                                                    \
}
DEFINE_MEAN(orc_mean_generate_region_f32, float, leaf_copy_f32, conv_axis_f32)
DEFINE_MEAN(orc_mean_generate_region_f64, double, leaf_copy_f64, conv_axis_f64)

/* ---- BinomialBlurFilter: src/filters/binomial_blur.rs:57-89, loop for
 * loop: fetch the request padded by `reps` on every axis; for each
 * repetition and axis, convolve_axis with [1/4, 1/2, 1/4] into the request
 * padded by (reps - rep - 1) + (1 on the axes not yet done in this
 * repetition), clipped to the bounds; the very last pass writes the request
 * itself. */
static const double ORC_BINOMIAL[3] = {0.25, 0.5, 0.25};

#define DEFINE_BINOMIAL(NAME, T, LEAF, CONV)                                     \
int NAME(int D, int ncomp, const T* img, const int64_t* img_index,               \
         const int64_t* img_size, uint64_t reps,                                 \
         const int64_t* req_index, const int64_t* req_size, T* out) {            \
    if (D < 1 || D > ORC_MAX_DIM || ncomp < 1) return -1;                        \
    orc_region bounds = region_make(D, img_index, img_size);                     \
    orc_region req = region_make(D, req_index, req_size);                        \
    uint64_t all[ORC_MAX_DIM] = {0};                                             \
    for (int d = 0; d < D; ++d) all[d] = reps;                                   \
    orc_region in_reg = region_clipped(D, region_padded(D, req, all), bounds);   \
    orc_region cur_reg;                                                          \
    T* cur = LEAF(D, ncomp, img, bounds, in_reg, &cur_reg);                      \
    for (uint64_t rep = 0; rep < reps; ++rep) {                                  \
        uint64_t remaining = reps - rep - 1;                                     \
        for (int axis = 0; axis < D; ++axis) {                                   \
            int is_last = rep + 1 == reps && axis + 1 == D;                      \
            orc_region out_reg;                                                  \
            if (is_last) {                                                       \
                out_reg = req;                                                   \
            } else {                                                             \
                uint64_t pad[ORC_MAX_DIM] = {0};                                 \
                for (int d = 0; d < D; ++d)                                      \
                    pad[d] = d > axis ? remaining + 1 : remaining;               \
                out_reg = region_clipped(D, region_padded(D, req, pad), bounds); \
            }                                                                    \
            int64_t n = region_len(D, out_reg);                                  \
            T* nxt = (T*)malloc((size_t)(n > 0 ? n : 1) * ncomp * sizeof(T));   \
            CONV(D, ncomp, cur, cur_reg, axis, ORC_BINOMIAL, 3, out_reg, nxt);   \
            free(cur);                                                           \
            cur = nxt;                                                           \
            cur_reg = out_reg;                                                   \
        }                                                                        \
    }                                                                            \
    memcpy(out, cur, (size_t)region_len(D, cur_reg) * ncomp * sizeof(T));        \
    free(cur);                                                                   \
    return 0;                                                                    \
}
DEFINE_BINOMIAL(orc_binomial_generate_region_f32, float, leaf_copy_f32, conv_axis_f32)
DEFINE_BINOMIAL(orc_binomial_generate_region_f64, double, leaf_copy_f64, conv_axis_f64)

/* ---- UnsharpMaskFilter: src/filters/edges.rs:406-452, loop for loop:
 * radius ceil(3 sigma) (index space), its own kernel formula; the input is
 * the request padded by the radius, clipped; every pass convolves the WHOLE
 * padded buffer (its output region is the buffer's region); then, per
 * requested pixel, in f64: edge = orig - blur, orig + amount * edge where
 * |edge| > threshold, else orig, cast to T.  Scalar pixels. */
void orc_index_kernel(double sigma, uint64_t radius, double* k) {
    uint64_t len = 2 * radius + 1;
    for (uint64_t i = 0; i < len; ++i) {
        double d = (double)i - (double)radius;
        k[i] = exp(-(d * d) / (2.0 * sigma * sigma));
    }
    double sum = 0.0;
    for (uint64_t i = 0; i < len; ++i) sum += k[i];
    for (uint64_t i = 0; i < len; ++i) k[i] /= sum;
}

#define DEFINE_UNSHARP(NAME, T, LEAF, CONV)                                      \
int NAME(int D, const T* img, const int64_t* img_index, const int64_t* img_size, \
         double sigma, double amount, double threshold,                          \
         const int64_t* req_index, const int64_t* req_size, T* out) {            \
    if (D < 1 || D > ORC_MAX_DIM) return -1;                                     \
    orc_region bounds = region_make(D, img_index, img_size);                     \
    orc_region req = region_make(D, req_index, req_size);                        \
    uint64_t radius = orc_kernel_radius(sigma, 1.0);                             \
    uint64_t radii[ORC_MAX_DIM] = {0};                                           \
    for (int d = 0; d < D; ++d) radii[d] = radius;                               \
    orc_region pad = region_clipped(D, region_padded(D, req, radii), bounds);    \
    orc_region in_reg;                                                           \
    T* input = LEAF(D, 1, img, bounds, pad, &in_reg);                            \
    uint64_t klen = 2 * radius + 1;                                              \
    double* kern = (double*)malloc(klen * sizeof(double));                       \
    orc_index_kernel(sigma, radius, kern);                                       \
    int64_t n = region_len(D, in_reg);                                           \
    T* blurred = (T*)malloc((size_t)(n > 0 ? n : 1) * sizeof(T));               \
    memcpy(blurred, input, (size_t)n * sizeof(T));                              \
    for (int axis = 0; axis < D; ++axis) {                                       \
        T* nxt = (T*)malloc((size_t)(n > 0 ? n : 1) * sizeof(T));               \
        CONV(D, 1, blurred, in_reg, axis, kern, klen, in_reg, nxt);              \
        free(blurred);                                                           \
        blurred = nxt;                                                           \
    }                                                                            \
    int64_t m = region_len(D, req);                                              \
    int64_t idx[ORC_MAX_DIM];                                                    \
    for (int d = 0; d < D; ++d) idx[d] = req.index[d];                           \
    for (int64_t o = 0; o < m; ++o) {                                            \
        int64_t f = flat_index(D, in_reg, idx);                                  \
        double orig = (double)input[f];                                          \
        double blur = (double)blurred[f];                                        \
        double edge = orig - blur;                                               \
        double enhanced = fabs(edge) > threshold ? orig + amount * edge : orig;  \
        out[o] = (T)enhanced;                                                    \
        for (int d = 0; d < D; ++d) {                                            \
            if (++idx[d] < req.index[d] + req.size[d]) break;                    \
            idx[d] = req.index[d];                                               \
        }                                                                        \
    }                                                                            \
    free(kern);                                                                  \
    free(blurred);                                                               \
    free(input);                                                                 \
    return 0;                                                                    \
}
DEFINE_UNSHARP(orc_unsharp_generate_region_f32, float, leaf_copy_f32, conv_axis_f32)
DEFINE_UNSHARP(orc_unsharp_generate_region_f64, double, leaf_copy_f64, conv_axis_f64)

/* ---- Recursive Gaussian: src/filters/recursive_gaussian.rs, loop for loop.
 * yvv_coeffs (33-47); causal_pass / anticausal_pass (55-86) over every line
 * of filter_axis (91-127), lines in iter_region order of the stripe region;
 * RecursiveGaussianFilter (axis = -1, axes 0..D-1) and
 * RecursiveGaussianAxisFilter (one axis): whole image in f64, then the
 * requested region cast to T (147-176, 219-240).  Scalar pixels. */
void orc_yvv_coeffs(double sigma, double* c) {
    sigma = fmax(sigma, 0.5);
    double q = sigma >= 2.5 ? 0.98711 * sigma - 0.96330
                            : 3.97156 - 4.14554 * sqrt(1.0 - 0.26891 * sigma);
    double q2 = q * q;
    double q3 = q2 * q;
    double b0 = 1.57825 + 2.44413 * q + 1.4281 * q2 + 0.422205 * q3;
    double b1 = 2.44413 * q + 2.85619 * q2 + 1.26661 * q3;
    double b2 = -(1.4281 * q2 + 1.26661 * q3);
    double b3 = 0.422205 * q3;
    c[0] = 1.0 - (b1 + b2 + b3) / b0;
    c[1] = b1 / b0;
    c[2] = b2 / b0;
    c[3] = b3 / b0;
}

static void orc_filter_axis(int D, double* data, orc_region r, int axis, double sigma_px) {
    int64_t size = r.size[axis];
    if (size <= 1) return;
    double c[4];
    orc_yvv_coeffs(sigma_px, c);
    double B = c[0], a1 = c[1], a2 = c[2], a3 = c[3];
    orc_region stripe = r;
    stripe.size[axis] = 1;
    int64_t nl = region_len(D, stripe);
    int64_t idx[ORC_MAX_DIM];
    for (int d = 0; d < D; ++d) idx[d] = stripe.index[d];
    int64_t* line = (int64_t*)malloc((size_t)size * sizeof(int64_t));
    for (int64_t l = 0; l < nl; ++l) {
        for (int64_t k = 0; k < size; ++k) {
            int64_t p[ORC_MAX_DIM];
            for (int d = 0; d < D; ++d) p[d] = idx[d];
            p[axis] += k;
            line[k] = flat_index(D, r, p);
        }
        double init = data[line[0]];
        double ym1 = init, ym2 = init, ym3 = init;
        for (int64_t i = 0; i < size; ++i) {
            double v = B * data[line[i]] + a1 * ym1 + a2 * ym2 + a3 * ym3;
            ym3 = ym2; ym2 = ym1; ym1 = v;
            data[line[i]] = v;
        }
        init = data[line[size - 1]];
        ym1 = init; ym2 = init; ym3 = init;
        for (int64_t i = size - 1; i >= 0; --i) {
            double v = B * data[line[i]] + a1 * ym1 + a2 * ym2 + a3 * ym3;
            ym3 = ym2; ym2 = ym1; ym1 = v;
            data[line[i]] = v;
        }
        for (int d = 0; d < D; ++d) {
            if (++idx[d] < stripe.index[d] + stripe.size[d]) break;
            idx[d] = stripe.index[d];
        }
    }
    free(line);
}

#define DEFINE_RECURSIVE(NAME, T)                                                \
int NAME(int D, const T* img, const int64_t* img_index, const int64_t* img_size, \
         const double* spacing, double sigma, int axis,                          \
         const int64_t* req_index, const int64_t* req_size, T* out) {            \
    if (D < 1 || D > ORC_MAX_DIM || axis < -1 || axis >= D) return -1;          \
    orc_region bounds = region_make(D, img_index, img_size);                     \
    orc_region req = region_make(D, req_index, req_size);                        \
    int64_t n = region_len(D, bounds);                                           \
    double* data = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));   \
    for (int64_t i = 0; i < n; ++i) data[i] = (double)img[i];                    \
    for (int a = 0; a < D; ++a)                                                  \
        if (axis < 0 || a == axis) orc_filter_axis(D, data, bounds, a, sigma / spacing[a]); \
    int64_t m = region_len(D, req);                                              \
    int64_t idx[ORC_MAX_DIM];                                                    \
    for (int d = 0; d < D; ++d) idx[d] = req.index[d];                           \
    for (int64_t o = 0; o < m; ++o) {                                            \
        out[o] = (T)data[flat_index(D, bounds, idx)];                            \
        for (int d = 0; d < D; ++d) {                                            \
            if (++idx[d] < req.index[d] + req.size[d]) break;                    \
            idx[d] = req.index[d];                                               \
        }                                                                        \
    }                                                                            \
    free(data);                                                                  \
    return 0;                                                                    \
}
DEFINE_RECURSIVE(orc_recursive_generate_region_f32, float)
DEFINE_RECURSIVE(orc_recursive_generate_region_f64, double)

/* Direction bin of the suppression step (src/filters/edges.rs:612-620) for
 * a direction r = atan2(gy, gx) in radians: to_degrees (one multiply by the
 * double nearest 180/pi), rem_euclid(180), then 0 below 22.5 or from 157.5,
 * 1 below 67.5, 2 below 112.5, else 3. */
static int canny_bin_of(double r) {
    double deg = r * (180.0 / 3.14159265358979323846);
    double angle = fmod(deg, 180.0);
    if (angle < 0) angle += 180.0;
    if (angle < 22.5 || angle >= 157.5) return 0;
    if (angle < 67.5) return 1;
    if (angle < 112.5) return 2;
    return 3;
}

/* The bin of each gradient (gx[k], gy[k]), with the host libm's atan2 as
 * the reference's f64::atan2 (src/filters/edges.rs:602). */
void orc_canny_bins(int64_t n, const double* gx, const double* gy, unsigned char* out) {
    for (int64_t k = 0; k < n; ++k) out[k] = (unsigned char)canny_bin_of(atan2(gy[k], gx[k]));
}

/* ---- CannyEdgeDetectionFilter: src/filters/edges.rs:548-664, loop for
 * loop (2-D): smooth the whole image with the index-space kernel (every
 * pass over the whole buffer), gradient by clamped central differences in
 * f64, magnitude sqrt(gx*gx + gy*gy) and direction atan2(gy, gx); non-
 * maximum suppression along the direction binned to 0/45/90/135 degrees;
 * hysteresis: strong = nms >= upper, then sweeps adding every pixel with
 * nms >= lower next to an edge (8 neighbours, clamped) until none changes;
 * the requested region as 1 / 0. */
#define DEFINE_CANNY(NAME, T, LEAF, CONV)                                        \
int NAME(const T* img, const int64_t* img_index, const int64_t* img_size,        \
         double sigma, double lower, double upper,                               \
         const int64_t* req_index, const int64_t* req_size, T* out) {            \
    const int D = 2;                                                             \
    orc_region bounds = region_make(D, img_index, img_size);                     \
    orc_region req = region_make(D, req_index, req_size);                        \
    orc_region full_reg;                                                         \
    T* full = LEAF(D, 1, img, bounds, bounds, &full_reg);                        \
    uint64_t radius = orc_kernel_radius(sigma, 1.0);                             \
    uint64_t klen = 2 * radius + 1;                                              \
    double* kern = (double*)malloc(klen * sizeof(double));                       \
    orc_index_kernel(sigma, radius, kern);                                       \
    int64_t n = region_len(D, full_reg);                                         \
    T* sm = (T*)malloc((size_t)(n > 0 ? n : 1) * sizeof(T));                     \
    memcpy(sm, full, (size_t)n * sizeof(T));                                     \
    for (int axis = 0; axis < D; ++axis) {                                       \
        T* nxt = (T*)malloc((size_t)(n > 0 ? n : 1) * sizeof(T));               \
        CONV(D, 1, sm, full_reg, axis, kern, klen, full_reg, nxt);               \
        free(sm);                                                                \
        sm = nxt;                                                                \
    }                                                                            \
    int64_t nx = bounds.size[0], ny = bounds.size[1];                            \
    int64_t ox = bounds.index[0], oy = bounds.index[1];                          \
    double* smd = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));     \
    for (int64_t i = 0; i < n; ++i) smd[i] = (double)sm[i];                      \
    double* mag = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));      \
    double* dir = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));      \
    for (int64_t j = 0; j < ny; ++j)                                             \
        for (int64_t i = 0; i < nx; ++i) {                                       \
            int64_t ip = i + 1 > nx - 1 ? nx - 1 : i + 1, im = i - 1 < 0 ? 0 : i - 1; \
            int64_t jp = j + 1 > ny - 1 ? ny - 1 : j + 1, jm = j - 1 < 0 ? 0 : j - 1; \
            double gx = smd[ip + j * nx] - smd[im + j * nx];                     \
            double gy = smd[i + jp * nx] - smd[i + jm * nx];                     \
            mag[i + j * nx] = sqrt(gx * gx + gy * gy);                           \
            dir[i + j * nx] = atan2(gy, gx);                                     \
        }                                                                        \
    double* nms = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));      \
    for (int64_t j = 0; j < ny; ++j)                                             \
        for (int64_t i = 0; i < nx; ++i) {                                       \
            int64_t f = i + j * nx;                                              \
            double m = mag[f];                                                   \
            const int bin = canny_bin_of(dir[f]);                                \
            const int64_t di = bin == 0 ? 1 : (bin == 1 ? 1 : (bin == 2 ? 0 : -1)); \
            const int64_t dj = bin == 0 ? 0 : 1;                                 \
            int64_t a_i = i + di, a_j = j + dj, b_i = i - di, b_j = j - dj;      \
            a_i = a_i < 0 ? 0 : (a_i > nx - 1 ? nx - 1 : a_i);                   \
            a_j = a_j < 0 ? 0 : (a_j > ny - 1 ? ny - 1 : a_j);                   \
            b_i = b_i < 0 ? 0 : (b_i > nx - 1 ? nx - 1 : b_i);                   \
            b_j = b_j < 0 ? 0 : (b_j > ny - 1 ? ny - 1 : b_j);                   \
            if (m >= mag[a_i + a_j * nx] && m >= mag[b_i + b_j * nx]) nms[f] = m; \
        }                                                                        \
    unsigned char* edges = (unsigned char*)calloc((size_t)(n > 0 ? n : 1), 1);   \
    for (int64_t f = 0; f < n; ++f) if (nms[f] >= upper) edges[f] = 1;          \
    int changed = 1;                                                             \
    while (changed) {                                                            \
        changed = 0;                                                             \
        for (int64_t j = 0; j < ny; ++j)                                         \
            for (int64_t i = 0; i < nx; ++i) {                                   \
                int64_t f = i + j * nx;                                          \
                if (!edges[f] && nms[f] >= lower) {                              \
                    for (int64_t dj = -1; dj <= 1; ++dj)                         \
                        for (int64_t di = -1; di <= 1; ++di) {                   \
                            int64_t ni = i + di, nj = j + dj;                    \
                            ni = ni < 0 ? 0 : (ni > nx - 1 ? nx - 1 : ni);       \
                            nj = nj < 0 ? 0 : (nj > ny - 1 ? ny - 1 : nj);       \
                            if (edges[ni + nj * nx]) { edges[f] = 1; changed = 1; } \
                        }                                                        \
                }                                                                \
            }                                                                    \
    }                                                                            \
    int64_t m = region_len(D, req), o = 0;                                       \
    for (int64_t j = req.index[1]; j < req.index[1] + req.size[1]; ++j)          \
        for (int64_t i = req.index[0]; i < req.index[0] + req.size[0]; ++i)      \
            out[o++] = edges[(i - ox) + (j - oy) * nx] ? (T)1.0 : (T)0.0;        \
    (void)m;                                                                     \
    free(edges); free(nms); free(dir); free(mag); free(smd);                     \
    free(sm); free(full); free(kern);                                            \
    return 0;                                                                    \
}
DEFINE_CANNY(orc_canny_generate_region_f32, float, leaf_copy_f32, conv_axis_f32)
DEFINE_CANNY(orc_canny_generate_region_f64, double, leaf_copy_f64, conv_axis_f64)

/* ---- adjoint in f64, by scattering the forward taps -------------------- */
/* No reference counterpart.  The forward operator is A = A_{D-1} ... A_0
 * (pass 0 first).  A^T = A_0^T ... A_{D-1}^T, so the passes are undone in
 * reverse axis order; each pass transposes convolve_axis literally: every
 * (output voxel, tap) pair that read src[clamp(...)] * w_k now adds
 * g_out * w_k into that same source voxel.  `gin` receives the gradient on
 * the clipped padded input region (the input_region_for_output of
 * `requested`), dense, axis 0 fastest. */
int orc_adjoint_f64(int D, int ncomp, const int64_t* img_index, const int64_t* img_size,
                    const double* spacing, double sigma, const int64_t* req_index,
                    const int64_t* req_size, const double* gout, double* gin,
                    int64_t* in_index_out, int64_t* in_size_out) {
    if (D < 1 || D > ORC_MAX_DIM || ncomp < 1) return -1;
    orc_region bounds = region_make(D, img_index, img_size);
    orc_region req = region_make(D, req_index, req_size);
    uint64_t radii[ORC_MAX_DIM] = {0};
    for (int d = 0; d < D; ++d) radii[d] = orc_kernel_radius(sigma, spacing[d]);
    orc_region in_reg = region_clipped(D, region_padded(D, req, radii), bounds);
    /* regions of every pass: reg[a] = input of pass a, reg[D] = requested */
    orc_region reg[ORC_MAX_DIM + 1];
    reg[0] = in_reg;
    for (int a = 0; a < D; ++a) {
        if (a + 1 < D) {
            uint64_t rem[ORC_MAX_DIM] = {0};
            for (int d = a + 1; d < D; ++d) rem[d] = radii[d];
            reg[a + 1] = region_clipped(D, region_padded(D, req, rem), bounds);
        } else {
            reg[a + 1] = req;
        }
    }
    int64_t n = region_len(D, req);
    double* g = (double*)malloc((size_t)(n > 0 ? n : 1) * ncomp * sizeof(double));
    memcpy(g, gout, (size_t)n * ncomp * sizeof(double));
    for (int axis = D - 1; axis >= 0; --axis) {
        orc_region sreg = reg[axis], oreg = reg[axis + 1];
        uint64_t klen = 2 * radii[axis] + 1;
        int64_t radius = (int64_t)radii[axis];
        double* kern = (double*)malloc(klen * sizeof(double));
        orc_build_kernel(sigma / spacing[axis], radii[axis], kern);
        int64_t ns = region_len(D, sreg), no = region_len(D, oreg);
        double* gs = (double*)calloc((size_t)(ns > 0 ? ns : 1) * ncomp, sizeof(double));
        int64_t lo = sreg.index[axis], hi = sreg.index[axis] + sreg.size[axis] - 1;
        int64_t idx[ORC_MAX_DIM];
        for (int d = 0; d < D; ++d) idx[d] = oreg.index[d];
        for (int64_t o = 0; o < no; ++o) {
            for (uint64_t k = 0; k < klen; ++k) {
                int64_t s[ORC_MAX_DIM];
                for (int d = 0; d < D; ++d) s[d] = idx[d];
                int64_t v = idx[axis] + (int64_t)k - radius;
                if (v < lo) v = lo;
                if (v > hi) v = hi;
                s[axis] = v;
                int64_t si = flat_index(D, sreg, s);
                for (int c = 0; c < ncomp; ++c) gs[si * ncomp + c] += g[o * ncomp + c] * kern[k];
            }
            for (int d = 0; d < D; ++d) {
                if (++idx[d] < oreg.index[d] + oreg.size[d]) break;
                idx[d] = oreg.index[d];
            }
        }
        free(kern);
        free(g);
        g = gs;
    }
    memcpy(gin, g, (size_t)region_len(D, in_reg) * ncomp * sizeof(double));
    free(g);
    for (int d = 0; d < D; ++d) { in_index_out[d] = in_reg.index[d]; in_size_out[d] = in_reg.size[d]; }
    return 0;
}

/* ---- "ref-faithful" threaded variant (CPU baseline) -------------------- */
/* Same arithmetic as orc_generate_region_f32, with the reference's loop
 * structure: leaf clipped copy (src/source.rs:42-58), a materialised
 * Vec<Index<D>> per pass (src/filters/conv.rs:21-22), a parallel map over
 * output voxels with a per-tap clamp and flat_index (conv.rs:24-37) run on a
 * pool of `nthreads` workers that pull dynamic chunks (the work-splitting
 * rayon's par_iter performs), and an order-preserving collect. */

typedef struct {
    int D, ncomp, axis;
    const float* src;
    orc_region sreg;
    const float* wf;
    int64_t klen, radius;
    const int64_t* indices; /* D int64 per voxel */
    int64_t n;
    float* out;
    int64_t chunk;
    int64_t next; /* atomic */
} ref_job;

static void* ref_worker(void* arg) {
    ref_job* j = (ref_job*)arg;
    const int D = j->D, nc = j->ncomp, axis = j->axis;
    const int64_t lo = j->sreg.index[axis], hi = j->sreg.index[axis] + j->sreg.size[axis] - 1;
    for (;;) {
        int64_t b = __atomic_fetch_add(&j->next, j->chunk, __ATOMIC_RELAXED);
        if (b >= j->n) break;
        int64_t e = b + j->chunk < j->n ? b + j->chunk : j->n;
        for (int64_t o = b; o < e; ++o) {
            const int64_t* oi = j->indices + o * D;
            float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int64_t k = 0; k < j->klen; ++k) {
                int64_t s[ORC_MAX_DIM];
                for (int d = 0; d < D; ++d) s[d] = oi[d];
                int64_t v = oi[axis] + k - j->radius;
                if (v < lo) v = lo;
                if (v > hi) v = hi;
                s[axis] = v;
                int64_t si = flat_index(D, j->sreg, s);
                for (int c = 0; c < nc; ++c) acc[c] = acc[c] + j->src[si * nc + c] * j->wf[k];
            }
            for (int c = 0; c < nc; ++c) j->out[o * nc + c] = acc[c];
        }
    }
    return NULL;
}

int orc_ref_faithful_f32(int D, int ncomp, const float* img, const int64_t* img_index,
                         const int64_t* img_size, const double* spacing, double sigma,
                         const int64_t* req_index, const int64_t* req_size, float* out,
                         int nthreads) {
    if (D < 1 || D > ORC_MAX_DIM || ncomp < 1 || ncomp > 4 || nthreads < 1) return -1;
    orc_region bounds = region_make(D, img_index, img_size);
    orc_region req = region_make(D, req_index, req_size);
    uint64_t radii[ORC_MAX_DIM] = {0};
    for (int d = 0; d < D; ++d) radii[d] = orc_kernel_radius(sigma, spacing[d]);
    orc_region in_reg = region_clipped(D, region_padded(D, req, radii), bounds);
    orc_region cur_reg;
    float* cur = leaf_copy_f32(D, ncomp, img, bounds, in_reg, &cur_reg);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int axis = 0; axis < D; ++axis) {
        int64_t klen = 2 * (int64_t)radii[axis] + 1;
        double* kern = (double*)malloc((size_t)klen * sizeof(double));
        float* wf = (float*)malloc((size_t)klen * sizeof(float));
        orc_build_kernel(sigma / spacing[axis], radii[axis], kern);
        for (int64_t k = 0; k < klen; ++k) wf[k] = (float)kern[k];
        orc_region out_reg;
        if (axis + 1 < D) {
            uint64_t rem[ORC_MAX_DIM] = {0};
            for (int d = axis + 1; d < D; ++d) rem[d] = radii[d];
            out_reg = region_clipped(D, region_padded(D, req, rem), bounds);
        } else {
            out_reg = req;
        }
        int64_t n = region_len(D, out_reg);
        /* materialised index list, built sequentially (conv.rs:21-22) */
        int64_t* indices = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * D * sizeof(int64_t));
        int64_t idx[ORC_MAX_DIM];
        for (int d = 0; d < D; ++d) idx[d] = out_reg.index[d];
        for (int64_t o = 0; o < n; ++o) {
            for (int d = 0; d < D; ++d) indices[o * D + d] = idx[d];
            for (int d = 0; d < D; ++d) {
                if (++idx[d] < out_reg.index[d] + out_reg.size[d]) break;
                idx[d] = out_reg.index[d];
            }
        }
        float* nxt = (float*)malloc((size_t)(n > 0 ? n : 1) * ncomp * sizeof(float));
        ref_job job = {D, ncomp, axis, cur, cur_reg, wf, klen, (int64_t)radii[axis],
                       indices, n, nxt, 1024, 0};
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, ref_worker, &job);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
        free(indices);
        free(kern);
        free(wf);
        free(cur);
        cur = nxt;
        cur_reg = out_reg;
    }
    free(th);
    memcpy(out, cur, (size_t)region_len(D, req) * ncomp * sizeof(float));
    free(cur);
    return 0;
}

/* ---- cache-friendly CPU separable Gaussian (bench comparison only) ------
 *
 * The same arithmetic as generate_region above for a 3-D f32 scalar image
 * and a request equal to the whole image: per axis, out = sum over k of
 * fl32(x[clamp(i + k - r)] * (float)w_k), added in order k = 0..2r to an
 * accumulator that starts at +0, no FMA, passes in axis order x, y, z.  The
 * loops run along x in every pass (contiguous rows, vectorised by the
 * compiler across x, which keeps each output's operation order), one plane
 * per work item over a pool of threads.  It is the honest CPU comparison a
 * straightforward optimised port gives, beside orc_ref_faithful_f32 (the
 * reference's own loop structure). */

typedef struct {
    const float* src;
    float* dst;
    int64_t nx, ny, nz;
    int axis;
    const float* wf;
    int64_t radius;
    int64_t next; /* atomic plane counter */
} opt_job;

__attribute__((target_clones("avx512f", "avx2", "default")))
static void opt_row_taps(float* acc, const float* row, float w, int64_t n) {
    for (int64_t x = 0; x < n; ++x) acc[x] = acc[x] + row[x] * w;
}

static void* opt_worker(void* arg) {
    opt_job* j = (opt_job*)arg;
    const int64_t nx = j->nx, ny = j->ny, nz = j->nz, r = j->radius, klen = 2 * r + 1;
    float* acc = (float*)malloc((size_t)nx * sizeof(float));
    float* pad = (float*)malloc((size_t)(nx + 2 * r) * sizeof(float));
    for (;;) {
        const int64_t z = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED);
        if (z >= nz) break;
        for (int64_t y = 0; y < ny; ++y) {
            float* out = j->dst + (z * ny + y) * nx;
            for (int64_t x = 0; x < nx; ++x) acc[x] = 0.0f;
            if (j->axis == 0) {
                /* the row with r clamped copies at each end, then shifted rows */
                const float* in = j->src + (z * ny + y) * nx;
                for (int64_t i = 0; i < nx + 2 * r; ++i) {
                    int64_t v = i - r;
                    v = v < 0 ? 0 : (v > nx - 1 ? nx - 1 : v);
                    pad[i] = in[v];
                }
                for (int64_t k = 0; k < klen; ++k) opt_row_taps(acc, pad + k, j->wf[k], nx);
            } else if (j->axis == 1) {
                for (int64_t k = 0; k < klen; ++k) {
                    int64_t v = y + k - r;
                    v = v < 0 ? 0 : (v > ny - 1 ? ny - 1 : v);
                    opt_row_taps(acc, j->src + (z * ny + v) * nx, j->wf[k], nx);
                }
            } else {
                for (int64_t k = 0; k < klen; ++k) {
                    int64_t v = z + k - r;
                    v = v < 0 ? 0 : (v > nz - 1 ? nz - 1 : v);
                    opt_row_taps(acc, j->src + (v * ny + y) * nx, j->wf[k], nx);
                }
            }
            memcpy(out, acc, (size_t)nx * sizeof(float));
        }
    }
    free(pad);
    free(acc);
    return NULL;
}

/* img: nx * ny * nz f32 (axis 0 fastest); out the same size; work: two
 * more images of that size (the x and y pass results), owned by the caller
 * so that repeated calls reuse touched pages. */
int orc_separable_opt_f32(const float* img, int64_t nx, int64_t ny, int64_t nz,
                          const double* spacing, double sigma, float* out, float* work,
                          int nthreads) {
    if (nx < 1 || ny < 1 || nz < 1 || nthreads < 1 || !work) return -1;
    const size_t n = (size_t)(nx * ny * nz);
    float* tmp[2] = {work, work + n};
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    const float* cur = img;
    for (int axis = 0; axis < 3; ++axis) {
        const uint64_t r = orc_kernel_radius(sigma, spacing[axis]);
        const int64_t klen = 2 * (int64_t)r + 1;
        double* kern = (double*)malloc((size_t)klen * sizeof(double));
        float* wf = (float*)malloc((size_t)klen * sizeof(float));
        orc_build_kernel(sigma / spacing[axis], r, kern);
        for (int64_t k = 0; k < klen; ++k) wf[k] = (float)kern[k];
        float* dst = axis == 2 ? out : tmp[axis];
        opt_job job = {cur, dst, nx, ny, nz, axis, wf, (int64_t)r, 0};
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, opt_worker, &job);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
        free(kern);
        free(wf);
        cur = dst;
    }
    free(th);
    return 0;
}
