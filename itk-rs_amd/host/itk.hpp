// itk.hpp -- C++ host side of the drop-in: the itk-rs pipeline protocol with
// GaussianFilter's pixel work done on the GPU through libitkg's C ABI.
//
// Mirrors, name for name:
//   Index / Size / Region            src/image.rs:1-77
//   Image (buffer + leaf source)     src/image.rs:79-174, src/source.rs:32-59
//   ImageSource protocol             src/source.rs:13-29 (duck-typed: any type
//                                    with largest_region / spacing / origin /
//                                    input_region_for_output / generate_region)
//   SourceRef (the blanket &S impl)  src/source.rs:62-75
//   GaussianFilter                   src/filters/gaussian.rs:13-91
//   MeanFilter, BoxMeanFilterN       src/filters/mean.rs:15-145
//   BinomialBlurFilter               src/filters/binomial_blur.rs:17-90
//   UnsharpMaskFilter                src/filters/edges.rs:376-452
//   CannyEdgeDetectionFilter         src/filters/edges.rs:526-664
//   RecursiveGaussian(Axis)Filter    src/filters/recursive_gaussian.rs:129-241
//   stream_full / split_region       src/source.rs:81-97, src/region_splitter.rs:8-39
//   VecPixel                         src/pixel.rs:61-99
//
// Errors: the reference's generate_region has no error channel and panics on
// bad input; here a failing C-ABI call throws itk::Error carrying the
// library's message.  There is no CPU fallback.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "itkg.h"

namespace itk {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != ITKG_OK) throw Error(rc, std::string("itkg: ") + itkg_last_error());
}

template <size_t D>
struct Index {
    std::array<int64_t, D> v{};
    int64_t& operator[](size_t d) { return v[d]; }
    int64_t operator[](size_t d) const { return v[d]; }
    bool operator==(const Index& o) const { return v == o.v; }
};

template <size_t D>
struct Size {
    std::array<uint64_t, D> v{};
    uint64_t operator[](size_t d) const { return v[d]; }
    bool operator==(const Size& o) const { return v == o.v; }
};

template <size_t D>
struct Region {
    Index<D> index;
    Size<D> size;

    static Region make(std::array<int64_t, D> idx, std::array<uint64_t, D> sz) {
        Region r;
        r.index.v = idx;
        r.size.v = sz;
        return r;
    }
    uint64_t linear_len() const {
        uint64_t n = 1;
        for (size_t d = 0; d < D; ++d) n *= size.v[d];
        return n;
    }
    Region padded_per_axis(const std::array<uint64_t, D>& radii) const {
        Region r = *this;
        for (size_t d = 0; d < D; ++d) {
            r.index.v[d] -= (int64_t)radii[d];
            r.size.v[d] += 2 * radii[d];
        }
        return r;
    }
    Region padded(uint64_t radius) const {
        std::array<uint64_t, D> r;
        r.fill(radius);
        return padded_per_axis(r);
    }
    Region clipped_to(const Region& b) const {
        Region r = *this;
        for (size_t d = 0; d < D; ++d) {
            const int64_t se = index.v[d] + (int64_t)size.v[d];
            const int64_t be = b.index.v[d] + (int64_t)b.size.v[d];
            const int64_t s = index.v[d] > b.index.v[d] ? index.v[d] : b.index.v[d];
            const int64_t e = se < be ? se : be;
            r.index.v[d] = s;
            r.size.v[d] = e - s > 0 ? (uint64_t)(e - s) : 0;
        }
        return r;
    }
    bool contains(const Index<D>& idx) const {
        for (size_t d = 0; d < D; ++d) {
            if (idx.v[d] < index.v[d] || idx.v[d] >= index.v[d] + (int64_t)size.v[d]) return false;
        }
        return true;
    }
    bool operator==(const Region& o) const { return index == o.index && size == o.size; }

    itkg_region to_c() const {
        itkg_region r{};
        for (size_t d = 0; d < D; ++d) {
            r.index[d] = index.v[d];
            r.size[d] = (int64_t)size.v[d];
        }
        return r;
    }
    static Region from_c(const itkg_region& c) {
        Region r;
        for (size_t d = 0; d < D; ++d) {
            r.index.v[d] = c.index[d];
            r.size.v[d] = (uint64_t)c.size[d];
        }
        return r;
    }
};

template <class T, size_t N>
struct VecPixel {
    std::array<T, N> v{};
    bool operator==(const VecPixel& o) const { return v == o.v; }
};

// Pixel layout handed to the C ABI: scalar type and component count.
template <class P>
struct PixelTraits;
template <>
struct PixelTraits<float> {
    static constexpr itkg_dtype dtype = ITKG_F32;
    static constexpr int ncomp = 1;
};
template <>
struct PixelTraits<double> {
    static constexpr itkg_dtype dtype = ITKG_F64;
    static constexpr int ncomp = 1;
};
template <class T, size_t N>
struct PixelTraits<VecPixel<T, N>> {
    static_assert(sizeof(VecPixel<T, N>) == N * sizeof(T), "VecPixel must be [T; N]");
    static constexpr itkg_dtype dtype = PixelTraits<T>::dtype;
    static constexpr int ncomp = (int)N;
};

template <class P, size_t D>
struct Image {
    using pixel_type = P;
    static constexpr size_t dim = D;
    Region<D> region;
    std::array<double, D> spacing_{};
    std::array<double, D> origin_{};
    std::vector<P> data;

    static Image allocate(Region<D> r, std::array<double, D> sp, std::array<double, D> org,
                          P fill) {
        Image im;
        im.region = r;
        im.spacing_ = sp;
        im.origin_ = org;
        im.data.assign(r.linear_len(), fill);
        return im;
    }
    size_t flat_index(const Index<D>& idx) const {
        size_t off = 0, stride = 1;
        for (size_t d = 0; d < D; ++d) {
            off += (size_t)(idx.v[d] - region.index.v[d]) * stride;
            stride *= region.size.v[d];
        }
        return off;
    }
    P get_pixel(const Index<D>& idx) const { return data.at(flat_index(idx)); }
    void set_pixel(const Index<D>& idx, P v) { data.at(flat_index(idx)) = v; }
    void copy_from(const Image& src) {
        for_each_index(src.region, [&](const Index<D>& i) {
            data.at(flat_index(i)) = src.data[src.flat_index(i)];
        });
    }

    // ImageSource for Image: the pipeline leaf (src/source.rs:32-59).
    Region<D> largest_region() const { return region; }
    std::array<double, D> spacing() const { return spacing_; }
    std::array<double, D> origin() const { return origin_; }
    Region<D> input_region_for_output(const Region<D>& r) const { return r; }
    Image generate_region(Region<D> requested) const {
        const Region<D> c = requested.clipped_to(region);
        Image out = allocate(c, spacing_, origin_, P{});
        for_each_index(c, [&](const Index<D>& i) { out.data[out.flat_index(i)] = get_pixel(i); });
        return out;
    }

    // iter_region (src/image.rs:143-174): axis 0 fastest.
    template <class F>
    static void for_each_index(const Region<D>& r, F&& f) {
        if (r.linear_len() == 0) return;
        Index<D> idx = r.index;
        for (;;) {
            f(idx);
            size_t d = 0;
            for (; d < D; ++d) {
                if (++idx.v[d] < r.index.v[d] + (int64_t)r.size.v[d]) break;
                idx.v[d] = r.index.v[d];
            }
            if (d == D) return;
        }
    }
};

// The blanket `impl ImageSource for &S`.
template <class S>
struct SourceRef {
    using pixel_type = typename S::pixel_type;
    static constexpr size_t dim = S::dim;
    const S* s;
    explicit SourceRef(const S& src) : s(&src) {}
    auto largest_region() const { return s->largest_region(); }
    auto spacing() const { return s->spacing(); }
    auto origin() const { return s->origin(); }
    auto input_region_for_output(const Region<dim>& r) const { return s->input_region_for_output(r); }
    auto generate_region(Region<dim> r) const { return s->generate_region(r); }
};

inline uint64_t kernel_radius(double sigma, double spacing) {
    return itkg_kernel_radius(sigma, spacing);
}

// GaussianFilter<S> with the reference's two public fields.
template <class S>
struct GaussianFilter {
    using pixel_type = typename S::pixel_type;
    static constexpr size_t dim = S::dim;
    using P = pixel_type;
    static constexpr size_t D = dim;

    S source;
    double sigma;
    // Arithmetic mode for every GaussianFilter (the reference has one mode;
    // EXACT reproduces it bit for bit).
    static inline itkg_arith arith = ITKG_EXACT;

    GaussianFilter(S src, double s) : source(std::move(src)), sigma(s) {}

    Region<D> largest_region() const { return source.largest_region(); }
    std::array<double, D> spacing() const { return source.spacing(); }
    std::array<double, D> origin() const { return source.origin(); }

    std::array<uint64_t, D> radii() const {
        std::array<uint64_t, D> r{};
        const auto sp = spacing();
        for (size_t d = 0; d < D; ++d) r[d] = kernel_radius(sigma, sp[d]);
        return r;
    }

    Region<D> input_region_for_output(const Region<D>& output) const {
        return output.padded_per_axis(radii()).clipped_to(source.largest_region());
    }

    Image<P, D> generate_region(Region<D> requested) const {
        const Region<D> bounds = source.largest_region();
        const Region<D> input_region = requested.padded_per_axis(radii()).clipped_to(bounds);
        const Image<P, D> input = source.generate_region(input_region);
        Image<P, D> out = Image<P, D>::allocate(requested, spacing(), origin(), P{});
        const auto sp = spacing();
        const itkg_region cb = bounds.to_c(), cr = requested.to_c();
        check(itkg_gaussian_generate_region_host(
            (int32_t)D, PixelTraits<P>::dtype, PixelTraits<P>::ncomp, sigma, sp.data(), &cb,
            input.data.data(), &cr, out.data.data(), arith));
        return out;
    }

    // Adjoint of generate_region(grad_output.region) (no counterpart in the
    // reference): the gradient w.r.t. the source's pixels, on
    // input_region_for_output(grad_output.region).
    Image<P, D> backward(const Image<P, D>& grad_output) const {
        Image<P, D> grad_input = Image<P, D>::allocate(
            input_region_for_output(grad_output.region), spacing(), origin(), P{});
        backward_into(grad_output, grad_input, false);
        return grad_input;
    }

    // Writes (accumulate false) or adds (true) that gradient into
    // grad_input, whose region must contain
    // input_region_for_output(grad_output.region); the rest of grad_input
    // is left as it is.  Streamed pieces accumulate into one image.
    void backward_into(const Image<P, D>& grad_output, Image<P, D>& grad_input,
                       bool accumulate) const {
        const auto sp = spacing();
        const itkg_region cb = source.largest_region().to_c(), cr = grad_output.region.to_c(),
                          cg = grad_input.region.to_c();
        check(itkg_gaussian_backward_host((int32_t)D, PixelTraits<P>::dtype,
                                          PixelTraits<P>::ncomp, sigma, sp.data(), &cb,
                                          grad_output.data.data(), &cr, grad_input.data.data(),
                                          &cg, accumulate ? 1 : 0));
    }
};

template <class S>
GaussianFilter<S> make_gaussian(S src, double sigma) {
    return GaussianFilter<S>(std::move(src), sigma);
}

// ---- the other convolve_axis filters (src/filters/mean.rs,
// src/filters/binomial_blur.rs) ------------------------------------------

// Arithmetic mode of MeanFilter, BoxMeanFilterN and BinomialBlurFilter
// (EXACT reproduces the reference bit for bit).
inline itkg_arith separable_arith = ITKG_EXACT;

// box_kernel (src/filters/mean.rs:15-18).
inline std::vector<double> box_kernel(uint64_t radius) {
    std::vector<double> w(2 * radius + 1);
    check(itkg_box_kernel(radius, w.data(), w.size()));
    return w;
}

// One generate_region of a separable clamp-to-edge filter with per-axis
// radii and weights (axis 0's weights first), from `input` (covering the
// padded request) to a new image of `requested`.
template <class P, size_t D>
Image<P, D> separable_generate(const std::array<uint64_t, D>& radii, const std::vector<double>& weights,
                               const Region<D>& bounds, const Image<P, D>& input,
                               const Region<D>& requested, const std::array<double, D>& spacing,
                               const std::array<double, D>& origin, itkg_arith arith) {
    Image<P, D> out = Image<P, D>::allocate(requested, spacing, origin, P{});
    const itkg_region cb = bounds.to_c(), cr = requested.to_c();
    check(itkg_separable_generate_region_host((int32_t)D, PixelTraits<P>::dtype, PixelTraits<P>::ncomp,
                                              radii.data(), weights.data(), &cb, input.data.data(), &cr,
                                              out.data.data(), arith));
    return out;
}

// BoxMeanFilterN<S, D> (src/filters/mean.rs:102-145): box mean with a
// half-width per axis.  MeanFilter below is the same with one radius.
template <class S>
struct BoxMeanFilterN {
    using pixel_type = typename S::pixel_type;
    static constexpr size_t dim = S::dim;
    using P = pixel_type;
    static constexpr size_t D = dim;

    S source;
    std::array<uint64_t, D> radii;

    BoxMeanFilterN(S src, std::array<uint64_t, D> r) : source(std::move(src)), radii(r) {}

    Region<D> largest_region() const { return source.largest_region(); }
    std::array<double, D> spacing() const { return source.spacing(); }
    std::array<double, D> origin() const { return source.origin(); }
    Region<D> input_region_for_output(const Region<D>& output) const {
        return output.padded_per_axis(radii).clipped_to(source.largest_region());
    }
    Image<P, D> generate_region(Region<D> requested) const {
        const Region<D> bounds = source.largest_region();
        const Image<P, D> input = source.generate_region(requested.padded_per_axis(radii).clipped_to(bounds));
        std::vector<double> w;
        for (size_t d = 0; d < D; ++d) {
            const auto k = box_kernel(radii[d]);
            w.insert(w.end(), k.begin(), k.end());
        }
        return separable_generate<P, D>(radii, w, bounds, input, requested, spacing(), origin(),
                                        separable_arith);
    }
};

// MeanFilter<S> (src/filters/mean.rs:28-84) with the reference's two public
// fields.
template <class S>
struct MeanFilter {
    using pixel_type = typename S::pixel_type;
    static constexpr size_t dim = S::dim;
    using P = pixel_type;
    static constexpr size_t D = dim;

    S source;
    uint64_t radius;

    MeanFilter(S src, uint64_t r) : source(std::move(src)), radius(r) {}

    std::array<uint64_t, D> radii() const {
        std::array<uint64_t, D> r{};
        r.fill(radius);
        return r;
    }
    Region<D> largest_region() const { return source.largest_region(); }
    std::array<double, D> spacing() const { return source.spacing(); }
    std::array<double, D> origin() const { return source.origin(); }
    Region<D> input_region_for_output(const Region<D>& output) const {
        return output.padded_per_axis(radii()).clipped_to(source.largest_region());
    }
    Image<P, D> generate_region(Region<D> requested) const {
        const Region<D> bounds = source.largest_region();
        const Image<P, D> input = source.generate_region(requested.padded_per_axis(radii()).clipped_to(bounds));
        std::vector<double> w;
        const auto k = box_kernel(radius);
        for (size_t d = 0; d < D; ++d) w.insert(w.end(), k.begin(), k.end());
        return separable_generate<P, D>(radii(), w, bounds, input, requested, spacing(), origin(),
                                        separable_arith);
    }
};

// BinomialBlurFilter<S> (src/filters/binomial_blur.rs:17-90): `repetitions`
// rounds of [1, 2, 1] / 4 on every axis, each one radius-1 separable
// generate_region whose request is the final request padded by the rounds
// still to come.
template <class S>
struct BinomialBlurFilter {
    using pixel_type = typename S::pixel_type;
    static constexpr size_t dim = S::dim;
    using P = pixel_type;
    static constexpr size_t D = dim;

    S source;
    uint64_t repetitions;

    BinomialBlurFilter(S src, uint64_t n) : source(std::move(src)), repetitions(n) {}

    static double effective_sigma_pixels(uint64_t n) { return std::sqrt((double)n / 4.0); }

    Region<D> largest_region() const { return source.largest_region(); }
    std::array<double, D> spacing() const { return source.spacing(); }
    std::array<double, D> origin() const { return source.origin(); }
    Region<D> input_region_for_output(const Region<D>& output) const {
        return output.padded(repetitions).clipped_to(source.largest_region());
    }
    Image<P, D> generate_region(Region<D> requested) const {
        const Region<D> bounds = source.largest_region();
        Image<P, D> cur = source.generate_region(requested.padded(repetitions).clipped_to(bounds));
        std::array<uint64_t, D> ones{};
        ones.fill(1);
        std::vector<double> w;
        for (size_t d = 0; d < D; ++d) w.insert(w.end(), {0.25, 0.5, 0.25});
        for (uint64_t rep = 0; rep < repetitions; ++rep) {
            const uint64_t remaining = repetitions - rep - 1;
            const Region<D> reg = remaining == 0 ? requested : requested.padded(remaining).clipped_to(bounds);
            cur = separable_generate<P, D>(ones, w, bounds, cur, reg, spacing(), origin(),
                                           separable_arith);
        }
        return cur;
    }
};

// UnsharpMaskFilter<S> (src/filters/edges.rs:376-452) with the reference's
// public fields; scalar f32 / f64 pixels.  The blur's arithmetic is
// separable_arith.
template <class S>
struct UnsharpMaskFilter {
    using pixel_type = typename S::pixel_type;
    static constexpr size_t dim = S::dim;
    using P = pixel_type;
    static constexpr size_t D = dim;
    static_assert(PixelTraits<P>::ncomp == 1, "unsharp mask takes scalar pixels");

    S source;
    double sigma;
    double amount = 0.5;
    double threshold = 0.0;

    UnsharpMaskFilter(S src, double s) : source(std::move(src)), sigma(s) {}
    UnsharpMaskFilter with_amount(double a) && {
        amount = a;
        return std::move(*this);
    }

    Region<D> largest_region() const { return source.largest_region(); }
    std::array<double, D> spacing() const { return source.spacing(); }
    std::array<double, D> origin() const { return source.origin(); }
    Region<D> input_region_for_output(const Region<D>& output) const {
        return output.padded(kernel_radius(sigma, 1.0)).clipped_to(source.largest_region());
    }
    Image<P, D> generate_region(Region<D> requested) const {
        const Region<D> bounds = source.largest_region();
        const Image<P, D> input = source.generate_region(input_region_for_output(requested));
        Image<P, D> out = Image<P, D>::allocate(requested, spacing(), origin(), P{});
        const itkg_region cb = bounds.to_c(), cr = requested.to_c();
        check(itkg_unsharp_mask_generate_region_host((int32_t)D, PixelTraits<P>::dtype, sigma, amount,
                                                     threshold, &cb, input.data.data(), &cr,
                                                     out.data.data(), separable_arith));
        return out;
    }
};

// CannyEdgeDetectionFilter<S> (src/filters/edges.rs:526-664) with the
// reference's public fields; 2-D scalar f32 / f64 pixels; the request as
// 1 / 0.  The smoothing's arithmetic is separable_arith.
template <class S>
struct CannyEdgeDetectionFilter {
    using pixel_type = typename S::pixel_type;
    static constexpr size_t dim = S::dim;
    using P = pixel_type;
    static constexpr size_t D = dim;
    static_assert(D == 2, "Canny takes 2-D images");
    static_assert(PixelTraits<P>::ncomp == 1, "Canny takes scalar pixels");

    S source;
    double sigma;
    double lower_threshold;
    double upper_threshold;

    CannyEdgeDetectionFilter(S src, double s, double lower, double upper)
        : source(std::move(src)), sigma(s), lower_threshold(lower), upper_threshold(upper) {}

    Region<D> largest_region() const { return source.largest_region(); }
    std::array<double, D> spacing() const { return source.spacing(); }
    std::array<double, D> origin() const { return source.origin(); }
    Region<D> input_region_for_output(const Region<D>&) const { return source.largest_region(); }
    Image<P, D> generate_region(Region<D> requested) const {
        const Region<D> bounds = source.largest_region();
        const Image<P, D> input = source.generate_region(bounds);
        Image<P, D> out = Image<P, D>::allocate(requested, spacing(), origin(), P{});
        const itkg_region cb = bounds.to_c(), cr = requested.to_c();
        check(itkg_canny_generate_region_host((int32_t)D, PixelTraits<P>::dtype, sigma,
                                              lower_threshold, upper_threshold, &cb,
                                              input.data.data(), &cr, out.data.data(),
                                              separable_arith));
        return out;
    }
};

// RecursiveGaussianFilter<S> (src/filters/recursive_gaussian.rs:129-177;
// axis = -1) and RecursiveGaussianAxisFilter<S> (recursive_gaussian.rs:
// 181-241; the `axis` field): the Young - van Vliet recursive Gaussian in
// f64 over the whole image; scalar f32 / f64 pixels.
template <class S>
struct RecursiveGaussianFilter {
    using pixel_type = typename S::pixel_type;
    static constexpr size_t dim = S::dim;
    using P = pixel_type;
    static constexpr size_t D = dim;
    static_assert(PixelTraits<P>::ncomp == 1, "the recursive Gaussian takes scalar pixels");

    S source;
    double sigma;
    int axis = -1;

    RecursiveGaussianFilter(S src, double s, int ax = -1) : source(std::move(src)), sigma(s), axis(ax) {}

    Region<D> largest_region() const { return source.largest_region(); }
    std::array<double, D> spacing() const { return source.spacing(); }
    std::array<double, D> origin() const { return source.origin(); }
    Region<D> input_region_for_output(const Region<D>&) const { return source.largest_region(); }
    Image<P, D> generate_region(Region<D> requested) const {
        const Region<D> bounds = source.largest_region();
        const Image<P, D> input = source.generate_region(bounds);
        Image<P, D> out = Image<P, D>::allocate(requested, spacing(), origin(), P{});
        const itkg_region cb = bounds.to_c(), cr = requested.to_c();
        const auto sp = spacing();
        check(itkg_recursive_gaussian_generate_region_host((int32_t)D, PixelTraits<P>::dtype, sigma,
                                                           sp.data(), axis, &cb, input.data.data(),
                                                           &cr, out.data.data()));
        return out;
    }
};

template <size_t D>
std::vector<Region<D>> split_region(const Region<D>& region, uint64_t num_pieces) {
    std::vector<itkg_region> buf(num_pieces < 1 ? 1 : num_pieces);
    uint64_t n = 0;
    const itkg_region c = region.to_c();
    check(itkg_split_region((int32_t)D, &c, num_pieces, buf.data(), buf.size(), &n));
    std::vector<Region<D>> out;
    for (uint64_t i = 0; i < n; ++i) out.push_back(Region<D>::from_c(buf[i]));
    return out;
}

template <class S>
Image<typename S::pixel_type, S::dim> stream_full(const S& source, uint64_t num_pieces) {
    using P = typename S::pixel_type;
    constexpr size_t D = S::dim;
    const Region<D> full = source.largest_region();
    Image<P, D> output = Image<P, D>::allocate(full, source.spacing(), source.origin(), P{});
    for (const Region<D>& piece : split_region(full, num_pieces)) {
        const Image<P, D> piece_data = source.generate_region(piece);
        output.copy_from(piece_data);
    }
    return output;
}

// UnaryFilter (src/filters/mod.rs:47-80): per-pixel map of an upstream, on
// the host.  Only here so that pipelines of the reference's tests compose.
template <class S, class F, class Q>
struct UnaryFilter {
    using pixel_type = Q;
    static constexpr size_t dim = S::dim;
    S source;
    F f;
    Region<dim> largest_region() const { return source.largest_region(); }
    std::array<double, dim> spacing() const { return source.spacing(); }
    std::array<double, dim> origin() const { return source.origin(); }
    Region<dim> input_region_for_output(const Region<dim>& r) const { return r; }
    Image<Q, dim> generate_region(Region<dim> requested) const {
        const auto in = source.generate_region(requested);
        Image<Q, dim> out = Image<Q, dim>::allocate(in.region, in.spacing_, in.origin_, Q{});
        for (size_t i = 0; i < in.data.size(); ++i) out.data[i] = f(in.data[i]);
        return out;
    }
};

template <class Q, class S, class F>
UnaryFilter<S, F, Q> make_unary(S src, F f) {
    return UnaryFilter<S, F, Q>{std::move(src), std::move(f)};
}

// ---- the casts around a Gaussian (examples/blur_png.rs:35-40) ------------

template <class T>
struct PTypeOf;
template <>
struct PTypeOf<uint8_t> {
    static constexpr itkg_ptype value = ITKG_PT_U8;
};
template <>
struct PTypeOf<uint16_t> {
    static constexpr itkg_ptype value = ITKG_PT_U16;
};
template <>
struct PTypeOf<int16_t> {
    static constexpr itkg_ptype value = ITKG_PT_I16;
};
template <>
struct PTypeOf<int32_t> {
    static constexpr itkg_ptype value = ITKG_PT_I32;
};
template <>
struct PTypeOf<float> {
    static constexpr itkg_ptype value = ITKG_PT_F32;
};
template <>
struct PTypeOf<double> {
    static constexpr itkg_ptype value = ITKG_PT_F64;
};

// UnaryFilter(p as f32) -> GaussianFilter{sigma} -> UnaryFilter(cast to Q)
// as one node on host images (itkg_cast_gaussian_cast_host): the pixels
// cross PCIe in their own types and the f32 intermediates stay on the GPU.
// out_mode ITKG_CONVERT_ROUND is `p.round() ... as Q` (for u8 the
// reference's clamp(0, 255) changes nothing the saturating cast does not),
// ITKG_CONVERT_AS is `p as Q`.
template <class Q, class S>
struct CastGaussianCast {
    using A = typename S::pixel_type;
    using pixel_type = Q;
    static constexpr size_t dim = S::dim;
    static constexpr size_t D = dim;
    S source;
    double sigma;
    itkg_convert_mode out_mode = ITKG_CONVERT_ROUND;
    // Arithmetic of the Gaussian inside (EXACT reproduces the reference).
    static inline itkg_arith arith = ITKG_EXACT;

    Region<D> largest_region() const { return source.largest_region(); }
    std::array<double, D> spacing() const { return source.spacing(); }
    std::array<double, D> origin() const { return source.origin(); }
    Region<D> input_region_for_output(const Region<D>& output) const {
        std::array<uint64_t, D> r{};
        const auto sp = spacing();
        for (size_t d = 0; d < D; ++d) r[d] = kernel_radius(sigma, sp[d]);
        return output.padded_per_axis(r).clipped_to(source.largest_region());
    }
    Image<Q, D> generate_region(Region<D> requested) const {
        const Image<A, D> input = source.generate_region(input_region_for_output(requested));
        Image<Q, D> out = Image<Q, D>::allocate(requested, spacing(), origin(), Q{});
        const auto sp = spacing();
        const itkg_region cb = source.largest_region().to_c(), cr = requested.to_c();
        check(itkg_cast_gaussian_cast_host((int32_t)D, PTypeOf<A>::value, PTypeOf<Q>::value, sigma,
                                           sp.data(), &cb, input.data.data(), &cr, out.data.data(),
                                           arith, out_mode));
        return out;
    }
};

template <class Q, class S>
CastGaussianCast<Q, S> make_cast_gaussian_cast(S src, double sigma,
                                               itkg_convert_mode out_mode = ITKG_CONVERT_ROUND) {
    return CastGaussianCast<Q, S>{std::move(src), sigma, out_mode};
}

}  // namespace itk
