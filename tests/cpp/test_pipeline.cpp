// The reference's pipeline tests (tests/pipeline.rs:19-153), the Gaussian
// unit tests (src/filters/gaussian.rs:107-127) and the MeanFilter and
// BinomialBlurFilter unit tests (src/filters/mean.rs:261-294,
// src/filters/binomial_blur.rs:97-128), written against the C++ host
// API (itk-rs_amd/host/itk.hpp), whose GaussianFilter runs on the GPU.
// Prints one "ok <name>" line per test; exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../itk-rs_amd/host/itk.hpp"

using namespace itk;

#define CHECK(cond, ...)                                           \
    do {                                                           \
        if (!(cond)) {                                             \
            std::printf("FAIL %s:%d: ", __func__, __LINE__);       \
            std::printf(__VA_ARGS__);                              \
            std::printf("\n");                                     \
            std::exit(1);                                          \
        }                                                          \
    } while (0)

template <size_t D>
Image<float, D> solid_image(std::array<uint64_t, D> size, float fill) {
    std::array<int64_t, D> z{};
    std::array<double, D> one, zero{};
    one.fill(1.0);
    return Image<float, D>::allocate(Region<D>::make(z, size), one, zero, fill);
}

static Image<float, 2> impulse_image_2d(uint64_t size) {
    const int64_t mid = (int64_t)(size / 2);
    auto img = solid_image<2>({size, size}, 0.0f);
    img.set_pixel(Index<2>{{mid, mid}}, 1.0f);
    return img;
}

static float sum(const std::vector<float>& v) {
    float s = 0.0f;
    for (float x : v) s += x;
    return s;
}

// src/filters/gaussian.rs:107-114
static void test_gaussian_impulse_energy_conserved() {
    auto filter = make_gaussian(impulse_image_2d(21), 1.5);
    auto out = filter.generate_region(filter.largest_region());
    const float s = sum(out.data);
    CHECK(std::fabs(s - 1.0f) < 0.01f, "energy not conserved: sum=%g", s);
}

// src/filters/gaussian.rs:116-127
static void test_gaussian_peak_at_centre() {
    auto filter = make_gaussian(impulse_image_2d(21), 1.5);
    auto out = filter.generate_region(filter.largest_region());
    const float peak = out.get_pixel(Index<2>{{10, 10}});
    for (float v : out.data) CHECK(v <= peak + 1e-6f, "pixel %g > peak %g", v, peak);
}

// tests/pipeline.rs:19-31
static void test_image_source_identity_2d() {
    auto img = solid_image<2>({8, 8}, 0.0f);
    img.set_pixel(Index<2>{{3, 5}}, 42.0f);
    const auto sub = Region<2>::make({2, 4}, {4, 3});
    auto cropped = img.generate_region(sub);
    CHECK(cropped.region == sub, "region");
    CHECK(cropped.get_pixel(Index<2>{{3, 5}}) == 42.0f, "marked pixel");
    CHECK(cropped.get_pixel(Index<2>{{2, 4}}) == 0.0f, "corner");
}

// tests/pipeline.rs:37-54
static void test_gaussian_impulse_2d() {
    auto filter = make_gaussian(impulse_image_2d(21), 1.5);
    auto out = filter.generate_region(filter.largest_region());
    const float s = sum(out.data);
    CHECK(std::fabs(s - 1.0f) < 0.01f, "energy not conserved: sum=%g", s);
    const float peak = out.get_pixel(Index<2>{{10, 10}});
    for (float v : out.data) CHECK(v <= peak + 1e-6f, "pixel %g > peak %g", v, peak);
}

// tests/pipeline.rs:60-72
static void test_gaussian_3d() {
    auto img = solid_image<3>({11, 11, 11}, 0.0f);
    img.set_pixel(Index<3>{{5, 5, 5}}, 1.0f);
    auto filter = make_gaussian(img, 1.0);
    auto out = filter.generate_region(filter.largest_region());
    const float s = sum(out.data);
    CHECK(std::fabs(s - 1.0f) < 0.05f, "3D energy not conserved: sum=%g", s);
}

// tests/pipeline.rs:78-100
static void test_vector_pixel_gaussian() {
    using V = VecPixel<float, 3>;
    auto img = Image<V, 2>::allocate(Region<2>::make({0, 0}, {11, 11}), {1.0, 1.0}, {0.0, 0.0}, V{});
    img.set_pixel(Index<2>{{5, 5}}, V{{1.0f, 2.0f, 3.0f}});
    auto filter = make_gaussian(img, 1.0);
    auto out = filter.generate_region(filter.largest_region());
    float s[3] = {0, 0, 0};
    for (const V& p : out.data)
        for (int c = 0; c < 3; ++c) s[c] += p.v[c];
    CHECK(std::fabs(s[0] - 1.0f) < 0.05f, "ch0 sum=%g", s[0]);
    CHECK(std::fabs(s[1] - 2.0f) < 0.05f, "ch1 sum=%g", s[1]);
    CHECK(std::fabs(s[2] - 3.0f) < 0.05f, "ch2 sum=%g", s[2]);
}

// tests/pipeline.rs:106-129 (the values are bit-identical, which the
// tolerance of the original also accepts)
static void test_streaming_matches_full() {
    auto img = solid_image<2>({20, 20}, 0.0f);
    img.set_pixel(Index<2>{{10, 10}}, 1.0f);
    auto filter = make_gaussian(img, 1.5);
    auto full = filter.generate_region(filter.largest_region());
    auto filter2 = make_gaussian(img, 1.5);
    auto streamed = stream_full(SourceRef<decltype(filter2)>(filter2), 4);
    CHECK(full.data.size() == streamed.data.size(), "size");
    for (size_t i = 0; i < full.data.size(); ++i) {
        CHECK(std::fabs(full.data[i] - streamed.data[i]) < 1e-6f, "mismatch %g vs %g",
              full.data[i], streamed.data[i]);
        CHECK(full.data[i] == streamed.data[i], "not bit-identical at %zu", i);
    }
}

// tests/pipeline.rs:135-153
static void test_unary_filter_pipeline() {
    auto gaussian = make_gaussian(impulse_image_2d(21), 1.5);
    auto threshold = make_unary<float>(gaussian, [](float p) { return p > 0.01f ? 1.0f : 0.0f; });
    auto out = threshold.generate_region(threshold.largest_region());
    size_t lit = 0;
    for (float v : out.data) lit += v > 0.5f;
    CHECK(lit > 0, "no pixels above threshold");
    CHECK(lit < 21 * 21, "all pixels above threshold");
}

// A request inside the image (a sub-region) equals the same crop of the
// full-image output: the halo is pulled from upstream, the clamp only acts at
// the true image bounds.
static void test_subregion_matches_full_crop() {
    auto img = solid_image<3>({19, 17, 23}, 0.0f);
    uint32_t s = 12345;
    for (float& v : img.data) {
        s = s * 1664525u + 1013904223u;
        v = (float)(s >> 8) / 16777216.0f;
    }
    auto filter = make_gaussian(img, 1.3);
    auto full = filter.generate_region(filter.largest_region());
    const auto sub = Region<3>::make({3, 2, 5}, {9, 11, 7});
    auto part = filter.generate_region(sub);
    Image<float, 3>::for_each_index(sub, [&](const Index<3>& i) {
        CHECK(part.get_pixel(i) == full.get_pixel(i), "sub-region voxel differs");
    });
}

static void test_error_is_thrown() {
    auto img = solid_image<2>({8, 8}, 0.0f);
    auto filter = make_gaussian(img, 1.0);
    bool thrown = false;
    try {
        filter.generate_region(Region<2>::make({4, 4}, {8, 8}));  // outside the image
    } catch (const Error& e) {
        thrown = e.code == ITKG_ERR_INVALID_ARG;
    }
    CHECK(thrown, "request outside the largest region must raise");
}

// src/filters/mean.rs:261-269
static void test_mean_constant_image() {
    MeanFilter<Image<float, 1>> f(solid_image<1>({10}, 5.0f), 2);
    auto out = f.generate_region(f.largest_region());
    for (float v : out.data) CHECK(std::fabs(v - 5.0f) < 1e-5f, "expected 5.0, got %g", v);
}

// src/filters/mean.rs:271-282
static void test_mean_ramp_interior() {
    auto img = solid_image<1>({20}, 0.0f);
    for (int64_t i = 0; i < 20; ++i) img.set_pixel(Index<1>{{i}}, (float)i);
    MeanFilter<Image<float, 1>> f(img, 2);
    auto out = f.generate_region(f.largest_region());
    for (int64_t i = 2; i < 18; ++i) {
        const float v = out.get_pixel(Index<1>{{i}});
        CHECK(std::fabs(v - (float)i) < 1e-4f, "at %lld: got %g", (long long)i, v);
    }
}

// src/filters/mean.rs:284-294
static void test_box_mean_2d_constant() {
    BoxMeanFilterN<Image<float, 2>> f(solid_image<2>({8, 8}, 3.0f), {1, 2});
    auto out = f.generate_region(f.largest_region());
    for (float v : out.data) CHECK(std::fabs(v - 3.0f) < 1e-5f, "expected 3.0, got %g", v);
}

// src/filters/binomial_blur.rs:97-107
static void test_binomial_constant_image_preserved() {
    BinomialBlurFilter<Image<float, 1>> f(solid_image<1>({20}, 3.0f), 4);
    auto out = f.generate_region(f.largest_region());
    for (int64_t i = 4; i < 16; ++i) {
        const float v = out.get_pixel(Index<1>{{i}});
        CHECK(std::fabs(v - 3.0f) < 1e-5f, "at %lld: %g", (long long)i, v);
    }
}

// src/filters/binomial_blur.rs:109-122
static void test_binomial_impulse_spreads() {
    auto img = solid_image<1>({21}, 0.0f);
    img.set_pixel(Index<1>{{10}}, 1.0f);
    BinomialBlurFilter<Image<float, 1>> f(img, 4);
    auto out = f.generate_region(f.largest_region());
    const float peak = out.get_pixel(Index<1>{{10}});
    CHECK(peak > 0.0f && peak < 1.0f, "peak=%g", peak);
    const float s = sum(out.data);
    CHECK(std::fabs(s - 1.0f) < 0.05f, "sum=%g", s);
}

// src/filters/binomial_blur.rs:124-128
static void test_binomial_effective_sigma() {
    const double s = BinomialBlurFilter<Image<float, 1>>::effective_sigma_pixels(4);
    CHECK(std::fabs(s - 1.0) < 1e-10, "sigma=%g", s);
}

// UnsharpMaskFilter (src/filters/edges.rs:406-452): amount 0 leaves the
// image as it is; on an impulse the centre is sharpened above its value and
// its neighbours go below zero.
static void test_unsharp_mask() {
    auto img = solid_image<2>({15, 15}, 0.0f);
    img.set_pixel(Index<2>{{7, 7}}, 1.0f);
    auto same = UnsharpMaskFilter<Image<float, 2>>(img, 1.0).with_amount(0.0);
    auto out0 = same.generate_region(same.largest_region());
    for (size_t i = 0; i < out0.data.size(); ++i)
        CHECK(out0.data[i] == img.data[i], "amount 0 changed pixel %zu", i);
    UnsharpMaskFilter<Image<float, 2>> f(img, 1.0);
    auto out = f.generate_region(f.largest_region());
    CHECK(out.get_pixel(Index<2>{{7, 7}}) > 1.0f, "centre %g", out.get_pixel(Index<2>{{7, 7}}));
    CHECK(out.get_pixel(Index<2>{{8, 7}}) < 0.0f, "neighbour %g", out.get_pixel(Index<2>{{8, 7}}));
}

// CannyEdgeDetectionFilter (src/filters/edges.rs:526-664): a vertical step
// between x = 19 and 20 gives edges only in those two columns, covering a
// whole column; a constant image gives none; a sub-request is a crop.
static void test_canny() {
    auto img = solid_image<2>({40, 30}, 0.0f);
    for (int64_t y = 0; y < 30; ++y)
        for (int64_t x = 20; x < 40; ++x) img.set_pixel(Index<2>{{x, y}}, 1.0f);
    CannyEdgeDetectionFilter<Image<float, 2>> f(img, 1.0, 0.1, 0.3);
    auto out = f.generate_region(f.largest_region());
    bool full_column = false;
    for (int64_t x = 0; x < 40; ++x) {
        int64_t on = 0;
        for (int64_t y = 0; y < 30; ++y) on += out.get_pixel(Index<2>{{x, y}}) == 1.0f;
        CHECK(on == 0 || x == 19 || x == 20, "edge pixels in column %lld", (long long)x);
        full_column = full_column || on == 30;
    }
    CHECK(full_column, "no whole edge column");
    const Region<2> sub = Region<2>::make({15, 5}, {10, 12});
    auto part = f.generate_region(sub);
    for (int64_t y = 5; y < 17; ++y)
        for (int64_t x = 15; x < 25; ++x)
            CHECK(part.get_pixel(Index<2>{{x, y}}) == out.get_pixel(Index<2>{{x, y}}),
                  "sub-request differs at %lld,%lld", (long long)x, (long long)y);
    CannyEdgeDetectionFilter<Image<float, 2>> c(solid_image<2>({40, 30}, 3.0f), 1.0, 0.1, 0.3);
    CHECK(sum(c.generate_region(c.largest_region()).data) == 0.0f, "edges on a constant image");
}

// src/filters/recursive_gaussian.rs:252-300
static void test_recursive_gaussian() {
    auto imp = solid_image<1>({51}, 0.0f);
    imp.set_pixel(Index<1>{{25}}, 1.0f);
    RecursiveGaussianFilter<Image<float, 1>> f(imp, 2.0);
    auto out = f.generate_region(f.largest_region());
    CHECK(std::fabs(sum(out.data) - 1.0f) < 0.02f, "energy=%g", sum(out.data));
    CHECK(out.get_pixel(Index<1>{{25}}) > out.get_pixel(Index<1>{{28}}), "peak not at the centre");
    RecursiveGaussianFilter<Image<float, 2>> c(solid_image<2>({15, 15}, 4.0f), 1.5);
    auto oc = c.generate_region(c.largest_region());
    for (int64_t i = 3; i < 12; ++i)
        for (int64_t j = 3; j < 12; ++j) {
            const float v = oc.get_pixel(Index<2>{{i, j}});
            CHECK(std::fabs(v - 4.0f) < 0.05f, "at [%lld,%lld]: %g", (long long)i, (long long)j, v);
        }
    auto imp31 = solid_image<1>({31}, 0.0f);
    imp31.set_pixel(Index<1>{{15}}, 1.0f);
    RecursiveGaussianFilter<Image<float, 1>> a(imp31, 2.0, 0);
    auto oa = a.generate_region(a.largest_region());
    CHECK(std::fabs(sum(oa.data) - 1.0f) < 0.02f, "single axis energy=%g", sum(oa.data));
}

// GaussianFilter::backward is the transpose of generate_region: column j of
// the forward operator is the forward response to the impulse at source
// pixel j, so (G^T y)_j = <G e_j, y>, built here from forward calls only.
template <class T>
static void check_backward_transpose(double sigma, Region<2> req) {
    const std::array<uint64_t, 2> size{11, 9};
    std::array<double, 2> one{1.0, 1.0}, zero{};
    const auto largest = Region<2>::make({0, 0}, size);
    Image<T, 2> y = Image<T, 2>::allocate(req, one, zero, T{});
    uint32_t st = 12345u;
    for (auto& v : y.data) {
        st = st * 1664525u + 1013904223u;
        v = (T)((st >> 8) * (1.0 / 16777216.0) - 0.5);
    }
    auto src = Image<T, 2>::allocate(largest, one, zero, T{});
    GaussianFilter<Image<T, 2>> f(src, sigma);
    const Image<T, 2> g = f.backward(y);
    CHECK(g.region == f.input_region_for_output(req), "gradient region");
    Image<T, 2>::for_each_index(largest, [&](const Index<2>& j) {
        Image<T, 2> e = Image<T, 2>::allocate(largest, one, zero, T{});
        e.set_pixel(j, (T)1);
        GaussianFilter<Image<T, 2>> fe(e, sigma);
        const Image<T, 2> col = fe.generate_region(req);
        double want = 0.0, mag = 0.0;
        for (size_t i = 0; i < col.data.size(); ++i) {
            want += (double)col.data[i] * (double)y.data[i];
            mag += std::fabs((double)col.data[i] * (double)y.data[i]);
        }
        const bool inside = g.region.contains(j);
        const double got = inside ? (double)g.get_pixel(j) : 0.0;
        const double tol = (sizeof(T) == 4 ? 2e-6 : 1e-14) * std::max(1.0, mag);
        CHECK(std::fabs(got - want) <= tol, "(%lld,%lld): backward %.9g vs transpose %.9g",
              (long long)j.v[0], (long long)j.v[1], got, want);
    });
}

static void test_gaussian_backward_transpose() {
    check_backward_transpose<float>(1.3, Region<2>::make({2, 1}, {6, 5}));
    check_backward_transpose<float>(0.8, Region<2>::make({0, 0}, {11, 9}));
    check_backward_transpose<double>(1.6, Region<2>::make({5, 3}, {6, 6}));
}

// backward_into with accumulate adds into a larger image and leaves the
// rest of it as it was
static void test_gaussian_backward_accumulate() {
    std::array<double, 2> one{1.0, 1.0}, zero{};
    const auto largest = Region<2>::make({0, 0}, {40, 30});
    auto src = Image<float, 2>::allocate(largest, one, zero, 0.0f);
    GaussianFilter<Image<float, 2>> f(src, 1.5);
    const auto req = Region<2>::make({10, 8}, {12, 9});
    Image<float, 2> y = Image<float, 2>::allocate(req, one, zero, 0.5f);
    const Image<float, 2> g = f.backward(y);
    Image<float, 2> acc = Image<float, 2>::allocate(largest, one, zero, 2.0f);
    f.backward_into(y, acc, true);
    f.backward_into(y, acc, true);
    Image<float, 2>::for_each_index(largest, [&](const Index<2>& i) {
        const float want = g.region.contains(i) ? 2.0f + 2.0f * g.get_pixel(i) : 2.0f;
        CHECK(std::fabs(acc.get_pixel(i) - want) <= 1e-6f, "(%lld,%lld): %g vs %g",
              (long long)i.v[0], (long long)i.v[1], acc.get_pixel(i), want);
    });
    // a constant gradient spreads with unit mass: the adjoint of a
    // normalised clamp-to-edge blur preserves the sum
    double s = 0.0;
    for (float v : g.data) s += v;
    CHECK(std::fabs(s - 0.5 * 12 * 9) < 1e-3, "gradient mass %g", s);
}

// examples/blur_png.rs:35-40: u8 -> f32, GaussianFilter, round and clamp
// back to u8.  The one-node GPU chain equals the three-node chain (CPU casts
// around the GPU Gaussian) bit for bit, on the whole image and on a
// sub-request.
static void test_blur_png_chain() {
    std::array<double, 2> one{1.0, 1.0}, zero{};
    const auto largest = Region<2>::make({0, 0}, {67, 45});
    auto img = Image<uint8_t, 2>::allocate(largest, one, zero, 0);
    uint32_t st = 7u;
    for (auto& v : img.data) {
        st = st * 1664525u + 1013904223u;
        v = (uint8_t)(st >> 24);
    }
    img.set_pixel(Index<2>{{20, 10}}, 255);
    auto as_f32 = make_unary<float>(img, [](uint8_t p) { return (float)p; });
    auto blurred = make_gaussian(as_f32, 1.7);
    auto as_u8 = make_unary<uint8_t>(blurred, [](float p) {
        return (uint8_t)std::fmin(std::fmax(std::round(p), 0.0f), 255.0f);
    });
    auto fused = make_cast_gaussian_cast<uint8_t>(img, 1.7);
    for (const auto& req : {largest, Region<2>::make({9, 4}, {30, 21})}) {
        const auto want = as_u8.generate_region(req);
        const auto got = fused.generate_region(req);
        CHECK(got.region == want.region, "region");
        for (size_t i = 0; i < got.data.size(); ++i)
            CHECK(got.data[i] == want.data[i], "pixel %zu: %d vs %d", i, got.data[i], want.data[i]);
    }
}

int main() {
#define RUN(t)                  \
    t();                        \
    std::printf("ok %s\n", #t); \
    std::fflush(stdout);
    RUN(test_image_source_identity_2d);
    RUN(test_gaussian_impulse_energy_conserved);
    RUN(test_gaussian_peak_at_centre);
    RUN(test_gaussian_impulse_2d);
    RUN(test_gaussian_3d);
    RUN(test_vector_pixel_gaussian);
    RUN(test_streaming_matches_full);
    RUN(test_unary_filter_pipeline);
    RUN(test_subregion_matches_full_crop);
    RUN(test_error_is_thrown);
    RUN(test_mean_constant_image);
    RUN(test_mean_ramp_interior);
    RUN(test_box_mean_2d_constant);
    RUN(test_binomial_constant_image_preserved);
    RUN(test_binomial_impulse_spreads);
    RUN(test_binomial_effective_sigma);
    RUN(test_unsharp_mask);
    RUN(test_canny);
    RUN(test_recursive_gaussian);
    RUN(test_gaussian_backward_transpose);
    RUN(test_gaussian_backward_accumulate);
    RUN(test_blur_png_chain);
    std::printf("all passed\n");
    return 0;
}
