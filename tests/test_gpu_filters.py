"""MeanFilter, BoxMeanFilterN and BinomialBlurFilter on the GPU, through
itkg_separable_generate_region: EXACT bit-identical to the oracle's
restatement of src/filters/mean.rs:59-83, 126-144 and
src/filters/binomial_blur.rs:57-89 on seeded inputs (whole images,
sub-region requests touching and away from the border, f32 and f64, one and
three components, fused and per-axis kernels); FAST within 1e-5 (max abs
error / max abs value) of EXACT; streamed pieces bit-identical to the whole
image; the reference's own unit tests; argument errors."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(777)


def rand(shape, ncomp=1, dtype=np.float32):
    shp = tuple(shape) + ((ncomp,) if ncomp > 1 else ())
    return RNG.random(shp).astype(dtype)


def _img(itk, x_np, dev, D):
    t = torch.from_numpy(x_np).to(dev)
    size = list(reversed(x_np.shape[:D]))
    return itk.Image(itk.Region([0] * D, size), (1.0,) * D, (0.0,) * D, t)


@pytest.fixture(params=["auto", "per_axis"])
def path(request, itk):
    itk.set_path(itk.PATH_AUTO if request.param == "auto" else itk.PATH_PER_AXIS)
    yield request.param
    itk.set_path(itk.PATH_AUTO)


@pytest.mark.parametrize("size,radii,req,ncomp,dtype", [
    ((64, 48, 40), (2, 2, 2), None, 1, np.float32),
    ((70, 33, 21), (5, 5, 3), ((3, 0, 4), (50, 33, 10)), 1, np.float32),
    ((40, 40, 30), (1, 1, 1), ((0, 10, 0), (40, 12, 30)), 3, np.float32),
    ((96, 80), (7, 7), None, 1, np.float32),
    ((50, 37, 12), (3, 1, 2), None, 1, np.float32),          # anisotropic: per-axis kernels
    ((30, 20, 10), (2, 2, 2), ((4, 4, 2), (20, 10, 6)), 1, np.float64),
    ((200,), (9,), ((20,), (100,)), 2, np.float32),
])
def test_mean_exact_matches_oracle(itk, orc, dev, path, size, radii, req, ncomp, dtype):
    D = len(size)
    x = rand(tuple(reversed(size)), ncomp, dtype)
    src = _img(itk, x, dev, D)
    f = itk.BoxMeanFilterN(src, list(radii)) if len(set(radii)) > 1 else itk.MeanFilter(src, radii[0])
    f.arith = itk.EXACT
    reg = src.region if req is None else itk.Region(list(req[0]), list(req[1]))
    out = f.generate_region(reg)
    torch.cuda.synchronize()
    ref = orc.mean_generate_region(x, [0] * D, list(radii), list(reg.index), list(reg.size), ncomp)
    assert np.array_equal(out.data.cpu().numpy(), ref)


@pytest.mark.parametrize("size,reps,req,ncomp", [
    ((64, 48, 40), 1, None, 1),
    ((50, 40, 30), 3, ((0, 5, 7), (25, 30, 10)), 1),
    ((33, 27, 19), 5, None, 3),
    ((128, 96), 4, ((100, 0), (28, 50)), 1),
    ((64,), 6, None, 1),
    ((20, 16, 12), 0, ((2, 2, 2), (10, 8, 6)), 1),
])
def test_binomial_exact_matches_oracle(itk, orc, dev, path, size, reps, req, ncomp):
    D = len(size)
    x = rand(tuple(reversed(size)), ncomp)
    src = _img(itk, x, dev, D)
    f = itk.BinomialBlurFilter(src, reps)
    f.arith = itk.EXACT
    reg = src.region if req is None else itk.Region(list(req[0]), list(req[1]))
    out = f.generate_region(reg)
    torch.cuda.synchronize()
    assert list(out.region.index) == list(reg.index) and list(out.region.size) == list(reg.size)
    ref = orc.binomial_generate_region(x, [0] * D, reps, list(reg.index), list(reg.size), ncomp)
    assert np.array_equal(out.data.cpu().numpy(), ref)


def test_binomial_generate_into_reuses_scratch_across_streams(itk, orc, dev):
    """generate_into keeps its two round buffers between calls; calls on two
    streams in turn, with sub-regions of different sizes, all match the
    oracle."""
    x = rand((40, 48, 56))
    src = _img(itk, x, dev, 3)
    f = itk.BinomialBlurFilter(src, 3)
    s2 = torch.cuda.Stream()
    for i, req in enumerate([((0, 0, 0), (56, 48, 40)), ((10, 5, 3), (20, 30, 12)),
                             ((0, 0, 0), (56, 48, 40))]):
        reg = itk.Region(list(req[0]), list(req[1]))
        inp = src.generate_region(f.input_region_for_output(reg))
        out = itk.Image.empty(reg, src.spacing, src.origin, 1, torch.float32, dev)
        with torch.cuda.stream(s2 if i % 2 else torch.cuda.current_stream()):
            f.generate_into(inp, out)
        torch.cuda.synchronize()
        ref = orc.binomial_generate_region(x, [0, 0, 0], 3, list(reg.index), list(reg.size))
        assert np.array_equal(out.data.cpu().numpy(), ref), i


@pytest.mark.parametrize("kind,arg", [("mean", 3), ("mean", 6), ("binomial", 3)])
def test_fast_within_tolerance(itk, dev, kind, arg):
    x = torch.rand((96, 112, 128), device=dev)
    f = (lambda m: itk.mean_filter(x, arg, arith=m)) if kind == "mean" else \
        (lambda m: itk.binomial_blur(x, arg, arith=m))
    ex, fa = f(itk.EXACT), f(itk.FAST)
    err = float((fa - ex).abs().max()) / float(ex.abs().max())
    assert err <= 1e-5, err


@pytest.mark.parametrize("kind", ["mean", "binomial"])
def test_streamed_equals_full(itk, dev, kind):
    x = torch.rand((60, 70, 80), device=dev)
    src = itk.Image(itk.Region([0, 0, 0], [80, 70, 60]), (1.0,) * 3, (0.0,) * 3, x)
    f = itk.MeanFilter(src, 4) if kind == "mean" else itk.BinomialBlurFilter(src, 3)
    f.arith = itk.FAST
    full = f.generate_region(src.region)
    streamed = itk.stream_full(f, 7)
    torch.cuda.synchronize()
    assert torch.equal(full.data, streamed.data)


def test_reference_unit_tests(itk, dev):
    # mean.rs:261-294 and binomial_blur.rs:97-128 through the GPU filters
    D1 = lambda t: itk.Image(itk.Region([0], [t.shape[0]]), (1.0,), (0.0,), t.to(dev))  # noqa: E731
    out = itk.MeanFilter(D1(torch.full((10,), 5.0)), 2).generate_region(itk.Region([0], [10]))
    assert bool(((out.data - 5.0).abs() < 1e-5).all())
    ramp = D1(torch.arange(20, dtype=torch.float32))
    out = itk.MeanFilter(ramp, 2).generate_region(ramp.region).data.cpu()
    assert all(abs(float(out[i]) - i) < 1e-4 for i in range(2, 18))
    sq = itk.Image(itk.Region([0, 0], [8, 8]), (1.0, 1.0), (0.0, 0.0), torch.full((8, 8), 3.0, device=dev))
    out = itk.BoxMeanFilterN(sq, [1, 2]).generate_region(sq.region)
    assert bool(((out.data - 3.0).abs() < 1e-5).all())
    c = D1(torch.full((20,), 3.0))
    out = itk.BinomialBlurFilter(c, 4).generate_region(c.region).data.cpu()
    assert all(abs(float(out[i]) - 3.0) < 1e-5 for i in range(4, 16))
    imp = torch.zeros(21)
    imp[10] = 1.0
    im = D1(imp)
    out = itk.BinomialBlurFilter(im, 4).generate_region(im.region).data.cpu()
    assert 0.0 < float(out[10]) < 1.0
    assert abs(float(out.sum()) - 1.0) < 0.05
    assert abs(itk.BinomialBlurFilter.effective_sigma_pixels(4) - 1.0) < 1e-10


def test_input_regions(itk, dev):
    src = itk.Image(itk.Region([0, 0], [30, 20]), (1.0, 1.0), (0.0, 0.0), torch.zeros((20, 30), device=dev))
    req = itk.Region([0, 5], [10, 4])
    assert itk.MeanFilter(src, 3).input_region_for_output(req) == itk.Region([0, 2], [13, 10])
    assert itk.BoxMeanFilterN(src, [1, 6]).input_region_for_output(req) == itk.Region([0, 0], [11, 15])
    assert itk.BinomialBlurFilter(src, 2).input_region_for_output(req) == itk.Region([0, 3], [12, 8])


def test_argument_errors(itk, dev):
    x = torch.rand((10, 12), device=dev)
    src = itk.Image(itk.Region([0, 0], [12, 10]), (1.0, 1.0), (0.0, 0.0), x)
    out = itk.Image.allocate(src.region, src.spacing, src.origin, 0.0, 1, x.dtype, dev)
    # asymmetric weights are refused (the adjoint kernels assume symmetry)
    with pytest.raises(itk.ItkgError) as e:
        itk._separable_into([1, 1], [[0.2, 0.5, 0.3], [0.25, 0.5, 0.25]], src.region, src, out, itk.EXACT)
    assert e.value.code == itk.ERR_INVALID_ARG
    with pytest.raises(itk.ItkgError) as e:
        itk.MeanFilter(src, 128).generate_region(src.region)
    assert e.value.code == itk.ERR_UNSUPPORTED
    with pytest.raises(itk.ItkgError) as e:
        itk.box_kernel(200)
    assert e.value.code == itk.ERR_UNSUPPORTED


# --- UnsharpMaskFilter (src/filters/edges.rs:406-452) -----------------------

@pytest.mark.parametrize("size,sigma,amount,thr,req,dtype", [
    ((64, 48, 40), 1.0, 0.5, 0.0, None, np.float32),
    ((70, 50, 30), 2.0, 1.5, 0.03, ((5, 0, 3), (40, 50, 20)), np.float32),
    ((96, 80), 1.4, 0.5, 0.0, ((0, 10), (96, 30)), np.float32),
    ((200,), 3.0, 0.8, 0.01, None, np.float32),
    ((30, 20, 10), 0.8, 0.5, 0.0, ((4, 4, 2), (20, 10, 6)), np.float64),
    ((20, 18, 16), 0.0, 0.5, 0.0, None, np.float32),    # NaN weights: the identity, as the reference
    ((20, 18, 16), -1.0, 0.5, 0.0, None, np.float32),   # radius 0: the identity
])
def test_unsharp_exact_matches_oracle(itk, orc, dev, path, size, sigma, amount, thr, req, dtype):
    D = len(size)
    x = rand(tuple(reversed(size)), 1, dtype)
    src = _img(itk, x, dev, D)
    f = itk.UnsharpMaskFilter(src, sigma).with_amount(amount)
    f.threshold = thr
    reg = src.region if req is None else itk.Region(list(req[0]), list(req[1]))
    out = f.generate_region(reg)
    torch.cuda.synchronize()
    ref = orc.unsharp_generate_region(x, [0] * D, sigma, amount, thr, list(reg.index), list(reg.size))
    assert np.array_equal(out.data.cpu().numpy(), ref)


@pytest.mark.parametrize("size,sigma,thr,req", [
    ((70, 50, 30), 2.0, 0.03, ((5, 0, 3), (40, 50, 20))),
    ((130, 129, 67), 1.0, 0.0, None),
    ((64, 70, 40), 5.0, 0.01, None),     # radius 15: EXACT as two launches, FAST fused
    ((50, 60, 33), 4.5, 0.0, None),      # radius 14
    ((96, 80), 1.4, 0.02, ((0, 10), (96, 30))),
    ((300, 211), 3.3, 0.0, None),
])
def test_unsharp_fused_step_equals_two_launches(itk, dev, size, sigma, thr, req):
    """f32 volumes and planes: the step in the blur's own launch (the fused
    kernel's unsharp modes) against the blur on the per-axis path followed
    by the per-pixel step kernel, bit for bit in EXACT and in FAST (the FAST
    blur has the same bits on both paths); thresholds that split the pixels,
    sub-requests, radii 3..15."""
    D = len(size)
    x = torch.from_numpy(rand(tuple(reversed(size)), 1, np.float32)).to(dev)
    src = itk.Image(itk.Region([0] * D, list(size)), (1.0,) * D, (0.0,) * D, x)
    reg = src.region if req is None else itk.Region(list(req[0]), list(req[1]))
    for arith in (itk.EXACT, itk.FAST):
        f = itk.UnsharpMaskFilter(src, sigma).with_amount(0.7)
        f.threshold = thr
        f.arith = arith
        fused = f.generate_region(reg).data
        itk.set_path(itk.PATH_PER_AXIS)
        try:
            two = f.generate_region(reg).data
        finally:
            itk.set_path(itk.PATH_AUTO)
        torch.cuda.synchronize()
        assert torch.equal(fused.view(torch.int32), two.view(torch.int32)), arith
        assert float((fused - x[src._slices(reg)]).abs().max()) > 0  # the step changed pixels


def test_unsharp_fast_and_host_entry(itk, dev):
    x = torch.rand((80, 96, 112), device=dev)
    src = itk.Image(itk.Region([0, 0, 0], [112, 96, 80]), (1.0,) * 3, (0.0,) * 3, x)
    ex = itk.UnsharpMaskFilter(src, 1.5)
    fa = itk.UnsharpMaskFilter(src, 1.5)
    fa.arith = itk.FAST
    a, b = ex.generate_region(src.region).data, fa.generate_region(src.region).data
    assert float((a - b).abs().max()) / float(a.abs().max()) <= 1e-5
    # the host-buffer entry point gives the device result bit for bit
    xh = x.cpu().numpy()
    out = np.empty_like(xh)
    reg = src.region.to_c()
    rc = itk.lib.itkg_unsharp_mask_generate_region_host(3, itk.F32, 1.5, 0.5, 0.0, ctypes.byref(reg),
                                                        xh.ctypes.data, ctypes.byref(reg),
                                                        out.ctypes.data, itk.EXACT)
    assert rc == 0, itk.lib.itkg_last_error()
    assert np.array_equal(out, a.cpu().numpy())


def test_unsharp_streamed_and_vector_refused(itk, dev):
    x = torch.rand((40, 50, 60), device=dev)
    src = itk.Image(itk.Region([0, 0, 0], [60, 50, 40]), (1.0,) * 3, (0.0,) * 3, x)
    f = itk.UnsharpMaskFilter(src, 1.2)
    assert torch.equal(f.generate_region(src.region).data, itk.stream_full(f, 5).data)
    v = itk.Image(src.region, src.spacing, src.origin, torch.rand((40, 50, 60, 3), device=dev))
    with pytest.raises(itk.ItkgError) as e:
        itk.UnsharpMaskFilter(v, 1.0).generate_region(v.region)
    assert e.value.code == itk.ERR_UNSUPPORTED


# --- Recursive Gaussian (src/filters/recursive_gaussian.rs) -----------------

@pytest.mark.parametrize("size,spacing,sigma,axis,req,dtype", [
    ((64, 48, 40), (1.0, 1.0, 1.0), 2.0, -1, None, np.float32),
    ((70, 50, 30), (0.7, 0.7, 2.5), 3.0, -1, ((5, 0, 3), (40, 50, 20)), np.float32),
    ((96, 80), (1.0, 1.0), 1.5, 1, None, np.float32),
    ((200,), (1.0,), 0.3, -1, ((20,), (100,)), np.float32),
    ((30, 20, 10), (1.0, 1.0, 1.0), 2.5, 0, ((4, 4, 2), (20, 10, 6)), np.float64),
    ((1, 40, 12), (1.0, 1.0, 1.0), 1.0, -1, None, np.float32),   # an axis of one sample
    # axis-0 rows of 32-sample chunks followed by a 16..31-sample tail, odd
    # lengths (lanes of one wave on different paths), f64 chunked reads of
    # the input, sub-requests; all axes and axis 0 alone
    ((50, 33, 17), (1.0, 1.0, 1.0), 2.0, -1, None, np.float32),
    ((50, 33, 17), (1.0, 1.0, 1.0), 2.0, 0, None, np.float32),
    ((83, 37, 21), (1.0, 0.8, 1.3), 1.7, -1, ((3, 2, 1), (77, 33, 19)), np.float32),
    ((83, 37, 21), (1.0, 1.0, 1.0), 2.4, 0, ((5, 0, 0), (61, 37, 21)), np.float32),
    ((81, 40, 9), (1.0, 1.0, 1.0), 1.3, -1, ((7, 3, 2), (70, 31, 5)), np.float64),
    ((81, 40, 9), (1.0, 1.0, 1.0), 2.2, 0, None, np.float64),
    ((113, 7), (1.0, 1.0), 2.6, -1, None, np.float32),
    ((95, 5, 3), (1.0, 1.0, 1.0), 1.1, 2, ((0, 1, 0), (95, 3, 3)), np.float32),
    # even rows: the in-place passes of the strided axes take two lines a
    # thread (f64 pixels, a 4-D image with two such passes)
    ((64, 40, 12), (1.0, 1.0, 1.0), 1.5, -1, ((2, 3, 1), (60, 35, 10)), np.float64),
    ((20, 9, 7, 5), (1.0, 1.0, 1.0, 1.0), 1.2, -1, None, np.float32),
])
def test_recursive_matches_oracle(itk, orc, dev, size, spacing, sigma, axis, req, dtype):
    D = len(size)
    x = rand(tuple(reversed(size)), 1, dtype)
    src = itk.Image(itk.Region([0] * D, list(size)), spacing, (0.0,) * D, torch.from_numpy(x).to(dev))
    f = itk.RecursiveGaussianFilter(src, sigma) if axis < 0 else \
        itk.RecursiveGaussianAxisFilter(src, sigma, axis)
    reg = src.region if req is None else itk.Region(list(req[0]), list(req[1]))
    out = f.generate_region(reg)
    torch.cuda.synchronize()
    ref = orc.recursive_generate_region(x, [0] * D, list(spacing), sigma, axis, list(reg.index),
                                        list(reg.size))
    assert np.array_equal(out.data.cpu().numpy(), ref)


@pytest.mark.parametrize("width,pitch,axis,dtype,req", [
    (83, 100, -1, np.float32, None),
    (83, 100, 0, np.float32, ((5, 2, 1), (70, 30, 8))),
    (50, 52, -1, np.float32, ((3, 0, 0), (45, 37, 11))),
    (17, 20, 0, np.float32, None),
    (83, 96, -1, np.float64, ((1, 1, 1), (80, 35, 9))),
    (69, 70, 0, np.float64, None),
    (128, 128, -1, np.float32, None),
])
def test_recursive_rows_of_a_window(itk, orc, dev, width, pitch, axis, dtype, req):
    """The input a window of `width` samples per row into rows of `pitch`
    samples (16-byte-aligned rows whose length is not a multiple of 4): axis
    0 goes through the wave-cooperative LDS path, with whole 16-byte vectors
    up to the row end and the last samples one by one; a short last chunk;
    all axes or axis 0 alone (which reads the input and writes the output
    itself); bit for bit the oracle."""
    rng = np.random.default_rng(800 + width + pitch)
    full = rng.random((11, 37, pitch)).astype(dtype)
    x = full[:, :, :width]
    src = itk.Image(itk.Region([0, 0, 0], [width, 37, 11]), (1.0, 1.0, 1.0), (0.0,) * 3,
                    torch.from_numpy(full).to(dev)[:, :, :width])
    f = itk.RecursiveGaussianFilter(src, 1.9) if axis < 0 else \
        itk.RecursiveGaussianAxisFilter(src, 1.9, axis)
    reg = src.region if req is None else itk.Region(list(req[0]), list(req[1]))
    out = f.generate_region(reg)
    torch.cuda.synchronize()
    ref = orc.recursive_generate_region(np.ascontiguousarray(x), [0, 0, 0], [1.0, 1.0, 1.0], 1.9,
                                        axis, list(reg.index), list(reg.size))
    assert np.array_equal(out.data.cpu().numpy(), ref)


def test_recursive_axis_1_beyond_65535(itk, orc, dev):
    """A 16 x 100000 image: the line kernels take 1-D grids, so axis 1 may
    exceed the 65535 rows of a grid dimension."""
    x = rand((100000, 16))
    src = itk.Image(itk.Region([0, 0], [16, 100000]), (1.0, 1.0), (0.0, 0.0), torch.from_numpy(x).to(dev))
    reg = itk.Region([3, 99000], [10, 1000])
    out = itk.RecursiveGaussianFilter(src, 1.5).generate_region(reg).data.cpu().numpy()
    ref = orc.recursive_generate_region(x, [0, 0], [1.0, 1.0], 1.5, -1, [3, 99000], [10, 1000])
    assert np.array_equal(out, ref)


def test_trim_scratch_keeps_results(itk, dev):
    """Trimming the scratch pool between calls changes nothing but the memory
    the pool holds."""
    x = torch.rand((40, 50, 60), device=dev)
    itk.set_path(itk.PATH_PER_AXIS)  # the per-axis path allocates its intermediates
    try:
        a = itk.gaussian(x, 2.0)
        torch.cuda.synchronize()
        itk.trim_scratch(0)
        b = itk.gaussian(x, 2.0)
        torch.cuda.synchronize()
        itk.trim_scratch(1 << 20)
    finally:
        itk.set_path(itk.PATH_AUTO)
    assert torch.equal(a, b)


def test_recursive_reference_tests_and_host_entry(itk, dev):
    # recursive_gaussian.rs:252-300 through the GPU filter
    imp = torch.zeros(51)
    imp[25] = 1.0
    im = itk.Image(itk.Region([0], [51]), (1.0,), (0.0,), imp.to(dev))
    out = itk.RecursiveGaussianFilter(im, 2.0).generate_region(im.region).data.cpu()
    assert abs(float(out.sum()) - 1.0) < 0.02 and float(out[25]) > float(out[28])
    c = itk.Image(itk.Region([0, 0], [15, 15]), (1.0, 1.0), (0.0, 0.0), torch.full((15, 15), 4.0, device=dev))
    out = itk.RecursiveGaussianFilter(c, 1.5).generate_region(c.region).data.cpu()
    assert bool(((out[3:12, 3:12] - 4.0).abs() < 0.05).all())
    # the host-buffer entry point equals the device one
    x = torch.rand((30, 40, 50), device=dev)
    src = itk.Image(itk.Region([0, 0, 0], [50, 40, 30]), (1.0,) * 3, (0.0,) * 3, x)
    want = itk.RecursiveGaussianFilter(src, 2.2).generate_region(itk.Region([5, 0, 3], [30, 40, 20]))
    xh = x.cpu().numpy()
    got = np.empty((20, 40, 30), np.float32)
    req = itk.Region([5, 0, 3], [30, 40, 20]).to_c()
    rc = itk.lib.itkg_recursive_gaussian_generate_region_host(
        3, itk.F32, 2.2, (ctypes.c_double * 3)(1.0, 1.0, 1.0), -1, ctypes.byref(src.region.to_c()),
        xh.ctypes.data, ctypes.byref(req), got.ctypes.data)
    assert rc == 0, itk.lib.itkg_last_error()
    assert np.array_equal(got, want.data.cpu().numpy())
    with pytest.raises(itk.ItkgError):
        itk.RecursiveGaussianAxisFilter(src, 1.0, 3).generate_region(src.region)


# --- CannyEdgeDetectionFilter (src/filters/edges.rs:548-664) ---------------

def _canny_images():
    rng = np.random.default_rng(21)
    yield "noise", rng.random((90, 120)).astype(np.float32), 1.0, 0.02, 0.06
    yy, xx = np.mgrid[0:128, 0:160]
    disc = (((xx - 80.0) ** 2 + (yy - 60.0) ** 2) < 40.0 ** 2).astype(np.float32)
    yield "disc", disc + 0.05 * rng.random(disc.shape).astype(np.float32), 1.5, 0.05, 0.2
    ramp = (xx / 160.0 + 0.3 * np.sin(yy / 6.0)).astype(np.float32)
    yield "waves", ramp, 1.2, 0.01, 0.04
    spiral = np.zeros((200, 200), np.float32)
    t = np.linspace(0, 6 * np.pi, 4000)
    sx = (100 + 2.8 * t * np.cos(t)).astype(int)
    sy = (100 + 2.8 * t * np.sin(t)).astype(int)
    spiral[np.clip(sy, 0, 199), np.clip(sx, 0, 199)] = 1.0   # one long chain for hysteresis
    yield "spiral", spiral, 0.8, 0.02, 0.5
    # a step edge that is strong only at its left end and weak along 1170
    # more pixels: hysteresis must carry it across 38 tiles of 32 pixels
    # (several batches of passes between flag read-backs); weak bars that
    # touch no strong edge stay out, a separate strong block stays in
    step = np.zeros((64, 1216), np.float32)
    step[32:, :] = np.interp(np.arange(1216), [0, 12, 42], [1.0, 1.0, 0.1])[None, :]
    step[8:20, 1100:1110] = 1.0
    for k in range(0, 1000, 80):
        step[4:28, 100 + k:104 + k] += 0.08
    yield "long chain", step, 1.0, 0.01, 0.3


@pytest.mark.parametrize("case", list(range(5)))
def test_canny_matches_oracle(itk, orc, dev, case):
    name, x, sigma, lower, upper = list(_canny_images())[case]
    ny, nx = x.shape
    src = itk.Image(itk.Region([0, 0], [nx, ny]), (1.0, 1.0), (0.0, 0.0), torch.from_numpy(x).to(dev))
    f = itk.CannyEdgeDetectionFilter(src, sigma, lower, upper)
    for req in (src.region, itk.Region([nx // 4, ny // 3], [nx // 2, ny // 2])):
        out = f.generate_region(req)
        torch.cuda.synchronize()
        ref = orc.canny_generate_region(x, [0, 0], sigma, lower, upper, list(req.index), list(req.size))
        assert np.array_equal(out.data.cpu().numpy(), ref), (name, req)
    assert float(f.generate_region(src.region).data.sum()) > 0, name   # the case has edges


@pytest.mark.parametrize("shape,sigma,lower,upper", [
    ((1, 37), 1.0, 0.02, 0.1),
    ((45, 1), 1.0, 0.02, 0.1),
    ((3, 2), 0.7, 0.0, 0.05),
    ((17, 129), 1.3, 0.0, 0.15),     # lower 0: every suppressed pixel (nms 0.0) is weak
    ((70, 65), 1.0, -1.0, 0.2),
])
def test_canny_thin_images_and_low_thresholds(itk, orc, dev, shape, sigma, lower, upper):
    """Images one or a few pixels across (every gradient and suppression
    neighbour clamped, the 64 x 16 tiles mostly outside the image) and lower
    thresholds at or below zero, where suppressed pixels count as weak:
    bit for bit the oracle."""
    rng = np.random.default_rng(sum(shape))
    x = rng.random(shape).astype(np.float32)
    ny, nx = shape
    src = itk.Image(itk.Region([0, 0], [nx, ny]), (1.0, 1.0), (0.0, 0.0), torch.from_numpy(x).to(dev))
    out = itk.CannyEdgeDetectionFilter(src, sigma, lower, upper).generate_region(src.region)
    torch.cuda.synchronize()
    ref = orc.canny_generate_region(x, [0, 0], sigma, lower, upper, [0, 0], [nx, ny])
    assert np.array_equal(out.data.cpu().numpy(), ref), shape


def test_canny_f64_and_errors(itk, orc, dev):
    rng = np.random.default_rng(5)
    x = rng.random((50, 60))
    src = itk.Image(itk.Region([0, 0], [60, 50]), (1.0, 1.0), (0.0, 0.0), torch.from_numpy(x).to(dev))
    out = itk.CannyEdgeDetectionFilter(src, 1.0, 0.02, 0.08).generate_region(src.region)
    ref = orc.canny_generate_region(x, [0, 0], 1.0, 0.02, 0.08, [0, 0], [60, 50])
    assert np.array_equal(out.data.cpu().numpy(), ref)
    v = itk.Image(itk.Region([0, 0, 0], [8, 8, 8]), (1.0,) * 3, (0.0,) * 3, torch.rand((8, 8, 8), device=dev))
    with pytest.raises(itk.ItkgError) as e:
        itk.CannyEdgeDetectionFilter(v, 1.0, 0.1, 0.2).generate_region(v.region)
    assert e.value.code == itk.ERR_UNSUPPORTED
