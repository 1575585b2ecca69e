"""Seeded random configurations through the GPU paths, each checked
against the oracle: EXACT bit for bit, FAST within 1e-5 (max abs error /
max abs value).  Shapes 1- to 3-D with odd extents, sigmas from below one
pixel to radii past the fused set, anisotropic spacings, sub-region
requests (inside and touching the border), 1 or 3 components, f32 and f64,
and both kernel paths; then MeanFilter / BinomialBlurFilter the same way;
the adjoint against the oracle's f64 transpose; random batched plans
against plain launches, forward and adjoint; the host-buffer entry point
at staged sizes (pinned or pageable) against the oracle; random casts over
strided windows against the oracle's Rust casts."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

FAST_TOL = 1e-5


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    D = int(rng.integers(1, 4))
    size = [int(rng.integers(1, 70 if D == 3 else (300 if D == 2 else 3000))) for _ in range(D)]
    sigma = float(rng.choice([0.3, 0.8, 1.3, 2.0, 3.1, 5.2, 6.5]))
    spacing = [float(rng.choice([0.5, 0.7, 1.0, 1.0, 1.0, 2.5])) for _ in range(D)]
    if rng.random() < 0.5:
        spacing = [spacing[0]] * D
    lo = [int(rng.integers(0, s)) for s in size]
    req = ([l for l in lo], [int(rng.integers(1, s - l + 1)) for s, l in zip(size, lo)])
    if rng.random() < 0.3:
        req = ([0] * D, list(size))
    ncomp = 3 if rng.random() < 0.25 else 1
    dtype = np.float64 if rng.random() < 0.15 else np.float32
    return D, size, sigma, spacing, req, ncomp, dtype


@pytest.mark.parametrize("seed", range(40))
def test_random_gaussian_parity(itk, orc, dev, seed):
    D, size, sigma, spacing, req, ncomp, dtype = _case(seed)
    rng = np.random.default_rng(seed)
    x = rng.random(tuple(reversed(size)) + ((ncomp,) if ncomp > 1 else ())).astype(dtype)
    src = itk.Image(itk.Region([0] * D, size), tuple(spacing), (0.0,) * D, torch.from_numpy(x).to(dev))
    reg = itk.Region(list(req[0]), list(req[1]))
    ref = orc.generate_region(x, [0] * D, spacing, sigma, list(reg.index), list(reg.size), ncomp)
    for path in (itk.PATH_AUTO, itk.PATH_PER_AXIS):
        itk.set_path(path)
        try:
            f = itk.GaussianFilter(src, sigma)
            f.arith = itk.EXACT
            got = f.generate_region(reg).data.cpu().numpy()
            assert np.array_equal(got, ref), (seed, path, D, size, sigma, spacing, req, ncomp, dtype)
            f.arith = itk.FAST
            fast = f.generate_region(reg).data.cpu().numpy()
            scale = float(np.abs(ref).max()) or 1.0
            assert float(np.abs(fast.astype(np.float64) - ref).max()) / scale <= FAST_TOL, (seed, path)
        finally:
            itk.set_path(itk.PATH_AUTO)


@pytest.mark.parametrize("seed", range(16))
def test_random_recursive_parity(itk, orc, dev, seed):
    """RecursiveGaussianFilter / RecursiveGaussianAxisFilter on random shapes
    (odd extents, rows with and without a chunk tail), spacings, sigmas,
    axes, sub-requests and dtypes, bit for bit against the oracle."""
    rng = np.random.default_rng(7000 + seed)
    D = int(rng.integers(1, 4))
    size = [int(rng.integers(1, 120 if D < 3 else 60)) for _ in range(D)]
    spacing = [float(rng.choice([0.6, 1.0, 1.0, 1.7])) for _ in range(D)]
    sigma = float(rng.choice([0.3, 0.9, 1.6, 2.4, 4.1]))
    axis = int(rng.integers(-1, D))
    dtype = np.float64 if rng.random() < 0.3 else np.float32
    lo = [int(rng.integers(0, s)) for s in size]
    reg = itk.Region(lo, [int(rng.integers(1, s - l + 1)) for s, l in zip(size, lo)])
    if rng.random() < 0.3:
        reg = itk.Region([0] * D, list(size))
    x = rng.random(tuple(reversed(size))).astype(dtype)
    src = itk.Image(itk.Region([0] * D, size), tuple(spacing), (0.0,) * D, torch.from_numpy(x).to(dev))
    f = itk.RecursiveGaussianFilter(src, sigma) if axis < 0 else \
        itk.RecursiveGaussianAxisFilter(src, sigma, axis)
    got = f.generate_region(reg).data.cpu().numpy()
    ref = orc.recursive_generate_region(x, [0] * D, spacing, sigma, axis, list(reg.index),
                                        list(reg.size))
    assert np.array_equal(got, ref), (seed, D, size, spacing, sigma, axis, dtype, reg)


@pytest.mark.parametrize("seed", range(20))
def test_random_filters_parity(itk, orc, dev, seed):
    D, size, _, _, req, ncomp, dtype = _case(seed + 500)
    rng = np.random.default_rng(seed + 500)
    x = rng.random(tuple(reversed(size)) + ((ncomp,) if ncomp > 1 else ())).astype(dtype)
    src = itk.Image(itk.Region([0] * D, size), (1.0,) * D, (0.0,) * D, torch.from_numpy(x).to(dev))
    reg = itk.Region(list(req[0]), list(req[1]))
    radii = [int(rng.integers(0, 9)) for _ in range(D)]
    f = itk.BoxMeanFilterN(src, radii)
    got = f.generate_region(reg).data.cpu().numpy()
    ref = orc.mean_generate_region(x, [0] * D, radii, list(reg.index), list(reg.size), ncomp)
    assert np.array_equal(got, ref), (seed, D, size, radii, req, ncomp, dtype)
    reps = int(rng.integers(0, 6))
    b = itk.BinomialBlurFilter(src, reps)
    got = b.generate_region(reg).data.cpu().numpy()
    ref = orc.binomial_generate_region(x, [0] * D, reps, list(reg.index), list(reg.size), ncomp)
    assert np.array_equal(got, ref), (seed, D, size, reps, req, ncomp, dtype)


@pytest.mark.parametrize("seed", range(20))
def test_random_adjoint_parity(itk, orc, dev, seed):
    """GaussianFilter.backward of a random sub-request against the oracle's
    f64 transpose (scattered taps): within 1e-5 of its max |value| in f32,
    1e-12 in f64."""
    D, size, sigma, spacing, req, ncomp, dtype = _case(seed + 200)
    rng = np.random.default_rng(seed + 200)
    src = itk.Image(itk.Region([0] * D, size), tuple(spacing), (0.0,) * D, None)
    reg = itk.Region(list(req[0]), list(req[1]))
    g = rng.standard_normal(tuple(reversed(reg.size)) + ((ncomp,) if ncomp > 1 else ())).astype(dtype)
    want, ii, isz = orc.adjoint_f64(g, [0] * D, size, spacing, sigma, list(reg.index), list(reg.size),
                                    ncomp)
    f = itk.GaussianFilter(src, sigma)
    for path in (itk.PATH_AUTO, itk.PATH_PER_AXIS):
        itk.set_path(path)
        try:
            gimg = itk.Image(reg, tuple(spacing), (0.0,) * D, torch.from_numpy(g).to(dev))
            gin = f.backward(gimg)
            assert list(gin.region.index) == ii and list(gin.region.size) == isz
            got = gin.data.cpu().numpy().astype(np.float64)
            scale = float(np.abs(want).max()) or 1.0
            tol = 1e-12 if dtype == np.float64 else 1e-5
            assert float(np.abs(got - want).max()) / scale <= tol, (seed, path, D, size, sigma, req)
        finally:
            itk.set_path(itk.PATH_AUTO)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("direction", [0, 1])
def test_random_batch_equals_plain(itk, dev, seed, direction):
    """A batched plan of random uneven volumes (random sigma and spacing per
    volume, some past the fused radii) equals each volume's plain launch bit
    for bit, before and after autotuning."""
    rng = np.random.default_rng(seed + 300)
    nvol = int(rng.integers(1, 9))
    vols, sig, sps = [], [], []
    for _ in range(nvol):
        shape = tuple(int(v) for v in rng.integers(4, 60, size=3))
        vols.append(torch.from_numpy(rng.random(shape).astype(np.float32)).to(dev))
        sig.append(float(rng.uniform(0.3, 4.0)))
        sps.append((float(rng.uniform(0.5, 2.0)),) * 3)
    outs = [torch.full_like(v, float("nan")) for v in vols]
    plan = itk.BatchPlan(vols, outs, sig, sps, direction=direction, arith=itk.FAST)
    s = torch.cuda.Stream()
    for tune in (False, True):
        if tune:
            plan.autotune(s)
        plan.run(s)
        s.synchronize()
        for v, o, sg, sp in zip(vols, outs, sig, sps):
            ref = itk.gaussian(v, sg, sp, arith=itk.FAST) if direction == 0 else \
                itk.gaussian_backward(v, sg, sp)
            torch.cuda.synchronize()
            assert torch.equal(o, ref), (seed, tune, tuple(v.shape), sg, sp)
    plan.close()


@pytest.mark.parametrize("seed", range(6))
def test_random_host_staging_parity(itk, orc, dev, seed):
    # host buffers large enough to run as staged pieces (over 32 MiB of f32),
    # odd extents, sub-requests: EXACT against the oracle bit for bit,
    # through the pinned ring or pageable copies at random
    import ctypes
    import os
    rng = np.random.default_rng(3000 + seed)
    size = [int(rng.integers(150, 330)), int(rng.integers(150, 330)), int(rng.integers(120, 260))]
    sigma = float(rng.choice([0.8, 1.3, 2.0, 3.1]))
    lo = [int(rng.integers(0, s // 3)) for s in size]
    req = (lo, [int(rng.integers(s // 2, s - l + 1)) for s, l in zip(size, lo)])
    x = rng.standard_normal(tuple(reversed(size))).astype(np.float32)
    largest = itk.Region([0, 0, 0], size)
    r = itk.Region(*req)
    in_reg = itk.GaussianFilter(itk.Image(largest, (1.0,) * 3, (0.0,) * 3, None), sigma) \
        .input_region_for_output(r)
    sl = tuple(slice(in_reg.index[d], in_reg.index[d] + in_reg.size[d]) for d in reversed(range(3)))
    xin = np.ascontiguousarray(x[sl])
    out = np.empty(tuple(reversed(r.size)), dtype=np.float32)
    mode = "pageable" if rng.random() < 0.4 else None
    old = os.environ.pop("ITKG_HOST_STAGING", None)
    if mode:
        os.environ["ITKG_HOST_STAGING"] = mode
    try:
        rc = itk.lib.itkg_gaussian_generate_region_host(
            3, itk.F32, 1, sigma, (ctypes.c_double * 3)(1.0, 1.0, 1.0),
            ctypes.byref(largest.to_c()), xin.ctypes.data, ctypes.byref(r.to_c()),
            out.ctypes.data, itk.EXACT)
    finally:
        os.environ.pop("ITKG_HOST_STAGING", None)
        if old is not None:
            os.environ["ITKG_HOST_STAGING"] = old
    assert rc == 0, itk.lib.itkg_last_error()
    ref = orc.generate_region(x, [0, 0, 0], [1.0] * 3, sigma, list(r.index), list(r.size))
    assert np.array_equal(out, ref), (seed, size, req, sigma, mode)


_CAST_TYPES = [(np.uint8, torch.uint8), (np.int16, torch.int16), (np.int32, torch.int32),
               (np.float32, torch.float32), (np.float64, torch.float64)]


@pytest.mark.parametrize("seed", range(12))
def test_random_cast_parity(itk, orc, dev, seed):
    # random type pairs, modes, shapes, components and windows of larger
    # buffers (strided views) against the oracle's Rust casts, bit for bit
    rng = np.random.default_rng(4000 + seed)
    D = int(rng.integers(1, 4))
    ncomp = int(rng.choice([1, 1, 2, 3]))
    (snp, st), (dnp, dt) = [_CAST_TYPES[i] for i in rng.choice(len(_CAST_TYPES), 2)]
    mode = int(rng.integers(0, 2))
    big = [int(rng.integers(4, 90 if D < 3 else 40)) for _ in range(D)]
    shape = tuple(reversed(big)) + ((ncomp,) if ncomp > 1 else ())
    if np.issubdtype(snp, np.floating):
        x = (rng.standard_normal(shape) * rng.choice([1.0, 300.0, 1e6])).astype(snp)
    else:
        info = np.iinfo(snp)
        x = rng.integers(info.min, int(info.max) + 1, shape, dtype=np.int64).astype(snp)
    img = itk.Image(itk.Region([0] * D, big), (1.0,) * D, (0.0,) * D, torch.from_numpy(x).to(dev))
    lo = [int(rng.integers(0, b)) for b in big]
    sub = itk.Region(lo, [int(rng.integers(1, b - l + 1)) for b, l in zip(big, lo)])
    # the output is the same window of a second buffer
    ob = torch.zeros(shape, dtype=dt, device=dev)
    obig = itk.Image(img.region, img.spacing, img.origin, ob)
    win = itk.Image(sub, img.spacing, img.origin, obig.data[obig._slices(sub)])
    itk.CastFilter(img, dt, mode).generate_into(img, win)
    torch.cuda.synchronize()
    v = orc.rust_round(x) if (mode == 1 and np.issubdtype(snp, np.floating)) else x
    ref = np.zeros(shape, dtype=dnp)
    sl = tuple(slice(sub.index[d], sub.index[d] + sub.size[d]) for d in reversed(range(D)))
    ref[sl] = orc.rust_as(v[sl], dnp)
    got = ob.cpu().numpy()
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), (seed, snp, dnp, mode, big, ncomp)
