"""GPU parity of the forward path (GaussianFilter::generate_region) against
the oracle: bit-exact in EXACT mode on every case, within 1e-5 (max abs error
over max abs value) in FAST mode."""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FAST_TOL = 1e-5


def _img(itk, arr, index, spacing, dev):
    D = len(spacing)
    t = torch.from_numpy(np.ascontiguousarray(arr)).to(dev)
    size = list(reversed(arr.shape[:D]))
    return itk.Image(itk.Region(index, size), tuple(spacing), (0.0,) * D, t)


def _bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    u = np.uint32 if a.dtype == np.float32 else np.uint64
    return a.shape == b.shape and np.array_equal(a.view(u), b.view(u))


def _rel_err(a, b):
    return float(np.abs(a.astype(np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


CASES = [
    # numpy shape, image index, spacing, sigma, requested index, requested size
    ((9, 13), [0, 0], [1.0, 1.0], 1.5, [0, 0], [13, 9]),
    ((40, 37), [3, -2], [1.0, 0.7], 1.0, [5, 0], [30, 21]),
    ((17,), [5], [1.0], 1.5, [7], [6]),
    ((7, 8, 9), [0, 0, 0], [1.0, 1.0, 1.0], 2.0, [0, 0, 0], [9, 8, 7]),
    ((3, 2, 4), [0, 0, 0], [1.0, 1.0, 1.0], 2.5, [0, 0, 0], [4, 2, 3]),
    ((1, 1, 5), [0, 0, 0], [1.0, 1.0, 1.0], 1.0, [0, 0, 0], [5, 1, 1]),
    ((33, 70, 66), [0, 0, 0], [1.0, 1.0, 1.0], 2.0, [0, 0, 0], [66, 70, 33]),
    ((33, 70, 66), [-4, 2, 1], [1.0, 1.0, 1.0], 2.0, [3, 9, 7], [50, 41, 20]),
    ((25, 30, 140), [0, 0, 0], [0.7, 0.7, 2.5], 3.0, [0, 0, 0], [140, 30, 25]),
    ((4, 3, 5, 6), [0, 0, 0, 0], [1.0, 1.0, 1.0, 2.0], 1.0, [1, 0, 1, 0], [4, 3, 2, 3]),
]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("ncomp", [1, 3])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("path", ["auto", "per_axis"])
def test_exact_matches_oracle(itk, orc, dev, case, ncomp, dtype, path):
    shape, ii, sp, sigma, ri, rs = case
    rng = np.random.default_rng(hash((shape, ncomp)) & 0xFFFF)
    arr = rng.random(shape + ((ncomp,) if ncomp > 1 else ())).astype(dtype)
    itk.set_path(itk.PATH_AUTO if path == "auto" else itk.PATH_PER_AXIS)
    try:
        f = itk.GaussianFilter(_img(itk, arr, ii, sp, dev), sigma)
        out = f.generate_region(itk.Region(ri, rs))
        torch.cuda.synchronize()
    finally:
        itk.set_path(itk.PATH_AUTO)
    ref = orc.generate_region(arr, ii, sp, sigma, ri, rs, ncomp)
    assert out.region == itk.Region(ri, rs)
    assert _bits_equal(out.data.cpu().numpy(), ref)


@pytest.mark.parametrize("r", list(range(1, 16)))
@pytest.mark.parametrize("arith", ["exact", "fast"])
def test_fused_radii_match_per_axis(itk, orc, dev, r, arith):
    """Every fused 3-D instantiation (isotropic radius 1..15) against the per-axis
    kernel on a volume with ragged tile edges and a sub-region request."""
    sigma = r / 3.0 - 1e-3
    assert itk.kernel_radius(sigma, 1.0) == r
    rng = np.random.default_rng(r)
    arr = rng.random((45, 51, 70)).astype(np.float32)
    img = _img(itk, arr, [0, 0, 0], [1.0, 1.0, 1.0], dev)
    f = itk.GaussianFilter(img, sigma)
    f.arith = itk.EXACT if arith == "exact" else itk.FAST
    req = itk.Region([2, 5, 3], [66, 40, 41])
    itk.set_path(itk.PATH_FUSED)
    try:
        fused = f.generate_region(req).data.cpu().numpy()
        full = f.generate_region(f.largest_region()).data.cpu().numpy()
    finally:
        itk.set_path(itk.PATH_PER_AXIS)
    try:
        per = f.generate_region(req).data.cpu().numpy()
    finally:
        itk.set_path(itk.PATH_AUTO)
    if arith == "exact":
        assert _bits_equal(fused, per)
        ref = orc.generate_region(arr, [0, 0, 0], [1.0] * 3, sigma, [0, 0, 0], [70, 51, 45])
        assert _bits_equal(full, ref)
    else:
        assert np.array_equal(fused, per)      # same FMA chain; equal up to the sign of zero
        ref = orc.generate_region(arr, [0, 0, 0], [1.0] * 3, sigma, [2, 5, 3], [66, 40, 41])
        assert _rel_err(fused, ref) <= FAST_TOL


def test_strided_input_view(itk, orc, dev):
    """The input may be a window of a larger buffer (non-dense strides)."""
    rng = np.random.default_rng(7)
    big = rng.random((30, 40, 50)).astype(np.float32)
    t = torch.from_numpy(big).to(dev)
    win = t[4:25, 6:33, 5:44]                  # strided view
    img = itk.Image(itk.Region([5, 6, 4], [39, 27, 21]), (1.0,) * 3, (0.0,) * 3, win)
    f = itk.GaussianFilter(img, 1.6)
    out = f.generate_region(itk.Region([8, 9, 7], [30, 18, 12])).data.cpu().numpy()
    ref = orc.generate_region(np.ascontiguousarray(big[4:25, 6:33, 5:44]), [5, 6, 4], [1.0] * 3,
                              1.6, [8, 9, 7], [30, 18, 12])
    assert _bits_equal(out, ref)


def test_fast_mode_tolerance_c3_shape(itk, orc, dev):
    """VecPixel<f32,3> anisotropic configuration at reduced size, FAST mode."""
    from itk_rs_amd import workloads as wl
    v, sigma, sp = wl.c3("cpu", shape=(40, 64, 64))
    arr = v.numpy()
    for arith in (itk.EXACT, itk.FAST):
        t = torch.from_numpy(arr).to(dev)
        out = itk.gaussian(t, sigma, sp, ndim=3, arith=arith).cpu().numpy()
        ref = orc.generate_region(arr, [0, 0, 0], sp, sigma, [0, 0, 0], [64, 64, 40], 3)
        if arith == itk.EXACT:
            assert _bits_equal(out, ref)
        else:
            assert _rel_err(out, ref) <= FAST_TOL


def test_c1_512sq_exact(itk, orc, dev):
    from itk_rs_amd import workloads as wl
    x, sigma, sp = wl.c1("cpu")
    out = itk.gaussian(x.to(dev), sigma, sp).cpu().numpy()
    ref = orc.generate_region(x.numpy(), [0, 0], sp, sigma, [0, 0], [512, 512])
    assert _bits_equal(out, ref)
    fast = itk.gaussian(x.to(dev), sigma, sp, arith=itk.FAST).cpu().numpy()
    assert _rel_err(fast, ref) <= FAST_TOL


def _golden():
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as fh:
        return json.load(fh)


def test_c2_512cube_golden_and_streaming(itk, dev):
    """512^3 f32, sigma 2: full output hash equals the oracle's (computed
    offline from the same seeded input), sampled voxels match, and the 8-piece
    stream_full output is bit-identical to the full-image output."""
    from itk_rs_amd import workloads as wl
    gold = _golden()["c2"]
    x, sigma, sp = wl.c2(dev)
    assert hashlib.sha256(x.cpu().numpy().tobytes()).hexdigest() == gold["input_sha256"]
    img = itk.Image(itk.Region([0, 0, 0], [512, 512, 512]), sp, (0.0,) * 3, x)
    f = itk.GaussianFilter(img, sigma)
    full = f.generate_region(f.largest_region()).data
    streamed = itk.stream_full(f, 8).data
    torch.cuda.synchronize()
    a = full.cpu().numpy()
    assert hashlib.sha256(a.tobytes()).hexdigest() == gold["output_sha256"]
    idx = np.array(gold["sample_flat_index"], dtype=np.int64)
    vals = np.array(gold["sample_value_bits"], dtype=np.uint32)
    assert np.array_equal(a.reshape(-1)[idx].view(np.uint32), vals)
    assert torch.equal(full.view(torch.int32), streamed.view(torch.int32))
    fast = f.__class__(img, sigma)
    fast.arith = itk.FAST
    fout = fast.generate_region(fast.largest_region()).data
    err = float((fout.double() - full.double()).abs().max() / full.abs().max())
    assert err <= FAST_TOL
    fstream = itk.stream_full(fast, 8).data
    assert torch.equal(fout, fstream)


def test_empty_request_is_empty(itk, dev):
    img = _img(itk, np.zeros((5, 6, 7), np.float32), [0, 0, 0], [1.0] * 3, dev)
    f = itk.GaussianFilter(img, 1.0)
    out = f.generate_region(itk.Region([1, 1, 1], [0, 3, 3]))
    assert out.data.numel() == 0


def test_cpp_host_pipeline():
    """The reference's pipeline tests through the C++ host API."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "cpp", "test_pipeline")
    assert os.path.exists(exe), "build the C++ driver first (make -C tests/cpp)"
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "all passed" in res.stdout


PER_AXIS_CASES = [
    # rows longer than one staged segment (64 * 8 outputs), large radius, clamp at both ends
    ((1500,), [0], [1.0], 7.0, [100], [1200]),
    ((1300,), [-40], [0.5], 2.0, [-40], [1300]),
    # radius beyond the fused set along every axis of a 2-D and a 3-D image
    ((70, 600), [0, 0], [1.0, 1.0], 9.0, [0, 0], [600, 70]),
    ((40, 50, 30), [2, -3, 1], [1.0, 0.5, 2.0], 6.0, [3, -1, 4], [25, 44, 33]),
    # anisotropic in-plane radii (no fused instantiation)
    ((20, 33, 90), [0, 0, 0], [0.6, 1.3, 1.0], 2.2, [0, 0, 0], [90, 33, 20]),
    # 4-D, largest radius along the slowest axis; tiny extents
    ((9, 5, 6, 70), [0, 0, 0, 0], [1.0, 1.0, 1.0, 0.25], 2.0, [3, 0, 1, 2], [60, 6, 4, 7]),
    ((2, 1, 3), [0, 0, 0], [1.0, 1.0, 1.0], 11.0, [0, 0, 0], [3, 1, 2]),
]


@pytest.mark.parametrize("case", PER_AXIS_CASES)
@pytest.mark.parametrize("ncomp", [1, 2, 4])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_per_axis_runs_exact_and_fast(itk, orc, dev, case, ncomp, dtype):
    """The per-axis kernels (strided-axis runs and staged rows) against the
    oracle: EXACT bit-exact, FAST within tolerance."""
    shape, ii, sp, sigma, ri, rs = case
    rng = np.random.default_rng(len(shape) * 100 + ncomp)
    arr = rng.random(shape + ((ncomp,) if ncomp > 1 else ())).astype(dtype)
    ref = orc.generate_region(arr, ii, sp, sigma, ri, rs, ncomp)
    itk.set_path(itk.PATH_PER_AXIS)
    try:
        for arith in (itk.EXACT, itk.FAST):
            f = itk.GaussianFilter(_img(itk, arr, ii, sp, dev), sigma)
            f.arith = arith
            out = f.generate_region(itk.Region(ri, rs)).data.cpu().numpy()
            if arith == itk.EXACT:
                assert _bits_equal(out, ref)
            else:
                assert _rel_err(out, ref) <= (FAST_TOL if dtype == np.float32 else 1e-12)
    finally:
        itk.set_path(itk.PATH_AUTO)


def test_per_axis_strided_output_view(itk, orc, dev):
    """Output written through a window of a larger buffer on the per-axis path."""
    rng = np.random.default_rng(9)
    arr = rng.random((12, 20, 300)).astype(np.float32)
    img = _img(itk, arr, [0, 0, 0], [1.0, 1.0, 1.0], dev)
    f = itk.GaussianFilter(img, 6.0)        # radius 18: per-axis
    req = itk.Region([10, 2, 1], [270, 15, 9])
    big = torch.full((14, 19, 280), float("nan"), device=dev)
    win = big[2:11, 1:16, 5:275]
    out = itk.Image(req, (1.0,) * 3, (0.0,) * 3, win)
    f.generate_into(img, out)
    torch.cuda.synchronize()
    ref = orc.generate_region(arr, [0, 0, 0], [1.0] * 3, 6.0, [10, 2, 1], [270, 15, 9])
    assert _bits_equal(win.cpu().numpy(), ref)
    rest = big.clone()
    rest[2:11, 1:16, 5:275] = 0.0
    assert torch.isnan(rest).sum().item() == rest.numel() - win.numel()  # nothing else written


@pytest.mark.parametrize("sigma", [0.3, 0.4, 0.9, 1.2, 1.5, 2.0, 3.4])
def test_fused_tile_shapes_agree(itk, orc, dev, monkeypatch, sigma):
    """FAST forward on 64 x 64, 64 x 32 and 128 x 32 tiles (the last built
    for radii 1..5; at radius 6 and 11 the request falls back to 64 x 64)
    and the adjoint on the first two: every voxel follows the same FMA chain,
    so the shapes agree bit for bit; all within tolerance of the oracle.
    EXACT on 64 x 64 and 128 x 32 tiles (built for radii 1..4): bit for bit
    the oracle.  x = 130: a 128-wide tile and a 2-column one; y = 90: a
    partial last row of tiles."""
    rng = np.random.default_rng(31)
    arr = rng.random((37, 90, 130)).astype(np.float32)
    x = torch.from_numpy(arr).to(dev)
    outs, grads, exact = {}, {}, {}
    for shape in ("64", "32", "128"):
        monkeypatch.setenv("ITKG_FUSED_SHAPE", shape)
        outs[shape] = itk.gaussian(x, sigma, (1.0, 1.0, 1.0), arith=itk.FAST)
        grads[shape] = itk.gaussian_backward(x, sigma, (1.0, 1.0, 1.0))
        exact[shape] = itk.gaussian(x, sigma, (1.0, 1.0, 1.0), arith=itk.EXACT)
    torch.cuda.synchronize()
    assert torch.equal(outs["64"], outs["32"])
    assert torch.equal(outs["64"], outs["128"])
    assert torch.equal(grads["64"], grads["32"])
    # EXACT (64 x 64 tiles, or 128 x 32 at radii 1..4) bit for bit against the oracle
    ref_exact = orc.generate_region(arr, [0, 0, 0], [1.0] * 3, sigma, [0, 0, 0], [130, 90, 37])
    for shape in ("64", "128"):
        assert _bits_equal(exact[shape].cpu().numpy(), ref_exact), shape
    ref = orc.generate_region(arr, [0, 0, 0], [1.0] * 3, sigma, [0, 0, 0], [130, 90, 37])
    assert _rel_err(outs["64"].cpu().numpy(), ref) <= FAST_TOL


@pytest.mark.parametrize("r", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("arith", ["fast", "exact"])
def test_wide_tiles_subregion(itk, orc, dev, monkeypatch, r, arith):
    """128 x 32 tiles (FAST radii 1..5, EXACT radii 1..4; EXACT radius 5
    falls back to 64 x 64) on a sub-region request of a volume whose extents
    are not multiples of the tile (x = 197: a last tile 69 columns wide,
    both 64-column halves in use; y = 75): FAST gives the same values as the
    per-axis FAST kernels and is within tolerance of the oracle, EXACT is
    bit for bit the oracle."""
    sigma = r / 3.0 - 1e-3
    rng = np.random.default_rng(100 + r)
    arr = rng.random((33, 75, 197)).astype(np.float32)
    img = _img(itk, arr, [0, 0, 0], [1.0, 1.0, 1.0], dev)
    f = itk.GaussianFilter(img, sigma)
    f.arith = itk.FAST if arith == "fast" else itk.EXACT
    req = itk.Region([3, 4, 2], [190, 69, 29])
    monkeypatch.setenv("ITKG_FUSED_SHAPE", "128")
    itk.set_path(itk.PATH_FUSED)
    try:
        wide = f.generate_region(req).data.cpu().numpy()
    finally:
        itk.set_path(itk.PATH_PER_AXIS)
    try:
        per = f.generate_region(req).data.cpu().numpy()
    finally:
        itk.set_path(itk.PATH_AUTO)
    ref = orc.generate_region(arr, [0, 0, 0], [1.0] * 3, sigma, [3, 4, 2], [190, 69, 29])
    if arith == "fast":
        assert np.array_equal(wide, per)  # same FMA chain; equal up to the sign of zero
        assert _rel_err(wide, ref) <= FAST_TOL
    else:
        assert _bits_equal(wide, ref)
        assert _bits_equal(per, ref)


@pytest.mark.parametrize("sigma", [1.2, 2.0])
def test_fused_tile_shapes_agree_interleaved(itk, dev, monkeypatch, sigma):
    """Three interleaved components (VecPixel<f32,3>): the strided output
    path of both tile shapes and of the per-axis kernels agrees bit for bit
    (x = 200: a partial last tile)."""
    rng = np.random.default_rng(32)
    arr = rng.random((21, 40, 200, 3)).astype(np.float32)
    x = torch.from_numpy(arr).to(dev)
    outs = {}
    for shape in ("64", "32", "128"):
        monkeypatch.setenv("ITKG_FUSED_SHAPE", shape)
        outs[shape] = itk.gaussian(x, sigma, (1.0, 1.0, 1.0), ndim=3, arith=itk.FAST)
    torch.cuda.synchronize()
    assert torch.equal(outs["64"], outs["32"])
    assert torch.equal(outs["64"], outs["128"])
    monkeypatch.delenv("ITKG_FUSED_SHAPE")
    per = {}
    itk.set_path(itk.PATH_PER_AXIS)
    try:
        per = itk.gaussian(x, sigma, (1.0, 1.0, 1.0), ndim=3, arith=itk.FAST)
    finally:
        itk.set_path(itk.PATH_AUTO)
    assert torch.equal(outs["64"], per)  # the per-axis FAST kernels follow the same FMA chain


@pytest.mark.parametrize("r", [1, 3, 6, 7, 9])
def test_per_axis_row_and_group_passes(itk, orc, dev, r):
    """The per-axis path's axis-0 pass (outputs staged through LDS, stored 64
    consecutive pixels per instruction) and its strided passes over rows of
    a multiple of 4 pixels (4 adjacent pixels per lane, below radius 8), on
    a sub-region request: EXACT bit for bit the oracle, FAST equal to the
    fused kernel's FMA chain, the adjoint within 1e-5 of the f64 transpose;
    x = 128 takes the 4-pixel groups, x = 130 the one-pixel lanes."""
    sigma = r / 3.0 - 1e-3
    for nx in (128, 130):
        rng = np.random.default_rng(700 + r + nx)
        arr = rng.random((19, 40, nx)).astype(np.float32)
        img = _img(itk, arr, [0, 0, 0], [1.0, 1.0, 1.0], dev)
        req = itk.Region([4, 3, 2], [nx - 8, 33, 15])
        f = itk.GaussianFilter(img, sigma)
        ref = orc.generate_region(arr, [0, 0, 0], [1.0] * 3, sigma, list(req.index), list(req.size))
        itk.set_path(itk.PATH_PER_AXIS)
        try:
            f.arith = itk.EXACT
            ex = f.generate_region(req).data.cpu().numpy()
            f.arith = itk.FAST
            fa = f.generate_region(req).data.cpu().numpy()
            g = rng.standard_normal((15, 33, nx - 8)).astype(np.float32)
            gimg = itk.Image(req, (1.0,) * 3, (0.0,) * 3, torch.from_numpy(g).to(dev))
            ga = f.backward(gimg).data.cpu().numpy()
        finally:
            itk.set_path(itk.PATH_AUTO)
        assert _bits_equal(ex, ref), nx
        fused = f.generate_region(req).data.cpu().numpy()
        assert np.array_equal(fa, fused), nx
        gref, _, _ = orc.adjoint_f64(g.astype(np.float64), [0, 0, 0], [nx, 40, 19], [1.0] * 3, sigma,
                                     list(req.index), list(req.size))
        assert _rel_err(ga, gref) <= FAST_TOL, nx


def test_gaussian_from_threads_and_alternating_layouts(itk, orc, dev):
    """gaussian() keeps its view structs per thread and layout: four threads
    filtering distinct images of one layout at once, and one thread
    alternating layouts (2-D, 3-D, interleaved components, f64), each equal
    bit for bit to the result of a lone call; the 2-D image against the
    oracle."""
    import threading
    g = torch.Generator(device=dev).manual_seed(31)
    imgs = [torch.rand((96, 130), generator=g, device=dev) for _ in range(4)]
    want = [itk.gaussian(x, 1.5, arith=itk.EXACT) for x in imgs]
    torch.cuda.synchronize()
    ref = orc.generate_region(imgs[0].cpu().numpy(), [0, 0], [1.0, 1.0], 1.5, [0, 0], [130, 96])
    assert np.array_equal(want[0].cpu().numpy(), ref)
    got = [None] * 4
    errors = []

    def run(k):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for _ in range(20):
                    got[k] = itk.gaussian(imgs[k], 1.5, arith=itk.EXACT)
            s.synchronize()
        except Exception as e:  # reported below, in the main thread
            errors.append(e)

    threads = [threading.Thread(target=run, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for k in range(4):
        assert torch.equal(got[k], want[k]), k
    others = [torch.rand((20, 30, 40), generator=g, device=dev),
              torch.rand((30, 40, 3), generator=g, device=dev),
              torch.rand((96, 130), generator=g, device=dev, dtype=torch.float64)]
    alone = [itk.gaussian(others[0], 1.2), itk.gaussian(others[1], 1.2, ndim=2),
             itk.gaussian(others[2], 1.2)]
    for _ in range(3):
        for x, a in zip([imgs[0]] + others, [want[0]] + alone):
            nd = 2 if x.dim() == 3 and x.shape[-1] == 3 else None
            sigma = 1.5 if x is imgs[0] else 1.2
            assert torch.equal(itk.gaussian(x, sigma, ndim=nd, arith=itk.EXACT), a)
