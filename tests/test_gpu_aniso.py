"""Anisotropic spacing on the fused kernel: slices thicker than the in-plane
spacing give a z radius below the x/y radius (kernel_radius per axis,
src/filters/gaussian.rs:24-27), finer slices one above it, and every pair
built (x/y radius 2..12 with z radius 1 .. x/y - 1; x/y radius 1..11 with z
radius x/y + 1 .. x/y + 4, at most 15) runs on the fused one-round-trip
kernel (itk.PATH_FUSED, which fails where the fused kernel is not built).
EXACT is bit for bit the oracle (generate_region, gaussian.rs:59-90); FAST follows
the per-axis FAST kernels' FMA chain bit for bit and is within 1e-5 (max abs
error over max abs value) of the oracle; the adjoint is within 1e-5 of the
oracle's f64 transpose."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TOL = 1e-5
PAIRS = [(r, z) for r in range(2, 13) for z in range(1, r)] + \
        [(r, z) for r in range(1, 12) for z in range(r + 1, min(r + 4, 15) + 1)]

# (z, y, x) volume; a sub-region request (index x, y, z; size x, y, z) whose
# input region is clipped at the image's low x/y/z edges and inside its high ones
SHAPE = (23, 70, 90)
REQ_INDEX, REQ_SIZE = [1, 4, 2], [80, 60, 17]


def _rel(a, b):
    return float(np.abs(a.astype(np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


def _sigma_spacing(rxy, rz):
    """sigma and spacing (1, 1, sz) with radii ceil(3 sigma) = rxy and
    ceil(3 sigma / sz) = rz."""
    sigma = rxy / 3.0 - 1e-3
    sz = 3.0 * sigma / (rz - 0.5)
    return sigma, [1.0, 1.0, sz]


def _run(itk, path, fn):
    itk.set_path(path)
    try:
        out = fn()
        torch.cuda.synchronize()
        return out
    finally:
        itk.set_path(itk.PATH_AUTO)


@pytest.mark.parametrize("rxy", range(1, 13))
def test_anisotropic_radii_on_fused_kernel(itk, orc, dev, rxy):
    rng = np.random.default_rng(600 + rxy)
    arr = rng.random(SHAPE).astype(np.float32)
    x = torch.from_numpy(arr).to(dev)
    size = list(reversed(SHAPE))
    req = itk.Region(REQ_INDEX, REQ_SIZE)
    g = rng.standard_normal(tuple(reversed(REQ_SIZE))).astype(np.float32)
    for rz in [z for r, z in PAIRS if r == rxy]:
        sigma, sp = _sigma_spacing(rxy, rz)
        assert [orc.kernel_radius(sigma, s) for s in sp] == [rxy, rxy, rz]
        img = itk.Image(itk.Region([0, 0, 0], size), tuple(sp), (0.0,) * 3, x)
        f = itk.GaussianFilter(img, sigma)
        ref = orc.generate_region(arr, [0, 0, 0], sp, sigma, REQ_INDEX, REQ_SIZE)
        # EXACT, fused: the oracle's bits
        f.arith = itk.EXACT
        ex = _run(itk, itk.PATH_FUSED, lambda: f.generate_region(req).data.cpu().numpy())
        assert np.array_equal(ex.view(np.uint32), ref.view(np.uint32)), (rxy, rz)
        # FAST, fused and per-axis: the same FMA chain
        f.arith = itk.FAST
        fa = _run(itk, itk.PATH_FUSED, lambda: f.generate_region(req).data.cpu().numpy())
        pa = _run(itk, itk.PATH_PER_AXIS, lambda: f.generate_region(req).data.cpu().numpy())
        assert np.array_equal(fa, pa), (rxy, rz)
        assert _rel(fa, ref) <= TOL, (rxy, rz)
        # adjoint, fused, against the f64 transpose
        gimg = itk.Image(req, tuple(sp), (0.0,) * 3, torch.from_numpy(g).to(dev))
        ga = _run(itk, itk.PATH_FUSED, lambda: f.backward(gimg))
        gref, in_i, in_s = orc.adjoint_f64(g.astype(np.float64), [0, 0, 0], size, sp, sigma,
                                           REQ_INDEX, REQ_SIZE)
        assert ga.region == itk.Region(in_i, in_s)
        assert _rel(ga.data.cpu().numpy(), gref) <= TOL, (rxy, rz)


def test_anisotropic_streaming_and_64_row_tiles(itk, dev, monkeypatch):
    """radii (6, 6, 3) (sigma 2 at spacing 1 x 1 x 2) over 192 x 128 x 97:
    64-row and 32-row tiles agree bit for bit, and stream_full (4 pieces, each
    its own fused launch) is bit-identical to the full request."""
    g = torch.Generator(device=dev).manual_seed(12)
    x = torch.rand((97, 128, 192), generator=g, device=dev)
    img = itk.Image(itk.Region([0, 0, 0], [192, 128, 97]), (1.0, 1.0, 2.0), (0.0,) * 3, x)
    for arith in (itk.FAST, itk.EXACT):
        f = itk.GaussianFilter(img, 2.0)
        f.arith = arith
        full = _run(itk, itk.PATH_FUSED, lambda: f.generate_region(f.largest_region()).data)
        streamed = _run(itk, itk.PATH_FUSED, lambda: itk.stream_full(f, 4).data)
        assert torch.equal(full.view(torch.int32), streamed.view(torch.int32))
        if arith == itk.FAST:
            monkeypatch.setenv("ITKG_FUSED_SHAPE", "32")
            small = _run(itk, itk.PATH_FUSED, lambda: f.generate_region(f.largest_region()).data)
            monkeypatch.delenv("ITKG_FUSED_SHAPE")
            assert torch.equal(full.view(torch.int32), small.view(torch.int32))


@pytest.mark.parametrize("shape", ["64", "32"])
def test_anisotropic_pairs_both_tile_shapes(itk, orc, dev, monkeypatch, shape):
    """Every anisotropic pair on 64-row and on 32-row tiles (the cost model
    picks one per call): FAST equal bit for bit to the per-axis FMA chain,
    the adjoint within 1e-5 of the f64 transpose (a 150 x 140 x 21 volume,
    whole request)."""
    rng = np.random.default_rng(650)
    shp = (21, 140, 150)
    arr = rng.random(shp).astype(np.float32)
    x = torch.from_numpy(arr).to(dev)
    size = list(reversed(shp))
    g = rng.standard_normal(shp).astype(np.float32)
    monkeypatch.setenv("ITKG_FUSED_SHAPE", shape)
    for rxy, rz in PAIRS:
        sigma, sp = _sigma_spacing(rxy, rz)
        img = itk.Image(itk.Region([0, 0, 0], size), tuple(sp), (0.0,) * 3, x)
        f = itk.GaussianFilter(img, sigma)
        f.arith = itk.FAST
        req = f.largest_region()
        fa = _run(itk, itk.PATH_FUSED, lambda: f.generate_region(req).data)
        pa = _run(itk, itk.PATH_PER_AXIS, lambda: f.generate_region(req).data)
        assert torch.equal(fa, pa), (rxy, rz)
        gimg = itk.Image(req, tuple(sp), (0.0,) * 3, torch.from_numpy(g).to(dev))
        ga = _run(itk, itk.PATH_FUSED, lambda: f.backward(gimg).data.cpu().numpy())
        gref, _, _ = orc.adjoint_f64(g.astype(np.float64), [0, 0, 0], size, sp, sigma, [0, 0, 0], size)
        assert _rel(ga, gref) <= TOL, (rxy, rz)


@pytest.mark.parametrize("direction", [0, 1])
def test_batched_plan_with_anisotropic_volumes(itk, orc, dev, direction):
    """A batched plan whose volumes mix isotropic spacing (the single
    persistent launch) and anisotropic spacing (plain fused launches on the
    plan's side streams): every output equals its plain call bit for bit,
    on the first run and on graph replays, and the forward matches the
    oracle within 1e-5."""
    g = torch.Generator(device=dev).manual_seed(21)
    shapes = [(40, 70, 90), (33, 64, 128), (25, 80, 66), (48, 48, 48)]
    spacings = [(1.0, 1.0, 1.0), (1.0, 1.0, 2.5), (0.8, 0.8, 1.6), (1.5, 1.5, 0.9)]
    sigmas = [1.5, 2.0, 1.2, 2.2]
    vols = [torch.rand(s, generator=g, device=dev) for s in shapes]
    outs = [torch.full_like(v, float("nan")) for v in vols]
    plan = itk.BatchPlan(vols, outs, sigmas, spacings, direction=direction, arith=itk.FAST)
    s = torch.cuda.Stream()
    for rep in range(3):
        for o in outs:
            o.fill_(float("nan"))
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            plan.run(s)
        s.synchronize()
        for v, o, sigma, sp in zip(vols, outs, sigmas, spacings):
            ref = (itk.gaussian(v, sigma, sp, arith=itk.FAST) if direction == 0
                   else itk.gaussian_backward(v, sigma, sp))
            torch.cuda.synchronize()
            assert torch.equal(o, ref), (rep, v.shape, sp)
    if direction == 0:
        for v, o, sigma, sp in zip(vols, outs, sigmas, spacings):
            arr = v.cpu().numpy()
            size = list(reversed(arr.shape))
            ref = orc.generate_region(arr, [0, 0, 0], list(sp), sigma, [0, 0, 0], size)
            assert _rel(o.cpu().numpy(), ref) <= TOL, sp
    plan.close()
