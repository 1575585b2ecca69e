"""itkg_gaussian_generate_regions: several requests over one input view
equal the same requests made one call at a time, bit for bit, FAST and
EXACT -- arbitrary (overlapping, unequal, border-touching) sub-regions
through the one-launch path, and the per-piece fallback (2-D, f64,
interleaved components, a radius without a persistent form, an empty
request) -- and EXACT pieces equal the oracle."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _multi(itk, src, sigma, regions, arith, dtype=torch.float32, ncomp=1):
    D = src.ndim
    outs = [itk.Image.empty(r, src.spacing, src.origin, ncomp, dtype, src.data.device) for r in regions]
    views = (itk.c_view * len(outs))(*[o.view() for o in outs])
    dt = itk.F32 if dtype == torch.float32 else itk.F64
    rc = itk.lib.itkg_gaussian_generate_regions(
        D, dt, ncomp, float(sigma), itk._doubles(src.spacing), ctypes.byref(src.region.to_c()),
        ctypes.byref(src.view()), len(outs), views, arith, itk._stream_ptr(src.data.device))
    assert rc == 0, itk.lib.itkg_last_error()
    torch.cuda.synchronize()
    return outs


def _single(itk, src, sigma, r, arith):
    f = itk.GaussianFilter(src, sigma)
    f.arith = arith
    return f.generate_region(r)


@pytest.mark.parametrize("arith", [0, 1])
def test_arbitrary_requests_one_launch(itk, orc, dev, arith):
    rng = np.random.default_rng(8)
    size = [150, 131, 97]
    x = rng.standard_normal(tuple(reversed(size))).astype(np.float32)
    src = itk.Image(itk.Region([3, -2, 5], size), (1.0,) * 3, (0.0,) * 3, torch.from_numpy(x).to(dev))
    regs = [itk.Region([3, -2, 5], size),                       # whole image
            itk.Region([40, 10, 30], [60, 50, 40]),             # interior
            itk.Region([3, -2, 5], [10, 131, 1]),               # one plane at the border
            itk.Region([100, 90, 60], [53, 39, 42]),            # touching the far corner
            itk.Region([50, 20, 40], [64, 64, 60])]             # overlapping the second
    for sigma in (2.0, 1.2):
        outs = _multi(itk, src, sigma, regs, arith)
        for r, o in zip(regs, outs):
            want = _single(itk, src, sigma, r, arith).data
            assert torch.equal(o.data.view(torch.int32), want.view(torch.int32)), (sigma, r)
            if arith == 0 and r.size[2] <= 42:
                ref = orc.generate_region(x, [3, -2, 5], [1.0] * 3, sigma, list(r.index), list(r.size))
                assert np.array_equal(o.data.cpu().numpy(), ref)


@pytest.mark.parametrize("case", ["2d", "f64", "vec3", "r20", "empty"])
def test_fallback_per_piece(itk, dev, case):
    rng = np.random.default_rng(9)
    if case == "2d":
        x = torch.from_numpy(rng.random((300, 200)).astype(np.float32)).to(dev)
        src = itk.Image(itk.Region([0, 0], [200, 300]), (1.0, 1.0), (0.0, 0.0), x)
        regs = [itk.Region([0, 0], [200, 150]), itk.Region([0, 150], [200, 150])]
        sigma, dtype, nc = 2.0, torch.float32, 1
    elif case == "vec3":
        x = torch.from_numpy(rng.random((40, 50, 60, 3)).astype(np.float32)).to(dev)
        src = itk.Image(itk.Region([0, 0, 0], [60, 50, 40]), (1.0,) * 3, (0.0,) * 3, x)
        regs = [itk.Region([0, 0, 0], [60, 50, 20]), itk.Region([0, 0, 20], [60, 50, 20])]
        sigma, dtype, nc = 1.5, torch.float32, 3
    else:
        dtype = torch.float64 if case == "f64" else torch.float32
        x = torch.from_numpy(rng.random((60, 70, 80))).to(dtype).to(dev)
        src = itk.Image(itk.Region([0, 0, 0], [80, 70, 60]), (1.0,) * 3, (0.0,) * 3, x)
        regs = [itk.Region([0, 0, 0], [80, 70, 30]), itk.Region([0, 0, 30], [80, 70, 30])]
        if case == "empty":
            regs.append(itk.Region([5, 5, 5], [0, 10, 10]))
        sigma, nc = (6.5 if case == "r20" else 1.5), 1     # 6.5: radius 20, no persistent form
    for arith in (0, 1):
        outs = _multi(itk, src, sigma, regs, arith, dtype, nc)
        for r, o in zip(regs, outs):
            want = _single(itk, src, sigma, r, arith).data
            assert torch.equal(o.data, want), (case, arith, r)


def test_bad_views_refused(itk, dev):
    x = torch.rand((30, 30, 30), device=dev)
    src = itk.Image(itk.Region([0, 0, 0], [30, 30, 30]), (1.0,) * 3, (0.0,) * 3, x)
    small = itk.Image(itk.Region([0, 0, 0], [30, 30, 10]), (1.0,) * 3, (0.0,) * 3, x[:10])
    out = itk.Image.empty(itk.Region([0, 0, 0], [30, 30, 20]), (1.0,) * 3, (0.0,) * 3, 1,
                          torch.float32, dev)
    views = (itk.c_view * 1)(out.view())
    rc = itk.lib.itkg_gaussian_generate_regions(
        3, itk.F32, 1, 1.5, itk._doubles((1.0,) * 3), ctypes.byref(src.region.to_c()),
        ctypes.byref(small.view()), 1, views, 1, itk._stream_ptr(dev))
    assert rc == itk.ERR_INVALID_ARG and "input" in itk.lib.itkg_last_error().decode()
