"""itkg_gaussian_generate_region_noalloc (the allocation-free variant of
itkg_gaussian_generate_region) through the C ABI: bit-identical to the
allocating entry point on the per-axis path (whose intermediates live in the
static device array) and on the fused path (no scratch), on consecutive
calls whose second input is smaller than the first; a call whose
intermediates exceed the array is refused with a clear error."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _call(itk, name, src, out, sigma, arith):
    D = src.ndim
    dt, nc = itk._pixel_info(src.data, D)
    with torch.cuda.device(src.data.device):
        return getattr(itk.lib, name)(
            D, dt, nc, float(sigma), itk._doubles(src.spacing),
            ctypes.byref(src.largest_region().to_c()), ctypes.byref(src.view()),
            ctypes.byref(out.view()), arith, itk._stream_ptr(src.data.device))


def _both(itk, dev, shape, sigma, spacing, ncomp=1, dtype=torch.float32, arith=0, req=None):
    D = len(spacing)
    g = torch.Generator().manual_seed(sum(shape) + ncomp)
    full = tuple(shape) + ((ncomp,) if ncomp > 1 else ())
    x = torch.rand(full, generator=g, dtype=torch.float64).to(dtype).to(dev)
    src = itk.Image(itk.Region([0] * D, list(reversed(shape))), tuple(spacing), (0.0,) * D, x)
    reg = src.region if req is None else itk.Region(*req)
    outs = []
    for name in ("itkg_gaussian_generate_region", "itkg_gaussian_generate_region_noalloc"):
        o = itk.Image.empty(reg, src.spacing, src.origin, ncomp, dtype, dev)
        o.data.fill_(float("nan"))
        rc = _call(itk, name, src, o, sigma, arith)
        assert rc == 0, (name, itk.lib.itkg_last_error())
        outs.append(o.data)
    torch.cuda.synchronize()
    return outs


CASES = [
    # (shape zyx, sigma, spacing, ncomp, dtype, arith, request)
    ((50, 60, 70), 2.0, (1.0, 1.0, 1.0), 1, torch.float32, 0, None),
    ((50, 60, 70), 2.0, (1.0, 1.0, 1.0), 1, torch.float32, 1, None),
    ((30, 40, 45), 6.5, (1.0, 1.0, 1.0), 1, torch.float32, 1, None),   # radius 20: per-axis
    ((20, 30, 40), 1.5, (0.7, 1.0, 2.0), 3, torch.float32, 0, None),
    ((18, 22, 26), 1.2, (1.0, 1.0, 1.0), 1, torch.float64, 0, ([2, 3, 1], [20, 15, 12])),
    ((90, 120), 3.0, (1.0, 1.0), 1, torch.float32, 0, None),
]


@pytest.mark.parametrize("path", ["per_axis", "auto"])
def test_noalloc_matches_allocating_entry(itk, dev, path):
    itk.set_path(itk.PATH_PER_AXIS if path == "per_axis" else itk.PATH_AUTO)
    try:
        # two passes over the cases, each followed by a smaller input than
        # the one before it: the array keeps the previous call's contents
        for rep in range(2):
            for shape, sigma, sp, nc, dt, ar, req in CASES:
                a, b = _both(itk, dev, shape, sigma, sp, nc, dt, ar, req)
                assert torch.equal(a, b), (rep, path, shape, sigma, nc, dt, ar)
                small = tuple(max(1, s // 3) for s in shape)
                a, b = _both(itk, dev, small, sigma, sp, nc, dt, ar)
                assert torch.equal(a, b), (rep, path, small, sigma, nc, dt, ar)
    finally:
        itk.set_path(itk.PATH_AUTO)


def test_noalloc_fused_path_any_size(itk, dev):
    """The fused kernel takes no scratch: a 256^3 call (intermediates far
    beyond the array on the per-axis path) runs and matches."""
    a, b = _both(itk, dev, (256, 256, 256), 2.0, (1.0, 1.0, 1.0), arith=1)
    assert torch.equal(a, b)


def test_noalloc_refuses_oversized_per_axis_call(itk, dev):
    itk.set_path(itk.PATH_PER_AXIS)
    try:
        x = torch.zeros((64, 512, 512), device=dev)
        src = itk.Image(itk.Region([0, 0, 0], [512, 512, 64]), (1.0,) * 3, (0.0,) * 3, x)
        o = itk.Image.empty(src.region, src.spacing, src.origin, 1, torch.float32, dev)
        rc = _call(itk, "itkg_gaussian_generate_region_noalloc", src, o, 2.0, 0)
        assert rc == itk.ERR_UNSUPPORTED
        msg = itk.lib.itkg_last_error().decode()
        assert "scratch" in msg and str(64 << 20) in msg, msg
        # the allocating entry point serves the same call
        assert _call(itk, "itkg_gaussian_generate_region", src, o, 2.0, 0) == 0
        torch.cuda.synchronize()
    finally:
        itk.set_path(itk.PATH_AUTO)
