"""The host-buffer entry points (itkg_*_generate_region_host) on requests
large enough to run as overlapped pieces: the output equals the device-view
entry point's bit for bit, for 1-, 2- and 3-D f32 images, interleaved
components, f64, sub-region requests (whose input region is the padded
request) and the separable filters' host entry."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _host_call(itk, x, ndim, sigma, largest, req, ncomp, arith):
    """itkg_gaussian_generate_region_host on the dense input region of `req`."""
    lib = itk.lib
    D = ndim
    sp = (ctypes.c_double * D)(*([1.0] * D))
    in_reg = itk.GaussianFilter(itk.Image(largest, (1.0,) * D, (0.0,) * D, None), sigma) \
        .input_region_for_output(req)
    sl = tuple(slice(in_reg.index[d] - largest.index[d],
                     in_reg.index[d] - largest.index[d] + in_reg.size[d]) for d in reversed(range(D)))
    xin = np.ascontiguousarray(x[sl])
    out = np.empty(tuple(reversed(req.size)) + ((ncomp,) if ncomp > 1 else ()), dtype=x.dtype)
    dt = itk.F32 if x.dtype == np.float32 else itk.F64
    rc = lib.itkg_gaussian_generate_region_host(D, dt, ncomp, float(sigma), sp,
                                                ctypes.byref(largest.to_c()), xin.ctypes.data,
                                                ctypes.byref(req.to_c()), out.ctypes.data, arith)
    assert rc == 0, lib.itkg_last_error()
    return out


@pytest.mark.parametrize("shape,req,ncomp,dtype,sigma", [
    ((256, 256, 256), None, 1, np.float32, 2.0),                      # 64 MiB: 4 pieces
    ((300, 260, 210), ((5, 0, 17), (200, 260, 180)), 1, np.float32, 1.5),
    ((192, 128, 160), None, 3, np.float32, 1.0),                      # 47 MB, 3 components: 2 pieces
    ((256, 128, 160), None, 1, np.float64, 2.0),                      # 42 MB: 2 pieces
    ((4096, 3000), ((0, 100), (4096, 2800)), 1, np.float32, 3.0),     # 2-D
    ((24_000_000,), None, 1, np.float32, 4.0),                          # 1-D
])
def test_host_pipeline_matches_device(itk, dev, shape, req, ncomp, dtype, sigma):
    D = len(shape)
    rng = np.random.default_rng(5)
    x = rng.random(tuple(reversed(shape)) + ((ncomp,) if ncomp > 1 else ())).astype(dtype)
    largest = itk.Region([0] * D, list(shape))
    reg = largest if req is None else itk.Region(list(req[0]), list(req[1]))
    for arith in (itk.FAST, itk.EXACT):
        got = _host_call(itk, x, D, sigma, largest, reg, ncomp, arith)
        src = itk.Image(largest, (1.0,) * D, (0.0,) * D, torch.from_numpy(x).to(dev))
        f = itk.GaussianFilter(src, sigma)
        f.arith = arith
        want = f.generate_region(reg).data.cpu().numpy()
        assert np.array_equal(got, want), (shape, arith)


def test_separable_host_pipeline_matches_device(itk, dev):
    shape = (256, 256, 200)
    rng = np.random.default_rng(6)
    x = rng.random(tuple(reversed(shape))).astype(np.float32)
    largest = itk.Region([0, 0, 0], list(shape))
    radii = (ctypes.c_uint64 * 3)(3, 3, 3)
    w = itk.box_kernel(3) * 3
    out = np.empty_like(x)
    rc = itk.lib.itkg_separable_generate_region_host(
        3, itk.F32, 1, radii, (ctypes.c_double * len(w))(*w), ctypes.byref(largest.to_c()),
        x.ctypes.data, ctypes.byref(largest.to_c()), out.ctypes.data, itk.EXACT)
    assert rc == 0, itk.lib.itkg_last_error()
    src = itk.Image(largest, (1.0,) * 3, (0.0,) * 3, torch.from_numpy(x).to(dev))
    want = itk.MeanFilter(src, 3).generate_region(largest).data.cpu().numpy()
    assert np.array_equal(out, want)


def _staging(mode):
    """ITKG_HOST_STAGING for the calls inside the block (read at each call)."""
    import contextlib
    import os

    @contextlib.contextmanager
    def cm():
        old = os.environ.get("ITKG_HOST_STAGING")
        if mode is None:
            os.environ.pop("ITKG_HOST_STAGING", None)
        else:
            os.environ["ITKG_HOST_STAGING"] = mode
        try:
            yield
        finally:
            if old is None:
                os.environ.pop("ITKG_HOST_STAGING", None)
            else:
                os.environ["ITKG_HOST_STAGING"] = old
    return cm()


@pytest.mark.parametrize("shape,req,sigma", [
    ((256, 256, 200), None, 2.0),                                # 52 MB: 3 pieces, 4 chunks
    ((301, 257, 131), ((3, 1, 9), (290, 250, 117)), 1.3),        # ragged planes and chunks
])
def test_host_pipeline_pinned_pageable_and_oracle(itk, orc, dev, shape, req, sigma):
    D = len(shape)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(tuple(reversed(shape))).astype(np.float32)
    largest = itk.Region([0] * D, list(shape))
    reg = largest if req is None else itk.Region(list(req[0]), list(req[1]))
    with _staging(None):
        pinned = _host_call(itk, x, D, sigma, largest, reg, 1, itk.EXACT)
        assert itk.host_staging_pinned()
    with _staging("pageable"):
        pageable = _host_call(itk, x, D, sigma, largest, reg, 1, itk.EXACT)
        assert not itk.host_staging_pinned()
    assert np.array_equal(pinned, pageable)
    ref = orc.generate_region(x, [0] * D, [1.0] * D, sigma, list(reg.index), list(reg.size))
    assert np.array_equal(pinned, ref)


def test_host_pipeline_trim_then_reuse(itk, dev):
    # the ring is freed by trim and created again by the next call
    shape = (256, 256, 160)
    x = np.random.default_rng(12).random(tuple(reversed(shape))).astype(np.float32)
    largest = itk.Region([0, 0, 0], list(shape))
    with _staging(None):
        a = _host_call(itk, x, 3, 2.0, largest, largest, 1, itk.FAST)
        itk.trim_scratch(0)
        b = _host_call(itk, x, 3, 2.0, largest, largest, 1, itk.FAST)
        assert itk.host_staging_pinned()
    assert np.array_equal(a, b)


def test_host_pipeline_concurrent_callers(itk, dev):
    # one caller holds the pinned ring, the other stages pageable; both are exact
    import threading
    shape = (256, 256, 144)
    xs = [np.random.default_rng(20 + i).random(tuple(reversed(shape))).astype(np.float32)
          for i in range(2)]
    largest = itk.Region([0, 0, 0], list(shape))
    want = []
    for x in xs:
        src = itk.Image(largest, (1.0,) * 3, (0.0,) * 3, torch.from_numpy(x).to(dev))
        f = itk.GaussianFilter(src, 1.5)
        f.arith = itk.FAST
        want.append(f.generate_region(largest).data.cpu().numpy())
    got = [[None, None], [None, None]]
    errs = []

    def run(i):
        try:
            for rep in range(2):
                got[i][rep] = _host_call(itk, xs[i], 3, 1.5, largest, largest, 1, itk.FAST)
        except Exception as e:  # reported below
            errs.append(e)

    with _staging(None):
        th = [threading.Thread(target=run, args=(i,)) for i in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join(timeout=60)
    assert not errs, errs
    for i in range(2):
        for rep in range(2):
            assert np.array_equal(got[i][rep], want[i]), (i, rep)


def test_cast_chain_host_pieces(itk, dev):
    # a u8 volume large enough for several pieces: the host chain equals the
    # device-resident chain (itself checked against the oracle at small sizes)
    # through the pinned ring and through pageable copies
    shape = (320, 300, 400)  # 38 MB of u8 each way: 2 pieces
    img = np.random.default_rng(13).integers(0, 256, tuple(reversed(shape)),
                                             dtype=np.int64).astype(np.uint8)
    largest = itk.Region([0, 0, 0], list(shape))
    src = itk.Image(largest, (1.0,) * 3, (0.0,) * 3, torch.from_numpy(img).to(dev))
    chain = itk.CastFilter(itk.GaussianFilter(itk.CastFilter(src, torch.float32), 1.8),
                           torch.uint8, itk.CONVERT_ROUND)
    want = chain.generate_region(largest).data.cpu().numpy()
    with _staging(None):
        a = itk.cast_gaussian_cast_host(img, 1.8, np.uint8)
        assert itk.host_staging_pinned()
    with _staging("pageable"):
        b = itk.cast_gaussian_cast_host(img, 1.8, np.uint8)
    assert np.array_equal(a, want) and np.array_equal(b, want)
