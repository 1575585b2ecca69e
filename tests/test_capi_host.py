"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/itkg.h declares, its region planning equals the oracle's and the
reference's unit tests, and bad arguments come back as error codes."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol(itk):
    hdr = open(os.path.join(ROOT, "include", "itkg.h")).read()
    declared = set(re.findall(r"\b(itkg_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(itk.EXPORTS)
    raw = ctypes.CDLL(itk.LIB_PATH)
    for name in declared:
        assert hasattr(raw, name), name


def test_version(itk):
    assert b"gfx950" in itk.lib.itkg_version()


@pytest.mark.parametrize("sigma,sp", [(1.5, 1.0), (2.0, 1.0), (3.0, 0.7), (3.0, 2.5), (-1.0, 1.0),
                                      (float("nan"), 1.0), (0.0, 1.0), (4.0, 0.5), (1.0, 3.0)])
def test_radius_equals_oracle(itk, orc, sigma, sp):
    assert itk.kernel_radius(sigma, sp) == orc.kernel_radius(sigma, sp)


@pytest.mark.parametrize("sigma_px,r", [(1.5, 5), (2.0, 6), (4.2857142857142865, 13), (8.0, 24)])
def test_weights_equal_oracle_bitwise(itk, orc, sigma_px, r):
    a = np.array(itk.build_kernel(sigma_px, r))
    b = orc.build_kernel(sigma_px, r)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_input_region_for_output(itk):
    # src/filters/gaussian.rs:52-57 on an 8-piece slab of the 512^3 volume
    src = itk.Image(itk.Region([0, 0, 0], [512, 512, 512]), (1.0,) * 3, (0.0,) * 3, None)
    f = itk.GaussianFilter(src, 2.0)
    slab = itk.Region([0, 0, 64], [512, 512, 64])
    assert f.input_region_for_output(slab) == itk.Region([0, 0, 58], [512, 512, 76])
    first = itk.Region([0, 0, 0], [512, 512, 64])
    assert f.input_region_for_output(first) == itk.Region([0, 0, 0], [512, 512, 70])


def test_region_algebra(itk):
    # src/image.rs:180-195
    r = itk.Region([10, 20], [5, 8])
    p = r.padded(2)
    assert p.index == (8, 18) and p.size == (9, 12)
    b = itk.Region([0, 0], [10, 10])
    c = itk.Region([-2, 3], [8, 10]).clipped_to(b)
    assert c.index == (0, 3) and c.size == (6, 7)


def test_split_region_reference_cases(itk):
    # src/region_splitter.rs:43-86
    R = itk.Region
    p = list(itk.split_region(R([0, 0], [4, 6]), 3))
    assert len(p) == 3 and [q.size[1] for q in p] == [2, 2, 2]
    p = list(itk.split_region(R([0, 0], [4, 7]), 3))
    assert sum(q.size[1] for q in p) == 7 and all(q.size[1] > 0 for q in p)
    assert [q.size[1] for q in p] == [3, 2, 2]   # remainder goes to the first pieces
    p = list(itk.split_region(R([0, 0], [4, 2]), 10))
    assert len(p) == 2 and sum(q.size[1] for q in p) == 2
    p = list(itk.split_region(R([5], [10]), 4))
    assert sum(q.size[0] for q in p) == 10 and p[0].index[0] == 5
    assert list(itk.split_region(R([0, 0], [4, 0]), 3)) == []
    p = list(itk.split_region(R([0, 0, 0], [512, 512, 512]), 8))
    assert [q.index[2] for q in p] == list(range(0, 512, 64))


def _call_generate(itk, ndim=3, dtype=0, ncomp=1, sigma=1.0, largest=None, req=None,
                   inreg=None, arith=0):
    L = itk.lib
    largest = largest or itk.Region([0] * ndim, [8] * ndim)
    req = req or largest
    inreg = inreg or largest
    vin, vout = itk.c_view(), itk.c_view()
    vin.data = 16
    vout.data = 32
    vin.region = inreg.to_c()
    vout.region = req.to_c()
    sp = (ctypes.c_double * max(ndim, 1))(*([1.0] * max(ndim, 1)))
    return L.itkg_gaussian_generate_region(ndim, dtype, ncomp, sigma, sp,
                                           ctypes.byref(largest.to_c()), ctypes.byref(vin),
                                           ctypes.byref(vout), arith, None)


def test_errors_without_touching_the_device(itk):
    R = itk.Region
    assert _call_generate(itk, ndim=5, largest=R([0] * 4, [8] * 4)) == itk.ERR_INVALID_ARG
    assert _call_generate(itk, dtype=7) == itk.ERR_INVALID_ARG
    assert _call_generate(itk, ncomp=9) == itk.ERR_INVALID_ARG
    assert _call_generate(itk, sigma=100.0) == itk.ERR_UNSUPPORTED          # radius 300
    assert b"radius" in itk.lib.itkg_last_error()
    # requested region outside the largest region
    assert _call_generate(itk, req=R([4, 0, 0], [8, 8, 8])) == itk.ERR_INVALID_ARG
    # input view that does not cover the padded input region
    assert _call_generate(itk, req=R([2, 2, 2], [4, 4, 4]),
                          inreg=R([2, 2, 2], [4, 4, 4])) == itk.ERR_INVALID_ARG
    # an empty request is a no-op and succeeds without any device work
    assert _call_generate(itk, req=R([0, 0, 0], [0, 8, 8])) == itk.OK
    assert itk.lib.itkg_set_path(9) == itk.ERR_INVALID_ARG


def test_host_mirror_sources_present():
    # the C++ host API and its test driver are part of the product tree
    assert os.path.exists(os.path.join(ROOT, "itk-rs_amd", "host", "itk.hpp"))
    assert os.path.exists(os.path.join(ROOT, "tests", "cpp", "test_pipeline.cpp"))
