"""Canny's direction binning next to the bin edges (src/filters/edges.rs:602,
612-620).  The thresholds the GPU kernels use there are generated by
itk-rs_amd/csrc/gen_canny_edges.py; these tests check the committed table
against a regeneration, each threshold against the reference's rounding
steps, and the oracle's bins (the host libm's atan2, as the reference's
f64::atan2) against the exact rule on gradients a few ulps either side of
every edge.

The exact rule is the bin of a correctly rounded atan2.  The host libm's
atan2 is within one ulp but not always correctly rounded: where it is off by
more than half an ulp right at an edge, the reference's bin (on that libm)
and the exact rule differ.  The tests show that every such difference is one
of those roundings, proved against the true angle."""
import math
import os
import sys
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "itk-rs_amd", "csrc"))
import gen_canny_edges as gen  # noqa: E402


def near_edge_gradients(ulps=4):
    """(gx, gy, edge index, upper half-plane) around every threshold, at
    several magnitudes, gy and gx each moved by up to `ulps` ulps."""
    out = []
    for k, (e, upper, r, (ch, cl), (sh, sl)) in enumerate(gen.thresholds()):
        for rho in (1e-3, 0.37, 1.0, 3.0, 1e3):
            gx0, gy0 = rho * ch, rho * sh
            for dx in range(-ulps, ulps + 1):
                for dy in range(-ulps, ulps + 1):
                    gx, gy = gx0, gy0
                    for _ in range(abs(dx)):
                        gx = math.nextafter(gx, math.inf if dx > 0 else -math.inf)
                    for _ in range(abs(dy)):
                        gy = math.nextafter(gy, math.inf if dy > 0 else -math.inf)
                    out.append((gx, gy, k % 4, upper))
    return out


def exact_bin(gx, gy, edge, upper):
    """The bin a correctly rounded atan2 gives next to `edge`: above the edge
    exactly when the true angle exceeds the threshold midpoint m, i.e. when
    gy cos(m) - gx sin(m) > 0 (evaluated in 80-digit decimal arithmetic)."""
    e, up, r, _, _ = [t for t in gen.thresholds() if t[1] == upper][edge]
    m = (Fraction(r) + Fraction(math.nextafter(r, -math.inf))) / 2
    s, c = gen._sin_cos(m, 80)
    getcontext().prec = 90
    side = Decimal(gy) * c - Decimal(gx) * s
    return (edge + 1) % 4 if side > 0 else edge


def test_table_is_the_generated_one():
    with open(os.path.join(ROOT, "itk-rs_amd", "csrc", "canny_edges.inc")) as fh:
        assert fh.read() == gen.render()


def test_thresholds_are_the_first_directions_on_each_edge():
    rows = gen.thresholds()
    assert len(rows) == 8
    for e, upper, r, (ch, cl), (sh, sl) in rows:
        below = math.nextafter(r, -math.inf)
        assert gen.angle_of(r) >= e > gen.angle_of(below)
        assert (r > 0) == upper
        # the (hi, lo) pairs are the cosine and sine of the midpoint
        m = (r + below) / 2
        assert abs(ch - math.cos(m)) < 1e-15 and abs(sh - math.sin(m)) < 1e-15
        assert abs(cl) <= math.ulp(ch) and abs(sl) <= math.ulp(sh)
        assert abs(ch * ch + sh * sh - 1.0) < 4e-16


def correctly_rounded(r, gx, gy):
    """Whether the double r is atan2(gy, gx) rounded to nearest: the true
    angle lies between the midpoints to r's neighbours (signs of
    sin(theta - midpoint) in 80-digit arithmetic)."""
    getcontext().prec = 90

    def above(t):
        s, c = gen._sin_cos(t, 80)
        return Decimal(gy) * c - Decimal(gx) * s > 0

    lo = (Fraction(r) + Fraction(math.nextafter(r, -math.inf))) / 2
    hi = (Fraction(r) + Fraction(math.nextafter(r, math.inf))) / 2
    return above(lo) and not above(hi)


def test_oracle_bins_follow_the_exact_rule_next_to_every_edge(orc):
    g = near_edge_gradients()
    gx = np.array([t[0] for t in g])
    gy = np.array([t[1] for t in g])
    bins = orc.canny_bins(gx, gy)
    want = np.array([exact_bin(*t) for t in g], dtype=np.uint8)
    differ = np.flatnonzero(bins != want).tolist()
    # every difference is a gradient where the host libm's atan2 is not
    # correctly rounded (Python's math.atan2 is the same libm call)
    for i in differ:
        assert not correctly_rounded(math.atan2(g[i][1], g[i][0]), g[i][0], g[i][1]), g[i]
    assert len(differ) <= len(g) // 100
    # both sides of every edge occur
    for k in range(4):
        sel = np.array([t[2] == k for t in g])
        assert {int(b) for b in want[sel]} == {k, (k + 1) % 4}


def test_oracle_bins_away_from_the_edges(orc):
    # far from an edge the bin is the sector of the angle
    rng = np.random.default_rng(5)
    theta = rng.uniform(-math.pi, math.pi, 20000)
    deg = np.degrees(theta) % 180.0
    far = np.abs(((deg - 22.5) % 45.0) - 22.5) < 22.0  # more than 0.5 degree from an edge
    rho = rng.uniform(0.01, 100.0, theta.size)[far]
    gx, gy = rho * np.cos(theta[far]), rho * np.sin(theta[far])
    sector = np.floor(((deg[far] + 22.5) % 180.0) / 45.0).astype(np.uint8)
    assert (orc.canny_bins(gx, gy) == sector).all()


@pytest.mark.gpu
def test_device_bins_equal_the_oracle(itk, orc, dev):
    import ctypes
    import torch
    g = near_edge_gradients()
    rng = np.random.default_rng(6)
    special = [(0.0, 0.0), (-0.0, 0.0), (0.0, -0.0), (-0.0, -0.0), (1.0, 0.0), (-1.0, 0.0),
               (0.0, 1.0), (0.0, -1.0), (1.0, 1.0), (-1.0, 1.0), (1.0, -1.0), (-1.0, -1.0),
               (3.0, 7.0), (1e-300, 1e-300), (1e300, -1e300), (5e-324, 0.0), (0.0, 5e-324),
               (-1.0, -0.0), (-0.0, -1.0), (math.nan, 1.0), (1.0, math.nan), (math.inf, 1.0),
               (1.0, -math.inf), (math.inf, math.inf), (-math.inf, -math.inf), (-math.inf, 0.0),
               (1e301, 3e301), (1e-271, -2e-271)]
    gx = np.concatenate([[t[0] for t in g], [s[0] for s in special], rng.standard_normal(50000)])
    gy = np.concatenate([[t[1] for t in g], [s[1] for s in special], rng.standard_normal(50000)])
    want = orc.canny_bins(gx, gy)
    # next to an edge the device follows the exact rule; the oracle differs
    # from it only where the host libm's atan2 is not correctly rounded
    exact = np.array([exact_bin(*t) for t in g], dtype=np.uint8)
    differ = np.flatnonzero(want[:len(g)] != exact)
    for i in differ:
        assert not correctly_rounded(math.atan2(g[i][1], g[i][0]), g[i][0], g[i][1])
    want[:len(g)] = exact
    dx = torch.from_numpy(gx).to(dev)
    dy = torch.from_numpy(gy).to(dev)
    out = torch.empty(gx.size, dtype=torch.uint8, device=dev)
    rc = itk.lib.itkg_canny_direction_bins(dx.data_ptr(), dy.data_ptr(), gx.size, out.data_ptr(),
                                           itk._stream_ptr(dev))
    assert rc == 0, itk.lib.itkg_last_error()
    got = out.cpu().numpy()
    assert (got == want).all(), np.flatnonzero(got != want)[:10]
