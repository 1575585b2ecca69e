"""Pins the oracle's restatements of the other convolve_axis filters --
MeanFilter / BoxMeanFilterN (src/filters/mean.rs:15-18, 59-83, 126-144),
BinomialBlurFilter (src/filters/binomial_blur.rs:57-89),
UnsharpMaskFilter (src/filters/edges.rs:406-452), the recursive Gaussian
(src/filters/recursive_gaussian.rs:33-240) and CannyEdgeDetectionFilter
(src/filters/edges.rs:548-664) -- before the GPU path is checked against
them.  UnsharpMaskFilter and Canny have no tests in the reference: their
pins are closed forms, identity cases, the dense operator, an independent
restatement and crop consistency.

Checks: the reference's own unit tests for both filters (mean.rs:261-294,
binomial_blur.rs:92-129), interior-impulse closed forms that are exact in
f32 (box products; binomial coefficients / 4^n), and the brute-force dense
per-axis operator in f64 with weights written out in the test, on requests
that touch the image border (where the binomial's per-repetition clamp
differs from one wide kernel).
"""
import math

import numpy as np
import pytest

RNG = np.random.default_rng(4321)


def rand(shape, ncomp=1, dtype=np.float32):
    shp = tuple(shape) + ((ncomp,) if ncomp > 1 else ())
    return RNG.random(shp).astype(dtype)


# --- box_kernel (mean.rs:15-18) ---------------------------------------------

@pytest.mark.parametrize("r", [0, 1, 2, 5, 13])
def test_box_kernel(orc, r):
    k = orc.box_kernel(r)
    assert len(k) == 2 * r + 1
    assert np.all(k == 1.0 / (2 * r + 1))


# --- the reference's own tests ----------------------------------------------

def test_mean_constant_image(orc):
    # mean.rs:261-269
    out = orc.mean_generate_region(np.full(10, 5.0, np.float32), [0], [2], [0], [10])
    assert np.all(np.abs(out - 5.0) < 1e-5)


def test_mean_ramp_interior(orc):
    # mean.rs:271-282: a ramp's symmetric-window mean is the centre value
    ramp = np.arange(20, dtype=np.float32)
    out = orc.mean_generate_region(ramp, [0], [2], [0], [20])
    for i in range(2, 18):
        assert abs(out[i] - i) < 1e-4


def test_box_mean_2d_constant(orc):
    # mean.rs:284-294
    out = orc.mean_generate_region(np.full((8, 8), 3.0, np.float32), [0, 0], [1, 2], [0, 0], [8, 8])
    assert np.all(np.abs(out - 3.0) < 1e-5)


def test_binomial_constant_image_preserved(orc):
    # binomial_blur.rs:97-107
    out = orc.binomial_generate_region(np.full(20, 3.0, np.float32), [0], 4, [0], [20])
    assert np.all(np.abs(out[4:16] - 3.0) < 1e-5)


def test_binomial_impulse_spreads(orc):
    # binomial_blur.rs:109-122
    x = np.zeros(21, np.float32)
    x[10] = 1.0
    out = orc.binomial_generate_region(x, [0], 4, [0], [21])
    assert 0.0 < out[10] < 1.0
    assert abs(float(out.sum()) - 1.0) < 0.05


# --- closed forms (exact in f32) --------------------------------------------

@pytest.mark.parametrize("D,n", [(1, 1), (1, 4), (2, 3), (3, 2)])
def test_binomial_impulse_closed_form(orc, D, n):
    # away from the border, n rounds of [1,2,1]/4 give C(2n, k) / 4^n on each
    # axis; every value and partial sum is a short dyadic fraction, so the f32
    # arithmetic is exact and the outer product is bit-exact
    N = 4 * n + 5
    c = N // 2
    x = np.zeros((N,) * D, np.float32)
    x[(c,) * D] = 1.0
    out = orc.binomial_generate_region(x, [0] * D, n, [0] * D, [N] * D)
    line = np.zeros(N)
    for k in range(2 * n + 1):
        line[c - n + k] = math.comb(2 * n, k) / 4.0 ** n
    want = line
    for _ in range(D - 1):
        want = np.multiply.outer(want, line)
    assert np.array_equal(out.astype(np.float64), want)


@pytest.mark.parametrize("radii", [[1], [3], [2, 1], [1, 2, 3]])
def test_mean_impulse_closed_form(orc, radii):
    # one nonzero input: each pass multiplies by fl32(1/len) once, so the
    # output is the running f32 product of the per-axis weights inside the
    # box and zero outside
    D = len(radii)
    N = 2 * max(radii) + 5
    c = N // 2
    x = np.zeros((N,) * D, np.float32)
    x[(c,) * D] = 1.0
    out = orc.mean_generate_region(x, [0] * D, radii, [0] * D, [N] * D)
    want = np.zeros((N,) * D, np.float32)
    val = np.float32(1.0)
    for a in range(D):
        val = np.float32(val * np.float32(1.0 / (2 * radii[a] + 1)))
    sl = tuple(slice(c - radii[D - 1 - i], c + radii[D - 1 - i] + 1) for i in range(D))
    want[sl] = val
    assert np.array_equal(out, want)


# --- dense operator (brute force, f64) --------------------------------------

@pytest.mark.parametrize("size,radii,req", [
    ((9, 7, 6), (1, 2, 3), ((0, 0, 0), (9, 7, 6))),
    ((12, 10, 8), (2, 2, 1), ((3, 1, 0), (5, 6, 4))),
    ((20, 15), (4, 0), ((0, 7), (20, 8))),
])
def test_mean_vs_dense(orc, size, radii, req):
    x = rand(tuple(reversed(size)), dtype=np.float64)
    out = orc.mean_generate_region(x, [0] * len(size), list(radii), list(req[0]), list(req[1]))
    mats, in_i, in_s = orc.dense_mean_operator([0] * len(size), list(size), list(radii),
                                               list(req[0]), list(req[1]))
    sub = x[tuple(slice(i, i + s) for i, s in zip(reversed(in_i), reversed(in_s)))]
    assert np.allclose(out, orc.apply_dense(mats, sub), rtol=0, atol=1e-12)
    out32 = orc.mean_generate_region(x.astype(np.float32), [0] * len(size), list(radii),
                                     list(req[0]), list(req[1]))
    assert np.max(np.abs(out32 - out)) < 1e-6


@pytest.mark.parametrize("size,reps,req", [
    ((9, 8, 7), 2, ((0, 0, 0), (9, 8, 7))),
    ((14, 11), 3, ((0, 2), (6, 9))),        # touches the x = 0 border
    ((16, 12, 10), 4, ((5, 0, 6), (6, 12, 4))),
    ((10,), 6, ((0,), (10,))),               # radius 6 exceeds half the image
])
def test_binomial_vs_dense(orc, size, reps, req):
    D = len(size)
    x = rand(tuple(reversed(size)), dtype=np.float64)
    out = orc.binomial_generate_region(x, [0] * D, reps, list(req[0]), list(req[1]))
    mats, in_i, in_s = orc.dense_binomial_operator([0] * D, list(size), reps,
                                                   list(req[0]), list(req[1]))
    sub = x[tuple(slice(i, i + s) for i, s in zip(reversed(in_i), reversed(in_s)))]
    assert np.allclose(out, orc.apply_dense(mats, sub), rtol=0, atol=1e-12)


def test_binomial_is_not_one_wide_kernel_at_the_border(orc):
    # n rounds of a clamped 3-tap kernel differ from one clamped 2n+1-tap
    # binomial kernel next to the border: the oracle must follow the rounds
    x = rand((12,), dtype=np.float64)
    n = 3
    out = orc.binomial_generate_region(x, [0], n, [0], [12])
    w = np.array([math.comb(2 * n, k) / 4.0 ** n for k in range(2 * n + 1)])
    wide = orc.apply_dense([orc.axis_matrix(w, 0, 12, 0, 12)], x)
    assert np.allclose(out[n:-n], wide[n:-n], atol=1e-12)
    assert not np.allclose(out[:n], wide[:n], atol=1e-9)


# --- UnsharpMaskFilter (edges.rs:406-452) ------------------------------------

@pytest.mark.parametrize("sigma", [0.5, 1.0, 1.7, 3.0])
def test_index_kernel_closed_form(orc, sigma):
    k = orc.index_kernel(sigma)
    r = math.ceil(3.0 * sigma)
    raw = [math.exp(-((i - r) ** 2) / (2.0 * sigma * sigma)) for i in range(2 * r + 1)]
    s = 0.0
    for v in raw:
        s += v
    assert len(k) == 2 * r + 1
    assert np.array_equal(k, np.array(raw) / s)
    assert np.array_equal(k, k[::-1])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_unsharp_identity_cases(orc, dtype):
    x = rand((9, 11, 13), dtype=dtype)
    idx, size = [0, 0, 0], [13, 11, 9]
    # amount 0: orig + 0 * edge = orig; threshold +inf: no pixel passes;
    # sigma < 0: radius 0, weight 1, so blur = orig and edge = 0
    assert np.array_equal(orc.unsharp_generate_region(x, idx, 1.2, 0.0, 0.0, idx, size), x)
    assert np.array_equal(orc.unsharp_generate_region(x, idx, 1.2, 0.7, math.inf, idx, size), x)
    assert np.array_equal(orc.unsharp_generate_region(x, idx, -1.0, 0.7, 0.0, idx, size), x)


@pytest.mark.parametrize("size,sigma,amount,thr,req", [
    ((10, 9, 8), 1.0, 0.5, 0.0, ((0, 0, 0), (10, 9, 8))),
    ((14, 12), 1.4, 1.5, 0.05, ((2, 0), (8, 12))),
    ((11, 10, 9), 0.8, 0.5, 0.02, ((3, 2, 1), (6, 5, 5))),
])
def test_unsharp_vs_dense(orc, size, sigma, amount, thr, req):
    D = len(size)
    x = rand(tuple(reversed(size)), dtype=np.float64)
    out = orc.unsharp_generate_region(x, [0] * D, sigma, amount, thr, list(req[0]), list(req[1]))
    r = math.ceil(3.0 * sigma)
    w = np.array([math.exp(-((i - r) ** 2) / (2.0 * sigma * sigma)) for i in range(2 * r + 1)])
    w = w / w.sum()
    in_i, in_s = orc.clipped(*orc.padded(list(req[0]), list(req[1]), [r] * D), [0] * D, list(size))
    mats = [orc.axis_matrix(w, req[0][a], req[1][a], in_i[a], in_s[a]) for a in range(D)]
    sub = x[tuple(slice(i, i + s) for i, s in zip(reversed(in_i), reversed(in_s)))]
    blur = orc.apply_dense(mats, sub)
    orig = x[tuple(slice(i, i + s) for i, s in zip(reversed(req[0]), reversed(req[1])))]
    edge = orig - blur
    want = np.where(np.abs(edge) > thr, orig + amount * edge, orig)
    near = np.abs(np.abs(edge) - thr) < 1e-9     # a threshold tie could go either way
    assert np.allclose(out[~near], want[~near], rtol=0, atol=1e-12)


def test_unsharp_subregion_is_crop_of_whole(orc):
    # each pass of the reference convolves the whole padded buffer; for the
    # requested pixels that equals the Gaussian path's shrinking pass regions,
    # so a sub-request is a crop of the whole-image result, bit for bit
    x = rand((20, 22, 24))
    full = orc.unsharp_generate_region(x, [0, 0, 0], 1.3, 0.6, 0.01, [0, 0, 0], [24, 22, 20])
    for req in (((0, 0, 0), (10, 22, 5)), ((7, 3, 9), (9, 11, 6)), ((20, 18, 16), (4, 4, 4))):
        sub = orc.unsharp_generate_region(x, [0, 0, 0], 1.3, 0.6, 0.01, list(req[0]), list(req[1]))
        crop = full[tuple(slice(i, i + s) for i, s in zip(reversed(req[0]), reversed(req[1])))]
        assert np.array_equal(sub, crop)


# --- Recursive Gaussian (recursive_gaussian.rs:33-240) ------------------------

def _impulse(n):
    x = np.zeros(n, np.float32)
    x[n // 2] = 1.0
    return x


def test_recursive_reference_tests(orc):
    # recursive_gaussian.rs:252-300
    out = orc.recursive_generate_region(_impulse(51), [0], [1.0], 2.0, -1, [0], [51])
    assert abs(float(out.sum()) - 1.0) < 0.02
    assert out[25] > out[28]
    c = np.full((15, 15), 4.0, np.float32)
    out = orc.recursive_generate_region(c, [0, 0], [1.0, 1.0], 1.5, -1, [0, 0], [15, 15])
    assert np.all(np.abs(out[3:12, 3:12] - 4.0) < 0.05)
    out = orc.recursive_generate_region(_impulse(31), [0], [1.0], 2.0, 0, [0], [31])
    assert abs(float(out.sum()) - 1.0) < 0.02


@pytest.mark.parametrize("sigma", [0.1, 0.5, 1.0, 2.4, 2.5, 3.7, 10.0])
def test_yvv_coefficients_clamp(orc, sigma):
    c = orc.yvv_coeffs(sigma)
    if sigma < 0.5:
        assert np.array_equal(c, orc.yvv_coeffs(0.5))  # clamped below at 0.5
    assert np.array_equal(orc.yvv_coeffs(float("nan")), orc.yvv_coeffs(0.5))  # f64::max drops NaN


# Young & van Vliet (Signal Processing 44, 1995) design the recursion from
# three poles of the scaled Gaussian, published as m0 = 1.16680 (real) and
# m1 +- i m2 = 1.10783 +- 1.40586 i; the denominator 1 - a1 x - a2 x^2 - a3 x^3
# (x = z^-1) vanishes where q (1 - x) = -m0 and -m1 +- i m2, with q the scale of
# the q(sigma) rule (q = 0.98711 sigma - 0.96330 from sigma 2.5 up, else
# 3.97156 - 4.14554 sqrt(1 - 0.26891 sigma)).  The poles recovered from the
# oracle's a_k therefore pin every constant of the b polynomials against the
# published pole constants, and the scale pins the q branch.  The b
# constants are rounded to about six digits, so the agreement loosens in
# proportion to q.
YVV_M0, YVV_M1, YVV_M2 = 1.16680, 1.10783, 1.40586


@pytest.mark.parametrize("sigma", [0.3, 0.6, 1.0, 1.7, 2.4, 2.5, 3.3, 5.0, 8.0])
def test_yvv_poles_are_the_published_ones(orc, sigma):
    B, a1, a2, a3 = orc.yvv_coeffs(sigma)
    v = 1.0 - np.roots([-a3, -a2, -a1, 1.0])          # q (1 - x) / q at the poles
    real = v[np.abs(v.imag) < 1e-9].real
    cplx = v[v.imag > 1e-9]
    assert real.size == 1 and cplx.size == 1
    q = -YVV_M0 / real[0]
    tol = 5e-5 + 7e-5 * q
    assert abs(cplx[0].real * q + YVV_M1) < tol
    assert abs(abs(cplx[0].imag) * q - YVV_M2) < tol
    s = max(sigma, 0.5)
    q_rule = 0.98711 * s - 0.96330 if s >= 2.5 else 3.97156 - 4.14554 * math.sqrt(1 - 0.26891 * s)
    assert abs(q - q_rule) < tol * q_rule
    # the gain B makes the response sum to one: B = 1 - a1 - a2 - a3 follows
    # from the b polynomials only up to their rounding (6e-6 of b0 q^2)
    assert abs(B - (1.0 - a1 - a2 - a3)) < 1e-12


def _yvv_closed_forms(c):
    """Variance and centre value of the zero-phase (causal then anticausal)
    impulse response on an unbounded line, from the coefficients' transfer
    function H(w) = B^2 / |1 - sum_k a_k e^{-ikw}|^2: the causal factor
    G(x) = B / (1 - sum a_k x^k) has mean m = sum k a_k / B and variance
    sum k^2 a_k / B + m^2 (its first two derivatives at x = 1), the two
    passes double the variance, and the centre value is the mean of H over a
    period (trapezoid rule, exact to rounding for this smooth periodic H)."""
    B, a = c[0], np.asarray(c[1:])
    k = np.arange(1, 4)
    m = (k * a).sum() / B
    var = 2.0 * ((k * k * a).sum() / B + m * m)
    w = 2.0 * np.pi * np.arange(1 << 16) / (1 << 16)
    A = 1.0 - (a[None, :] * np.exp(-1j * np.outer(w, k))).sum(axis=1)
    return var, float((B * B / np.abs(A) ** 2).mean())


@pytest.mark.parametrize("sigma", [0.6, 1.0, 1.7, 2.4, 2.5, 3.3, 5.0, 8.0])
def test_recursive_impulse_matches_transfer_function(orc, sigma):
    # an impulse in the middle of a line long enough that the boundary
    # extension never reaches it (the slowest pole decays by 1 / (1 + m0 / q)
    # per sample): the oracle's recursion has the closed-form variance and
    # centre value of its coefficients, and unit sum
    n = 2001
    x = np.zeros(n)
    x[n // 2] = 1.0
    h = orc.recursive_generate_region(x, [0], [1.0], sigma, -1, [0], [n])
    i = np.arange(n) - n // 2
    var, peak = _yvv_closed_forms(orc.yvv_coeffs(sigma))
    assert abs(h.sum() - 1.0) < 1e-12
    assert abs((h * i * i).sum() - var) < 1e-11 * var
    assert abs(h[n // 2] - peak) < 1e-11 * peak
    assert abs((h * i).sum()) < 1e-10                  # zero phase: no shift


@pytest.mark.parametrize("sigma,lo,hi", [(1.0, 1.08, 1.30), (1.5, 1.08, 1.30), (2.0, 1.08, 1.30),
                                         (3.0, 1.05, 1.20), (5.0, 1.05, 1.20)])
def test_recursive_impulse_is_gaussian_like(orc, sigma, lo, hi):
    # the third-order recursion approximates a Gaussian: unit sum, symmetric,
    # peak within 10% of a sampled Gaussian's, and an effective width a little
    # above sigma (1.24 sigma at 1, 1.11 sigma at 5 for this form); a wrong
    # branch of the q formula lands far outside these bands
    n = 161
    x = np.zeros(n)
    x[80] = 1.0
    h = orc.recursive_generate_region(x, [0], [1.0], sigma, -1, [0], [n])
    g = np.exp(-0.5 * ((np.arange(n) - 80) / sigma) ** 2)
    g /= g.sum()
    assert abs(h.sum() - 1.0) < 1e-9
    assert np.allclose(h, h[::-1], atol=1e-6 * h.max())   # zero-phase up to rounding and the ends
    assert np.abs(h - g).max() < 0.1 * g.max()
    i = np.arange(n)
    sd = math.sqrt(float((h * (i - 80) ** 2).sum()))
    assert lo * sigma < sd < hi * sigma


@pytest.mark.parametrize("shape,spacing,sigma,axis", [
    ((13, 9, 7), (1.0, 1.0, 1.0), 1.5, -1),
    ((12, 10), (0.5, 2.0), 2.0, -1),
    ((8, 11, 6), (1.0, 1.0, 1.0), 3.0, 1),
    ((40,), (1.0,), 0.3, -1),
    ((1, 10, 5), (1.0, 1.0, 1.0), 1.0, -1),      # an axis of one sample is left alone
])
def test_recursive_c_vs_numpy(orc, shape, spacing, sigma, axis):
    D = len(shape)
    x = rand(tuple(reversed(shape)), dtype=np.float64)
    a = orc.recursive_generate_region(x, [0] * D, list(spacing), sigma, axis, [0] * D, list(shape))
    b = orc.np_recursive(x, list(spacing), sigma, axis)
    assert np.array_equal(a, b)
    x32 = x.astype(np.float32)
    a32 = orc.recursive_generate_region(x32, [0] * D, list(spacing), sigma, axis, [0] * D, list(shape))
    assert np.array_equal(a32, orc.np_recursive(x32, list(spacing), sigma, axis).astype(np.float32))


# --- CannyEdgeDetectionFilter (edges.rs:548-664) -------------------------------

@pytest.mark.parametrize("seed,sigma,lower,upper", [(1, 1.0, 0.05, 0.15), (2, 1.5, 0.02, 0.1),
                                                    (3, 0.8, 0.1, 0.3)])
def test_canny_c_vs_numpy(orc, seed, sigma, lower, upper):
    # independent: the blur as the dense f64 operator of the index-space
    # kernel (whole-image passes clamp at the image border), then np_canny;
    # random data leave no magnitude or threshold ties for rounding to flip
    rng = np.random.default_rng(seed)
    x = rng.random((40, 52)).astype(np.float64)
    got = orc.canny_generate_region(x, [0, 0], sigma, lower, upper, [0, 0], [52, 40])
    r = math.ceil(3.0 * sigma)
    w = np.array([math.exp(-((i - r) ** 2) / (2.0 * sigma * sigma)) for i in range(2 * r + 1)])
    w = w / w.sum()
    blur = orc.apply_dense([orc.axis_matrix(w, 0, 52, 0, 52), orc.axis_matrix(w, 0, 40, 0, 40)], x)
    assert np.array_equal(got, orc.np_canny(blur, lower, upper))


def test_canny_step_and_constant(orc):
    x = np.zeros((30, 40), np.float32)
    x[:, 20:] = 1.0                                   # a vertical step between x = 19 and 20
    out = orc.canny_generate_region(x, [0, 0], 1.0, 0.1, 0.3, [0, 0], [40, 30])
    cols = np.nonzero(out.any(axis=0))[0]
    assert set(cols.tolist()) <= {19, 20} and len(cols) >= 1
    assert out[:, cols[0]].all()                      # the whole edge column is found
    c = np.full((30, 40), 3.0, np.float32)
    assert not orc.canny_generate_region(c, [0, 0], 1.0, 0.1, 0.3, [0, 0], [40, 30]).any()
    # upper above every magnitude: no strong pixel, no edges
    assert not orc.canny_generate_region(x, [0, 0], 1.0, 0.1, 1e9, [0, 0], [40, 30]).any()


def test_canny_subregion_is_crop(orc):
    rng = np.random.default_rng(9)
    x = rng.random((33, 45)).astype(np.float32)
    full = orc.canny_generate_region(x, [0, 0], 1.2, 0.05, 0.12, [0, 0], [45, 33])
    sub = orc.canny_generate_region(x, [0, 0], 1.2, 0.05, 0.12, [7, 4], [20, 15])
    assert np.array_equal(sub, full[4:19, 7:27])
