"""Pins the CPU oracle (oracle/gauss_oracle.c) before anything is checked
against it.

Each check is chosen so that a plausible mistake in the restatement fails
it: an independent numpy float32 restatement (bit-exact), scipy's
clamp-to-edge correlation (tolerance), the interior-impulse closed form
(bit-exact), the dense per-axis operator (brute force), the reference's own
property tests (src/filters/gaussian.rs:107-127, tests/pipeline.rs:37-129),
and special cases of sigma.
"""
import math

import numpy as np
import pytest

RNG = np.random.default_rng(1234)


def rand(shape, ncomp=1, dtype=np.float32):
    shp = tuple(shape) + ((ncomp,) if ncomp > 1 else ())
    return RNG.random(shp).astype(dtype)


# --- kernel_radius / build_kernel (src/filters/gaussian.rs:24-40) -----------

@pytest.mark.parametrize("sigma,spacing,r", [
    (1.5, 1.0, 5), (2.0, 1.0, 6), (3.0, 0.7, 13), (3.0, 2.5, 4), (1.8, 1.0, 6),
    (1.2, 1.0, 4), (1.0, 1.0, 3), (0.5, 2.0, 1), (4.0, 0.5, 24), (0.1, 1.0, 1),
])
def test_radius_config_table(orc, sigma, spacing, r):
    # radii of the five configurations and the batched extremes
    assert orc.kernel_radius(sigma, spacing) == r


def test_radius_special_cases(orc):
    assert orc.kernel_radius(-1.0, 1.0) == 0           # negative -> 0 (saturating cast)
    assert orc.kernel_radius(float("nan"), 1.0) == 0   # NaN -> 0
    assert orc.kernel_radius(0.0, 1.0) == 0
    assert orc.kernel_radius(1.0, 3.0) == 1            # 3 * (1/3) rounds to exactly 1.0
    # evaluation order 3*(s/sp), not (3s)/sp: pick sp so the two roundings
    # straddle an integer
    hits = 0
    for k in range(1, 2000):
        s, sp = 0.1 * k, 0.3
        a = math.ceil(3.0 * (s / sp))
        b = math.ceil((3.0 * s) / sp)
        if a != b:
            hits += 1
            assert orc.kernel_radius(s, sp) == a
    assert hits > 0


@pytest.mark.parametrize("sigma_px,r", [(1.5, 5), (2.0, 6), (4.2857142857142865, 13), (0.3, 1)])
def test_kernel_closed_form(orc, sigma_px, r):
    w = orc.build_kernel(sigma_px, r)
    assert len(w) == 2 * r + 1
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.array_equal(w, w[::-1])                       # exactly symmetric
    assert np.argmax(w) == r
    # neighbour ratio exp(-0.5 * ((k+1)^2 - k^2) / s^2)
    for k in range(r):
        ratio = w[r + k + 1] / w[r + k]
        assert abs(ratio - math.exp(-0.5 * (2 * k + 1) / sigma_px ** 2)) < 1e-13


# --- generate_region ---------------------------------------------------------

CASES = [
    # (numpy shape, img_index, spacing, sigma, req_index, req_size) ; sizes are axis-0-first
    ((9, 13), [0, 0], [1.0, 1.0], 1.5, [0, 0], [13, 9]),
    ((9, 13), [2, -1], [1.0, 0.7], 1.0, [4, 0], [6, 5]),
    ((7, 8, 9), [0, 0, 0], [1.0, 1.0, 1.0], 2.0, [0, 0, 0], [9, 8, 7]),
    ((7, 8, 9), [1, 2, 3], [0.7, 0.7, 2.5], 1.2, [2, 3, 4], [5, 4, 3]),
    ((3, 2, 4), [0, 0, 0], [1.0, 1.0, 1.0], 2.5, [0, 0, 0], [4, 2, 3]),  # extent < 2r+1
    ((1, 1, 5), [0, 0, 0], [1.0, 1.0, 1.0], 1.0, [0, 0, 0], [5, 1, 1]),
    ((17,), [5], [1.0], 1.5, [7], [6]),
    ((4, 3, 5, 6), [0, 0, 0, 0], [1.0, 1.0, 1.0, 2.0], 1.0, [1, 0, 1, 0], [4, 3, 2, 3]),
]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("ncomp", [1, 3])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_c_oracle_matches_numpy_restatement(orc, case, ncomp, dtype):
    shape, ii, sp, sigma, ri, rs = case
    img = rand(shape, ncomp, dtype)
    a = orc.generate_region(img, ii, sp, sigma, ri, rs, ncomp)
    b = orc.np_generate_region(img, ii, sp, sigma, ri, rs, ncomp)
    assert a.shape == b.shape
    assert np.array_equal(a, b)
    assert np.array_equal(np.signbit(a), np.signbit(b))


@pytest.mark.parametrize("case", CASES)
def test_c_oracle_matches_dense_operator(orc, case):
    shape, ii, sp, sigma, ri, rs = case
    img = rand(shape, 1, np.float64)
    size = list(reversed(shape))
    mats, in_i, in_s = orc.dense_operator(ii, size, sp, sigma, ri, rs)
    sl = tuple(slice(in_i[d] - ii[d], in_i[d] - ii[d] + in_s[d]) for d in reversed(range(len(sp))))
    y = orc.apply_dense(mats, img[sl])
    a64 = orc.generate_region(img, ii, sp, sigma, ri, rs)
    assert np.abs(y - a64).max() <= 1e-12
    a32 = orc.generate_region(img.astype(np.float32), ii, sp, sigma, ri, rs)
    assert np.abs(y - a32).max() <= 1e-5 * max(1.0, np.abs(y).max())


def test_c_oracle_matches_scipy_nearest(orc):
    ndi = pytest.importorskip("scipy.ndimage")
    img = rand((12, 15, 18), 1, np.float32)
    sp = [1.0, 0.8, 1.6]
    sigma = 1.7
    out = orc.generate_region(img, [0, 0, 0], sp, sigma, [0, 0, 0], [18, 15, 12])
    ref = img.astype(np.float64)
    for a in range(3):
        r = orc.kernel_radius(sigma, sp[a])
        w = orc.build_kernel(sigma / sp[a], r).astype(np.float32).astype(np.float64)
        ref = ndi.correlate1d(ref, w, axis=2 - a, mode="nearest")
    assert np.abs(out - ref).max() < 1e-6


def test_interior_impulse_closed_form(orc):
    n = 31
    img = np.zeros((n, n, n), np.float32)
    img[15, 15, 15] = 1.0
    sp = [1.0, 0.9, 1.3]
    sigma = 1.4
    out = orc.generate_region(img, [0, 0, 0], sp, sigma, [0, 0, 0], [n, n, n])
    w = [orc.build_kernel(sigma / s, orc.kernel_radius(sigma, s)).astype(np.float32) for s in sp]
    r = [orc.kernel_radius(sigma, s) for s in sp]
    exp = np.zeros_like(out)
    for k in range(-r[2], r[2] + 1):
        for j in range(-r[1], r[1] + 1):
            for i in range(-r[0], r[0] + 1):
                v = np.float32(w[0][r[0] - i]) * np.float32(w[1][r[1] - j])
                exp[15 + k, 15 + j, 15 + i] = np.float32(v) * np.float32(w[2][r[2] - k])
    assert np.array_equal(out.view(np.uint32), exp.view(np.uint32))


# --- the reference's own property tests ---------------------------------------

def _impulse(size, D=2):
    img = np.zeros((size,) * D, np.float32)
    img[(size // 2,) * D] = 1.0
    return img


def test_reference_impulse_2d(orc):
    # src/filters/gaussian.rs:107-127, tests/pipeline.rs:37-54
    out = orc.generate_region(_impulse(21), [0, 0], [1.0, 1.0], 1.5, [0, 0], [21, 21])
    assert abs(float(out.sum(dtype=np.float32)) - 1.0) < 0.01
    assert (out <= out[10, 10] + 1e-6).all()


def test_reference_gaussian_3d(orc):
    # tests/pipeline.rs:60-72
    out = orc.generate_region(_impulse(11, 3), [0, 0, 0], [1.0] * 3, 1.0, [0] * 3, [11] * 3)
    assert abs(float(out.sum(dtype=np.float32)) - 1.0) < 0.05


def test_reference_vector_pixel(orc):
    # tests/pipeline.rs:78-100
    img = np.zeros((11, 11, 3), np.float32)
    img[5, 5] = [1.0, 2.0, 3.0]
    out = orc.generate_region(img, [0, 0], [1.0, 1.0], 1.0, [0, 0], [11, 11], ncomp=3)
    for c, want in enumerate([1.0, 2.0, 3.0]):
        assert abs(float(out[..., c].sum(dtype=np.float32)) - want) < 0.05


def _stream(orc, img, ii, sp, sigma, pieces):
    D = len(sp)
    size = list(reversed(img.shape[:D]))
    out = np.zeros_like(img)
    slow = D - 1
    n = min(pieces, size[slow])
    base, rem = divmod(size[slow], n)
    start = ii[slow]
    for p in range(n):
        sz = base + (1 if p < rem else 0)
        ri = list(ii)
        rs = list(size)
        ri[slow], rs[slow] = start, sz
        part = orc.generate_region(img, ii, sp, sigma, ri, rs)
        out[start - ii[slow]:start - ii[slow] + sz] = part
        start += sz
    return out


def test_reference_streaming_matches_full(orc):
    # tests/pipeline.rs:106-129 -- and bit-identical, not just within 1e-6
    img = _impulse(20)
    full = orc.generate_region(img, [0, 0], [1.0, 1.0], 1.5, [0, 0], [20, 20])
    st = _stream(orc, img, [0, 0], [1.0, 1.0], 1.5, 4)
    assert np.array_equal(full.view(np.uint32), st.view(np.uint32))
    img3 = rand((23, 9, 10))
    full3 = orc.generate_region(img3, [0, 0, 0], [1.0, 1.0, 1.0], 2.0, [0, 0, 0], [10, 9, 23])
    st3 = _stream(orc, img3, [0, 0, 0], [1.0, 1.0, 1.0], 2.0, 8)
    assert np.array_equal(full3.view(np.uint32), st3.view(np.uint32))


# --- sigma special cases -----------------------------------------------------

def test_negative_sigma_is_identity(orc):
    img = rand((6, 7))
    img[0, 0] = -0.0
    out = orc.generate_region(img, [0, 0], [1.0, 1.0], -1.0, [0, 0], [7, 6])
    assert np.array_equal(out, img)          # values
    assert not np.signbit(out[0, 0])         # 0 + (-0 * 1) = +0


def test_zero_sigma_is_nan(orc):
    out = orc.generate_region(rand((5, 5)), [0, 0], [1.0, 1.0], 0.0, [0, 0], [5, 5])
    assert np.isnan(out).all()


def test_constant_image_stays_constant(orc):
    img = np.full((9, 10, 11), 0.37, np.float32)
    out = orc.generate_region(img, [0, 0, 0], [1.0, 1.0, 1.0], 2.0, [0, 0, 0], [11, 10, 9])
    assert np.abs(out - np.float32(0.37)).max() <= 13 * 3 * np.finfo(np.float32).eps


# --- adjoint (f64) -----------------------------------------------------------

@pytest.mark.parametrize("case", CASES)
def test_adjoint_is_dense_transpose(orc, case):
    shape, ii, sp, sigma, ri, rs = case
    size = list(reversed(shape))
    D = len(sp)
    g = RNG.standard_normal(tuple(reversed(rs)))
    gin, in_i, in_s = orc.adjoint_f64(g, ii, size, sp, sigma, ri, rs)
    mats, ci, cs = orc.dense_operator(ii, size, sp, sigma, ri, rs)
    assert (ci, cs) == (in_i, in_s)
    ref = orc.apply_dense(mats, g, transpose=True)
    assert np.abs(gin - ref).max() <= 1e-12
    # <A x, g> == <x, A^T g>
    x = RNG.standard_normal(tuple(reversed(in_s)))
    full = np.zeros(tuple(reversed(size)))
    sl = tuple(slice(in_i[d] - ii[d], in_i[d] - ii[d] + in_s[d]) for d in reversed(range(D)))
    full[sl] = x
    ax = orc.generate_region(full, ii, sp, sigma, ri, rs)
    assert abs(float((ax * g).sum()) - float((x * gin).sum())) < 1e-10


# --- the threaded ref-faithful variant (CPU baseline) --------------------------

@pytest.mark.parametrize("ncomp", [1, 3])
def test_ref_faithful_is_bit_identical(orc, ncomp):
    img = rand((14, 17, 19), ncomp)
    a = orc.generate_region(img, [0, 0, 0], [1.0, 0.8, 1.5], 1.6, [2, 1, 0], [15, 14, 14], ncomp)
    b = orc.ref_faithful(img, [0, 0, 0], [1.0, 0.8, 1.5], 1.6, [2, 1, 0], [15, 14, 14], ncomp, 4)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("shape,spacing,sigma", [((19, 23, 41), (1.0, 1.0, 1.0), 2.0),
                                                 ((12, 30, 17), (0.7, 1.3, 2.5), 1.6),
                                                 ((5, 4, 70), (1.0, 1.0, 1.0), 3.3)])
def test_separable_opt_is_bit_identical(orc, shape, spacing, sigma):
    """The bench's optimised CPU comparison (whole-image 3-D f32, rows along
    x in every pass, threads over planes) gives the oracle's bits, extents
    smaller than the kernel included."""
    rng = np.random.default_rng(sum(shape))
    img = rng.random(shape).astype(np.float32)
    nz, ny, nx = shape
    a = orc.generate_region(img, [0, 0, 0], list(spacing), sigma, [0, 0, 0], [nx, ny, nz])
    b = orc.separable_opt(img, list(spacing), sigma, 3)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
