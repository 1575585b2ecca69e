"""CPU oracle for the itk-rs separable Gaussian -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check in __graft_entry__.py and bench.py's
cpu_baseline leg may import this module, and only as the checker (or as the
timed CPU baseline).  The product path never imports it.

Three independent restatements live here:

* ``generate_region`` / ``adjoint_f64`` / ``ref_faithful``: ctypes calls into
  ``liboracle.so`` (gauss_oracle.c), the loop-for-loop C restatement of
  src/filters/gaussian.rs:24-90 and src/filters/conv.rs:11-41.
* ``np_generate_region``: a numpy float32 restatement written independently
  (vectorised over every axis but the pass axis, taps in order k = 0..2r,
  separate multiply and add).  It must agree with the C oracle to the bit.
* ``dense_operator``: the explicit f64 matrix of one clamp-to-edge pass
  (A[i, clamp(i + off + k - r)] += w_k), applied as a Kronecker product --
  the brute-force check for the forward and the explicit transpose for the
  adjoint.

The other convolve_axis filters: ``mean_generate_region`` and
``binomial_generate_region`` (C restatements of src/filters/mean.rs:59-83,
126-144 and src/filters/binomial_blur.rs:57-89), with
``dense_mean_operator`` / ``dense_binomial_operator`` as their brute-force
checks; ``unsharp_generate_region`` restates src/filters/edges.rs:406-452.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i64p = ctypes.POINTER(ctypes.c_int64)
        f64p = ctypes.POINTER(ctypes.c_double)
        L.orc_kernel_radius.restype = ctypes.c_uint64
        L.orc_kernel_radius.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_build_kernel.restype = None
        L.orc_build_kernel.argtypes = [ctypes.c_double, ctypes.c_uint64, f64p]
        for name in ("orc_generate_region_f32", "orc_generate_region_f64"):
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, i64p, i64p, f64p,
                           ctypes.c_double, i64p, i64p, ctypes.c_void_p]
        L.orc_ref_faithful_f32.restype = ctypes.c_int
        L.orc_ref_faithful_f32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, i64p, i64p,
                                           f64p, ctypes.c_double, i64p, i64p, ctypes.c_void_p,
                                           ctypes.c_int]
        L.orc_separable_opt_f32.restype = ctypes.c_int
        L.orc_separable_opt_f32.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.c_int64, f64p, ctypes.c_double,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.orc_box_kernel.restype = None
        L.orc_box_kernel.argtypes = [ctypes.c_uint64, f64p]
        for name in ("orc_mean_generate_region_f32", "orc_mean_generate_region_f64"):
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, i64p, i64p, u64p,
                           i64p, i64p, ctypes.c_void_p]
        for name in ("orc_binomial_generate_region_f32", "orc_binomial_generate_region_f64"):
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, i64p, i64p, ctypes.c_uint64,
                           i64p, i64p, ctypes.c_void_p]
        L.orc_index_kernel.restype = None
        L.orc_index_kernel.argtypes = [ctypes.c_double, ctypes.c_uint64, f64p]
        for name in ("orc_unsharp_generate_region_f32", "orc_unsharp_generate_region_f64"):
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_int, ctypes.c_void_p, i64p, i64p, ctypes.c_double,
                           ctypes.c_double, ctypes.c_double, i64p, i64p, ctypes.c_void_p]
        L.orc_yvv_coeffs.restype = None
        L.orc_yvv_coeffs.argtypes = [ctypes.c_double, f64p]
        for name in ("orc_recursive_generate_region_f32", "orc_recursive_generate_region_f64"):
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_int, ctypes.c_void_p, i64p, i64p, f64p, ctypes.c_double,
                           ctypes.c_int, i64p, i64p, ctypes.c_void_p]
        for name in ("orc_canny_generate_region_f32", "orc_canny_generate_region_f64"):
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_void_p, i64p, i64p, ctypes.c_double, ctypes.c_double,
                           ctypes.c_double, i64p, i64p, ctypes.c_void_p]
        L.orc_canny_bins.restype = None
        L.orc_canny_bins.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p]
        L.orc_adjoint_f64.restype = ctypes.c_int
        L.orc_adjoint_f64.argtypes = [ctypes.c_int, ctypes.c_int, i64p, i64p, f64p, ctypes.c_double,
                                      i64p, i64p, ctypes.c_void_p, ctypes.c_void_p, i64p, i64p]
        _lib = L
    return _lib


def _i64(v):
    a = (ctypes.c_int64 * len(v))(*[int(x) for x in v])
    return a


def _f64(v):
    return (ctypes.c_double * len(v))(*[float(x) for x in v])


# ---------------------------------------------------------------------------
# region algebra (src/image.rs:28-64)
# ---------------------------------------------------------------------------

def kernel_radius(sigma: float, spacing: float) -> int:
    """src/filters/gaussian.rs:24-27 (via the C oracle)."""
    return int(lib().orc_kernel_radius(float(sigma), float(spacing)))


def build_kernel(sigma_px: float, radius: int) -> np.ndarray:
    """src/filters/gaussian.rs:29-40 (via the C oracle), f64 weights."""
    out = np.zeros(2 * radius + 1, dtype=np.float64)
    lib().orc_build_kernel(float(sigma_px), int(radius),
                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def padded(index, size, radii):
    return ([i - r for i, r in zip(index, radii)], [s + 2 * r for s, r in zip(size, radii)])


def clipped(index, size, bidx, bsize):
    ni, ns = [], []
    for i, s, bi, bs in zip(index, size, bidx, bsize):
        st = max(i, bi)
        en = min(i + s, bi + bs)
        ni.append(st)
        ns.append(max(en - st, 0))
    return ni, ns


def _shape(size, ncomp):
    """numpy shape of a dense region buffer: axis 0 fastest -> reversed."""
    shp = tuple(int(s) for s in reversed(size))
    return shp + ((ncomp,) if ncomp > 1 else ())


# ---------------------------------------------------------------------------
# C oracle entry points
# ---------------------------------------------------------------------------

def generate_region(img: np.ndarray, img_index, spacing, sigma, req_index, req_size,
                    ncomp: int = 1) -> np.ndarray:
    """GaussianFilter{source: img, sigma}.generate_region(req) via the C oracle.

    ``img`` is a dense numpy array in numpy order (slowest axis first), with a
    trailing component axis when ncomp > 1.  The dtype selects f32 or f64.
    """
    D = len(spacing)
    img = np.ascontiguousarray(img)
    img_size = list(reversed(img.shape[:D]))
    fn = lib().orc_generate_region_f32 if img.dtype == np.float32 else lib().orc_generate_region_f64
    if img.dtype not in (np.float32, np.float64):
        raise TypeError("oracle supports float32 and float64 pixels")
    out = np.zeros(_shape(req_size, ncomp), dtype=img.dtype)
    rc = fn(D, ncomp, img.ctypes.data, _i64(img_index), _i64(img_size), _f64(spacing), float(sigma),
            _i64(req_index), _i64(req_size), out.ctypes.data)
    if rc != 0:
        raise ValueError("oracle rejected arguments")
    return out


def box_kernel(radius: int) -> np.ndarray:
    """src/filters/mean.rs:15-18 (via the C oracle)."""
    out = np.zeros(2 * radius + 1, dtype=np.float64)
    lib().orc_box_kernel(int(radius), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def _filter_call(stem, img, img_index, arg, req_index, req_size, ncomp):
    img = np.ascontiguousarray(img)
    if img.dtype not in (np.float32, np.float64):
        raise TypeError("oracle supports float32 and float64 pixels")
    D = len(img_index)
    img_size = list(reversed(img.shape[:D]))
    fn = getattr(lib(), stem + ("_f32" if img.dtype == np.float32 else "_f64"))
    out = np.zeros(_shape(req_size, ncomp), dtype=img.dtype)
    rc = fn(D, ncomp, img.ctypes.data, _i64(img_index), _i64(img_size), arg, _i64(req_index),
            _i64(req_size), out.ctypes.data)
    if rc != 0:
        raise ValueError("oracle rejected arguments")
    return out


def mean_generate_region(img: np.ndarray, img_index, radii, req_index, req_size,
                         ncomp: int = 1) -> np.ndarray:
    """BoxMeanFilterN{source: img, radii}.generate_region(req) (MeanFilter:
    equal radii) via the C oracle (src/filters/mean.rs:59-83, 126-144)."""
    arr = (ctypes.c_uint64 * len(radii))(*[int(r) for r in radii])
    return _filter_call("orc_mean_generate_region", img, img_index, arr, req_index, req_size, ncomp)


def binomial_generate_region(img: np.ndarray, img_index, reps, req_index, req_size,
                             ncomp: int = 1) -> np.ndarray:
    """BinomialBlurFilter{source: img, repetitions}.generate_region(req) via
    the C oracle (src/filters/binomial_blur.rs:57-89)."""
    return _filter_call("orc_binomial_generate_region", img, img_index, int(reps), req_index,
                        req_size, ncomp)


def index_kernel(sigma: float) -> np.ndarray:
    """src/filters/edges.rs:414-423 (via the C oracle): f64 weights, radius
    ceil(3 sigma)."""
    r = kernel_radius(sigma, 1.0)
    out = np.zeros(2 * r + 1, dtype=np.float64)
    lib().orc_index_kernel(float(sigma), r, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def unsharp_generate_region(img: np.ndarray, img_index, sigma, amount, threshold, req_index,
                            req_size) -> np.ndarray:
    """UnsharpMaskFilter{source: img, sigma, amount, threshold}.generate_region(req)
    via the C oracle (src/filters/edges.rs:406-452), scalar pixels."""
    img = np.ascontiguousarray(img)
    D = len(img_index)
    img_size = list(reversed(img.shape[:D]))
    fn = lib().orc_unsharp_generate_region_f32 if img.dtype == np.float32 else \
        lib().orc_unsharp_generate_region_f64
    out = np.zeros(_shape(req_size, 1), dtype=img.dtype)
    rc = fn(D, img.ctypes.data, _i64(img_index), _i64(img_size), float(sigma), float(amount),
            float(threshold), _i64(req_index), _i64(req_size), out.ctypes.data)
    if rc != 0:
        raise ValueError("oracle rejected arguments")
    return out


def yvv_coeffs(sigma: float) -> np.ndarray:
    """src/filters/recursive_gaussian.rs:33-47 (via the C oracle): B, a1, a2, a3."""
    out = np.zeros(4, dtype=np.float64)
    lib().orc_yvv_coeffs(float(sigma), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def recursive_generate_region(img: np.ndarray, img_index, spacing, sigma, axis, req_index,
                              req_size) -> np.ndarray:
    """RecursiveGaussianFilter (axis -1) / RecursiveGaussianAxisFilter (axis)
    .generate_region(req) via the C oracle (recursive_gaussian.rs:147-240)."""
    img = np.ascontiguousarray(img)
    D = len(img_index)
    img_size = list(reversed(img.shape[:D]))
    fn = lib().orc_recursive_generate_region_f32 if img.dtype == np.float32 else \
        lib().orc_recursive_generate_region_f64
    out = np.zeros(_shape(req_size, 1), dtype=img.dtype)
    rc = fn(D, img.ctypes.data, _i64(img_index), _i64(img_size), _f64(spacing), float(sigma),
            int(axis), _i64(req_index), _i64(req_size), out.ctypes.data)
    if rc != 0:
        raise ValueError("oracle rejected arguments")
    return out


def np_recursive(img: np.ndarray, spacing, sigma, axis) -> np.ndarray:
    """Independent numpy restatement of the recursive Gaussian on the whole
    image (f64 result): every line along an axis at once, positions in turn,
    the same products and sums in the same order."""
    D = len(spacing)
    y = np.asarray(img, dtype=np.float64).copy()
    for a in range(D):
        if axis >= 0 and a != axis:
            continue
        ax = D - 1 - a                      # numpy axis of image axis a
        n = y.shape[ax]
        if n <= 1:
            continue
        B, a1, a2, a3 = (float(v) for v in yvv_coeffs(sigma / spacing[a]))
        v = np.moveaxis(y, ax, 0)           # a view: writes go to y
        for order in (range(n), range(n - 1, -1, -1)):
            first = next(iter(order))
            ym1 = v[first].copy()
            ym2 = ym1.copy()
            ym3 = ym1.copy()
            for i in order:
                cur = B * v[i] + a1 * ym1 + a2 * ym2 + a3 * ym3
                ym3, ym2, ym1 = ym2, ym1, cur
                v[i] = cur
    return y


def canny_generate_region(img: np.ndarray, img_index, sigma, lower, upper, req_index,
                          req_size) -> np.ndarray:
    """CannyEdgeDetectionFilter{source: img, sigma, lower, upper}.generate_region(req)
    via the C oracle (src/filters/edges.rs:548-664), 2-D scalar."""
    img = np.ascontiguousarray(img)
    img_size = list(reversed(img.shape[:2]))
    fn = lib().orc_canny_generate_region_f32 if img.dtype == np.float32 else \
        lib().orc_canny_generate_region_f64
    out = np.zeros(_shape(req_size, 1), dtype=img.dtype)
    rc = fn(img.ctypes.data, _i64(img_index), _i64(img_size), float(sigma), float(lower),
            float(upper), _i64(req_index), _i64(req_size), out.ctypes.data)
    if rc != 0:
        raise ValueError("oracle rejected arguments")
    return out


def canny_bins(gx: np.ndarray, gy: np.ndarray) -> np.ndarray:
    """Direction bins (0: 0, 1: 45, 2: 90, 3: 135 degrees) of the suppression
    step of CannyEdgeDetectionFilter (src/filters/edges.rs:602, 612-620) for
    f64 gradients, with the host libm's atan2 as in the reference."""
    gx = np.ascontiguousarray(gx, dtype=np.float64)
    gy = np.ascontiguousarray(gy, dtype=np.float64)
    out = np.empty(gx.shape, dtype=np.uint8)
    lib().orc_canny_bins(gx.size, gx.ctypes.data, gy.ctypes.data, out.ctypes.data)
    return out


def np_canny(smoothed: np.ndarray, lower: float, upper: float) -> np.ndarray:
    """Independent restatement of Canny's steps after smoothing (whole 2-D
    image, numpy order [y, x]): vectorised gradient, magnitude, direction and
    suppression; hysteresis as the connected components (8-neighbour) of
    nms >= lower that hold a pixel with nms >= upper -- the fixed point the
    reference's sweeps reach, by a different algorithm."""
    from scipy import ndimage
    sm = smoothed.astype(np.float64)
    ny, nx = sm.shape
    xi = np.arange(nx)
    yi = np.arange(ny)
    gx = sm[:, np.minimum(xi + 1, nx - 1)] - sm[:, np.maximum(xi - 1, 0)]
    gy = sm[np.minimum(yi + 1, ny - 1), :] - sm[np.maximum(yi - 1, 0), :]
    mag = np.sqrt(gx * gx + gy * gy)
    ang = np.fmod(np.arctan2(gy, gx) * (180.0 / math.pi), 180.0)
    ang = np.where(ang < 0, ang + 180.0, ang)
    di = np.where((ang < 22.5) | (ang >= 157.5), 1, np.where(ang < 67.5, 1, np.where(ang < 112.5, 0, -1)))
    dj = np.where((ang < 22.5) | (ang >= 157.5), 0, 1)
    J, I = np.meshgrid(yi, xi, indexing="ij")
    ma = mag[np.clip(J + dj, 0, ny - 1), np.clip(I + di, 0, nx - 1)]
    mb = mag[np.clip(J - dj, 0, ny - 1), np.clip(I - di, 0, nx - 1)]
    nms = np.where((mag >= ma) & (mag >= mb), mag, 0.0)
    weak = nms >= lower
    lab, nlab = ndimage.label(weak, structure=np.ones((3, 3)))
    strong_labels = np.unique(lab[(nms >= upper) & weak])
    edges = np.isin(lab, strong_labels[strong_labels > 0]) | (nms >= upper)
    return edges.astype(smoothed.dtype)


def ref_faithful(img: np.ndarray, img_index, spacing, sigma, req_index, req_size, ncomp=1,
                 nthreads=None) -> np.ndarray:
    """The reference's loop structure on `nthreads` host threads (f32 only)."""
    D = len(spacing)
    img = np.ascontiguousarray(img, dtype=np.float32)
    img_size = list(reversed(img.shape[:D]))
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    out = np.zeros(_shape(req_size, ncomp), dtype=np.float32)
    rc = lib().orc_ref_faithful_f32(D, ncomp, img.ctypes.data, _i64(img_index), _i64(img_size),
                                    _f64(spacing), float(sigma), _i64(req_index), _i64(req_size),
                                    out.ctypes.data, int(nthreads))
    if rc != 0:
        raise ValueError("oracle rejected arguments")
    return out


def separable_opt(img: np.ndarray, spacing, sigma, nthreads=None, out=None,
                  work=None) -> np.ndarray:
    """The whole-image 3-D f32 Gaussian with the reference's arithmetic in a
    cache-friendly loop order (rows along x in every pass, planes over
    `nthreads` host threads): the optimised-CPU comparison of the bench.
    `out` (the image's shape) and `work` (two images' worth of f32) may be
    passed in so that repeated calls reuse their pages."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    if img.ndim != 3:
        raise ValueError("3-D scalar images only")
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    nz, ny, nx = img.shape
    if out is None:
        out = np.empty_like(img)
    if work is None:
        work = np.empty(2 * img.size, dtype=np.float32)
    if out.shape != img.shape or out.dtype != np.float32 or not out.flags.c_contiguous:
        raise ValueError("out: a contiguous f32 array of the image's shape")
    if work.size < 2 * img.size or work.dtype != np.float32 or not work.flags.c_contiguous:
        raise ValueError("work: a contiguous f32 array of twice the image's size")
    rc = lib().orc_separable_opt_f32(img.ctypes.data, nx, ny, nz, _f64(spacing), float(sigma),
                                     out.ctypes.data, work.ctypes.data, int(nthreads))
    if rc != 0:
        raise ValueError("oracle rejected arguments")
    return out


def adjoint_f64(gout: np.ndarray, img_index, img_size, spacing, sigma, req_index, req_size,
                ncomp=1):
    """Adjoint (transpose) of generate_region in f64 by scattered taps.

    Returns (gin, in_index, in_size) where gin lives on the clipped padded
    input region."""
    D = len(spacing)
    gout = np.ascontiguousarray(gout, dtype=np.float64)
    radii = [kernel_radius(sigma, sp) for sp in spacing]
    pi, ps = padded(req_index, req_size, radii)
    ii, isz = clipped(pi, ps, img_index, img_size)
    gin = np.zeros(_shape(isz, ncomp), dtype=np.float64)
    oi = (ctypes.c_int64 * D)()
    os_ = (ctypes.c_int64 * D)()
    rc = lib().orc_adjoint_f64(D, ncomp, _i64(img_index), _i64(img_size), _f64(spacing),
                               float(sigma), _i64(req_index), _i64(req_size), gout.ctypes.data,
                               gin.ctypes.data, oi, os_)
    if rc != 0:
        raise ValueError("oracle rejected arguments")
    assert list(oi) == ii and list(os_) == isz
    return gin, ii, isz


# ---------------------------------------------------------------------------
# independent numpy float32 restatement (checks the C oracle bit for bit)
# ---------------------------------------------------------------------------

def _np_radius(sigma, spacing):
    v = math.ceil(3.0 * (sigma / spacing)) if not math.isnan(sigma / spacing) else 0
    return max(int(v), 0)


def _np_kernel(sigma_px, r):
    # math.exp is the C library's exp (what Rust's f64::exp calls); numpy's
    # vectorised exp can differ in the last ulp.
    x = np.arange(2 * r + 1, dtype=np.float64) - float(r)
    t = x / sigma_px
    k = np.array([math.exp(v) for v in (-0.5 * (t * t))], dtype=np.float64)
    s = 0.0
    for v in k:  # sequential sum, as Iterator::sum folds
        s += float(v)
    return k / s


def np_generate_region(img: np.ndarray, img_index, spacing, sigma, req_index, req_size,
                       ncomp: int = 1) -> np.ndarray:
    """numpy restatement; float dtype of `img` is kept for every pass."""
    D = len(spacing)
    T = img.dtype.type
    size = list(reversed(img.shape[:D]))
    radii = [_np_radius(sigma, sp) for sp in spacing]
    pi, ps = padded(req_index, req_size, radii)
    in_i, in_s = clipped(pi, ps, img_index, size)
    # leaf clipped copy
    sl = tuple(slice(in_i[d] - img_index[d], in_i[d] - img_index[d] + in_s[d])
               for d in reversed(range(D)))
    cur = np.array(img[sl], dtype=img.dtype, copy=True)
    cur_i, cur_s = list(in_i), list(in_s)
    for a in range(D):
        w = _np_kernel(sigma / spacing[a], radii[a])
        if a + 1 < D:
            rem = [0] * D
            for d in range(a + 1, D):
                rem[d] = radii[d]
            oi, os_ = clipped(*padded(req_index, req_size, rem), img_index, size)
        else:
            oi, os_ = list(req_index), list(req_size)
        np_axis = D - 1 - a
        # the other axes of cur already equal the output's (see generate_region)
        lo, hi = cur_i[a], cur_i[a] + cur_s[a] - 1
        out_coords = np.arange(oi[a], oi[a] + os_[a])
        acc = np.zeros(_shape(os_, ncomp), dtype=img.dtype)
        for k in range(2 * radii[a] + 1):
            src = np.clip(out_coords + k - radii[a], lo, hi) - cur_i[a]
            taken = np.take(cur, src, axis=np_axis)
            acc = acc + taken * T(w[k])
        cur, cur_i, cur_s = acc, oi, os_
    return cur


# ---------------------------------------------------------------------------
# explicit operator (brute force forward, explicit transpose for the adjoint)
# ---------------------------------------------------------------------------

def axis_matrix(w: np.ndarray, out_lo: int, out_n: int, in_lo: int, in_n: int) -> np.ndarray:
    """Dense f64 matrix of one clamp-to-edge pass along one axis."""
    r = (len(w) - 1) // 2
    A = np.zeros((out_n, in_n), dtype=np.float64)
    for i in range(out_n):
        for k in range(len(w)):
            j = min(max(out_lo + i + k - r, in_lo), in_lo + in_n - 1) - in_lo
            A[i, j] += w[k]
    return A


def dense_operator(img_index, img_size, spacing, sigma, req_index, req_size):
    """Per-axis matrices [A_0, ..., A_{D-1}] and the input region they act on."""
    D = len(spacing)
    radii = [_np_radius(sigma, sp) for sp in spacing]
    in_i, in_s = clipped(*padded(req_index, req_size, radii), img_index, img_size)
    mats = []
    for a in range(D):
        w = _np_kernel(sigma / spacing[a], radii[a])
        mats.append(axis_matrix(w, req_index[a], req_size[a], in_i[a], in_s[a]))
    return mats, in_i, in_s


def dense_mean_operator(img_index, img_size, radii, req_index, req_size):
    """Per-axis f64 matrices of the box-mean filter (weights 1/(2r+1) written
    out here, not taken from the oracle) and the input region they act on."""
    D = len(radii)
    in_i, in_s = clipped(*padded(req_index, req_size, radii), img_index, img_size)
    mats = []
    for a in range(D):
        w = np.full(2 * radii[a] + 1, 1.0 / (2 * radii[a] + 1))
        mats.append(axis_matrix(w, req_index[a], req_size[a], in_i[a], in_s[a]))
    return mats, in_i, in_s


def dense_binomial_operator(img_index, img_size, reps, req_index, req_size):
    """Per-axis f64 matrices of `reps` rounds of [1, 2, 1] / 4: along each axis
    round t maps the request padded by reps - t (clipped) to the request padded
    by reps - t - 1; the product of the rounds' matrices, and the input
    region."""
    D = len(req_index)
    w = np.array([0.25, 0.5, 0.25])
    in_i, in_s = clipped(*padded(req_index, req_size, [reps] * D), img_index, img_size)
    mats = []
    for a in range(D):
        M = np.eye(in_s[a])
        lo, n = in_i[a], in_s[a]
        for t in range(reps):
            oi, os_ = clipped(*padded([req_index[a]], [req_size[a]], [reps - t - 1]),
                              [img_index[a]], [img_size[a]])
            M = axis_matrix(w, oi[0], os_[0], lo, n) @ M
            lo, n = oi[0], os_[0]
        mats.append(M)
    return mats, in_i, in_s


def apply_dense(mats, x: np.ndarray, transpose=False) -> np.ndarray:
    """Apply the separable operator (or its transpose) to x (numpy order)."""
    D = len(mats)
    y = np.asarray(x, dtype=np.float64)
    for a in range(D):
        M = mats[a].T if transpose else mats[a]
        np_axis = D - 1 - a
        y = np.moveaxis(np.tensordot(M, y, axes=([1], [np_axis])), 0, np_axis)
    return y


# ---------------------------------------------------------------------------
# UnaryFilter casts (src/filters/mod.rs:47-80; examples/blur_png.rs:35-37)
# ---------------------------------------------------------------------------

def rust_round(x: np.ndarray) -> np.ndarray:
    """f32::round / f64::round: nearest integer, halves away from zero (NaN
    and infinities unchanged).  x - trunc(x) is exact, so is the test."""
    x = np.asarray(x)
    t = np.trunc(x)
    with np.errstate(invalid="ignore"):
        d = x - t
        # select rather than add, so a zero result keeps its sign
        # (round(-0.3) is -0.0)
        r = np.where(d >= 0.5, t + 1, np.where(d <= -0.5, t - 1, t))
    return np.where(np.isfinite(x), r, x).astype(x.dtype)


def rust_as(x: np.ndarray, dtype) -> np.ndarray:
    """Rust's `x as T` for numeric x: integer -> integer keeps the low bits
    (narrowing) or extends by the source's signedness (widening); integer
    and float -> float round to nearest even; float -> integer truncates
    toward zero, saturates at T's range, and maps NaN to 0."""
    x = np.asarray(x)
    dt = np.dtype(dtype)
    if np.issubdtype(x.dtype, np.floating) and np.issubdtype(dt, np.integer):
        info = np.iinfo(dt)
        with np.errstate(invalid="ignore"):
            v = np.trunc(x.astype(np.float64))
            v = np.where(np.isnan(v), 0.0, np.clip(v, info.min, info.max))
        return v.astype(dt)
    return x.astype(dt)


def blur_png_chain(img_u8: np.ndarray, sigma: float, spacing=None) -> np.ndarray:
    """examples/blur_png.rs:35-40 on the C oracle: `p as f32`, the whole-image
    GaussianFilter, `p.round().clamp(0.0, 255.0) as u8`."""
    D = img_u8.ndim
    sp = list(spacing) if spacing is not None else [1.0] * D
    f = rust_as(img_u8, np.float32)
    size = list(reversed(img_u8.shape))
    g = generate_region(f, [0] * D, sp, sigma, [0] * D, size)
    r = rust_round(g)
    with np.errstate(invalid="ignore"):
        c = np.where(np.isnan(r), r, np.clip(r, np.float32(0.0), np.float32(255.0)))
    return rust_as(c.astype(np.float32), np.uint8)
