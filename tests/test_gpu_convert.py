"""GPU pixel conversions (the UnaryFilter casts around a Gaussian,
src/filters/mod.rs:47-80, examples/blur_png.rs:35-37) bit for bit against
the oracle's restatement of Rust's casts, and the blur_png chain (u8 -> f32,
Gaussian, round and cast to u8) against the oracle, device-resident and on
host buffers."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NAN, INF = float("nan"), float("inf")
TYPES = [np.uint8, np.int16, np.int32, np.float32, np.float64]
TORCH = {np.uint8: torch.uint8, np.int16: torch.int16, np.int32: torch.int32,
         np.float32: torch.float32, np.float64: torch.float64}
_PT = {np.uint8: 0, np.int16: 2, np.int32: 3, np.float32: 4, np.float64: 5}  # itkg_ptype


def _values(dt, rng, n=4096):
    dt = np.dtype(dt)
    if np.issubdtype(dt, np.floating):
        special = np.array([NAN, INF, -INF, 0.0, -0.0, 0.5, -0.5, 1.5, 2.5, -2.5, 0.49999997,
                            254.5, 255.5, 255.0, 256.0, -1.0, 32767.5, -32768.5, 65535.5,
                            2147483520.0, 2147483648.0, -2147483904.0, 1e30, -1e30, 16777217.0],
                           dtype=dt)
        body = (rng.standard_normal(n) * rng.choice([1.0, 100.0, 1e5, 3e9], n)).astype(dt)
        return np.concatenate([special, body])
    info = np.iinfo(dt)
    special = np.array([info.min, info.max, 0, 1, -1 if info.min < 0 else 2, 255, 256, 257],
                       dtype=np.int64).clip(info.min, info.max).astype(dt)
    body = rng.integers(info.min, int(info.max) + 1, n, dtype=np.int64).astype(dt)
    return np.concatenate([special, body])


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype != b.dtype or a.shape != b.shape:
        return False
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("src", TYPES)
@pytest.mark.parametrize("mode", [0, 1])
def test_every_type_pair_matches_rust_casts(itk, orc, dev, src, mode):
    rng = np.random.default_rng(2 * TYPES.index(src) + mode)
    full = _values(src, rng)
    # a dense run of odd length (four elements per lane, the last one in
    # block 0) and one of a multiple of four
    for x in (full, np.ascontiguousarray(full[: full.size // 4 * 4])):
        xin = itk.Image(itk.Region([0], [x.size]), (1.0,), (0.0,), torch.from_numpy(x).to(dev))
        for dst in TYPES:
            out = itk.Image.empty(xin.region, (1.0,), (0.0,), 1, TORCH[dst], dev)
            itk.CastFilter(xin, TORCH[dst], mode).generate_into(xin, out)
            torch.cuda.synchronize()
            v = orc.rust_round(x) if (mode == 1 and np.issubdtype(np.dtype(src), np.floating)) else x
            ref = orc.rust_as(v, dst)
            assert _same(out.data.cpu().numpy(), ref), (np.dtype(src), np.dtype(dst), mode, x.size)


@pytest.mark.parametrize("src", TYPES)
@pytest.mark.parametrize("mode", [0, 1])
def test_every_type_pair_on_windows(itk, orc, dev, src, mode):
    """64 rows of a wider buffer into 64 rows of another wider buffer: rows
    of a multiple of four elements at strides of multiples of four (four
    elements per lane, row by row) and rows of 63 at odd strides (one
    element per lane); the padding of the output buffer stays untouched."""
    rng = np.random.default_rng(40 + 2 * TYPES.index(src) + mode)
    vals = _values(src, rng, 5000)
    for w, pin, pout in ((64, 68, 72), (63, 65, 67)):
        xb = np.ascontiguousarray(vals[: 64 * pin].reshape(64, pin))
        img = itk.Image(itk.Region([0, 0], [pin, 64]), (1.0, 1.0), (0.0, 0.0),
                        torch.from_numpy(xb).to(dev))
        req = itk.Region([0, 0], [w, 64])
        win_in = itk.Image(req, (1.0, 1.0), (0.0, 0.0), img.data[img._slices(req)])
        x = xb[:, :w]
        v = orc.rust_round(x) if (mode == 1 and np.issubdtype(np.dtype(src), np.floating)) else x
        for dst in TYPES:
            ob = torch.zeros((64, pout), dtype=TORCH[dst], device=dev)
            big = itk.Image(itk.Region([0, 0], [pout, 64]), (1.0, 1.0), (0.0, 0.0), ob)
            out = itk.Image(req, (1.0, 1.0), (0.0, 0.0), big.data[big._slices(req)])
            rc = itk.lib.itkg_convert_pixels(2, _PT[src], _PT[dst], 1, ctypes.byref(win_in.view()),
                                             ctypes.byref(out.view()), mode, itk._stream_ptr(dev))
            assert rc == 0, itk.lib.itkg_last_error()
            torch.cuda.synchronize()
            got = ob.cpu().numpy()
            assert _same(np.ascontiguousarray(got[:, :w]), orc.rust_as(v, dst)), \
                (np.dtype(src), np.dtype(dst), mode, w)
            assert (got[:, w:] == 0).all()


def test_strided_subregion_and_components(itk, orc, dev):
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((9, 14, 17, 3)) * 300).astype(np.float32)
    big = torch.from_numpy(x).to(dev)
    img = itk.Image(itk.Region([2, -1, 4], [17, 14, 9]), (1.0,) * 3, (0.0,) * 3, big)
    req = itk.Region([5, 2, 6], [9, 8, 4])
    out_t = torch.zeros((6, 10, 12, 3), dtype=torch.uint8, device=dev)
    # the output is a window of a larger buffer (strides of the big buffer)
    out_big = itk.Image(itk.Region([4, 1, 5], [12, 10, 6]), (1.0,) * 3, (0.0,) * 3, out_t)
    win = out_big.data[out_big._slices(req)]
    out = itk.Image(req, (1.0,) * 3, (0.0,) * 3, win)
    rc = itk.lib.itkg_convert_pixels(3, itk.PT_F32, itk.PT_U8, 3, ctypes.byref(img.view()),
                                     ctypes.byref(out.view()), itk.CONVERT_ROUND,
                                     itk._stream_ptr(dev))
    assert rc == 0, itk.lib.itkg_last_error()
    torch.cuda.synchronize()
    sub = x[6 - 4:6 - 4 + 4, 2 + 1:2 + 1 + 8, 5 - 2:5 - 2 + 9, :]
    ref = orc.rust_as(orc.rust_round(sub), np.uint8)
    got = out_t.cpu().numpy()
    assert _same(got[1:5, 1:9, 1:10, :], ref)
    # nothing outside the window was written
    mask = np.ones(got.shape[:3], dtype=bool)
    mask[1:5, 1:9, 1:10] = False
    assert (got[mask] == 0).all()


@pytest.mark.parametrize("shape,sigma", [((300, 411), 2.0), ((64, 65), 1.5), ((40, 50, 60), 1.2)])
def test_blur_png_chain_resident_and_host(itk, orc, dev, shape, sigma):
    rng = np.random.default_rng(sum(shape))
    img = rng.integers(0, 256, shape, dtype=np.int64).astype(np.uint8)
    img[shape[0] // 3, :] = 255  # an edge, so rounding sees halves and extremes
    D = img.ndim
    ref = orc.blur_png_chain(img, sigma)
    src = itk.Image(itk.Region([0] * D, list(reversed(shape))), (1.0,) * D, (0.0,) * D,
                    torch.from_numpy(img).to(dev))
    chain = itk.CastFilter(itk.GaussianFilter(itk.CastFilter(src, torch.float32), sigma),
                           torch.uint8, itk.CONVERT_ROUND)
    got = chain.generate_region(chain.largest_region()).data.cpu().numpy()
    assert _same(got, ref)
    host = itk.cast_gaussian_cast_host(img, sigma, np.uint8)
    assert _same(host, ref)
    # FAST arithmetic: within one grey level, and equal almost everywhere
    fast = itk.cast_gaussian_cast_host(img, sigma, np.uint8, arith=itk.FAST)
    d = np.abs(fast.astype(np.int32) - ref.astype(np.int32))
    assert d.max() <= 1 and (d == 0).mean() > 0.99


def test_host_chain_u16_to_f32(itk, orc):
    rng = np.random.default_rng(9)
    img = rng.integers(0, 65536, (33, 47, 29), dtype=np.int64).astype(np.uint16)
    got = itk.cast_gaussian_cast_host(img, 1.7, np.float32, out_mode=itk.CONVERT_AS)
    f = orc.rust_as(img, np.float32)
    ref = orc.generate_region(f, [0, 0, 0], [1.0] * 3, 1.7, [0, 0, 0], [29, 47, 33])
    assert _same(got, ref)


def test_rows_beyond_the_grid_limit(itk, orc, dev):
    # axis 1 longer than 65535 rows: the whole image is one dense run (the
    # flat kernel, 350005 elements: a tail of one); its first four columns
    # are not (the generic conversion kernel)
    rng = np.random.default_rng(4)
    x = (rng.standard_normal((70001, 5)) * 200).astype(np.float32)
    img = itk.Image(itk.Region([0, 0], [5, 70001]), (1.0, 1.0), (0.0, 0.0), torch.from_numpy(x).to(dev))
    for req in (img.region, itk.Region([0, 0], [4, 70001])):
        for dst, mode in ((torch.uint8, itk.CONVERT_ROUND), (torch.int16, itk.CONVERT_AS)):
            out = itk.CastFilter(img, dst, mode).generate_region(req)
            xs = x[:, : req.size[0]]
            v = orc.rust_round(xs) if mode == itk.CONVERT_ROUND else xs
            ref = orc.rust_as(v, np.uint8 if dst == torch.uint8 else np.int16)
            assert _same(out.data.cpu().numpy(), ref), (dst, req)
