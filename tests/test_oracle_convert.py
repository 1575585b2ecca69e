"""The oracle's restatement of Rust's numeric casts (the UnaryFilter
closures around a Gaussian, examples/blur_png.rs:35-37) against known
answers that follow from Rust's documented rules: float -> integer `as`
truncates toward zero, saturates and sends NaN to 0; integer -> integer `as`
keeps the low bits when narrowing and extends by the source's signedness
when widening; to float rounds to nearest even; f32::round rounds halves
away from zero."""
import numpy as np

F = np.float32
NAN, INF = float("nan"), float("inf")


def test_float_to_u8_as(orc):
    x = np.array([NAN, -INF, -1.0, -0.7, 0.0, 0.4, 0.5, 0.9999, 1.5, 254.7, 255.0, 255.9,
                  256.0, 1e30, INF], dtype=F)
    want = [0, 0, 0, 0, 0, 0, 0, 0, 1, 254, 255, 255, 255, 255, 255]
    assert orc.rust_as(x, np.uint8).tolist() == want


def test_round_half_away_then_u8(orc):
    x = np.array([NAN, -INF, -1.0, -0.7, -0.5, 0.0, 0.4, 0.49999997, 0.5, 1.5, 2.5, 254.4,
                  254.5, 255.5, 1e30], dtype=F)
    r = orc.rust_round(x)
    assert np.isnan(r[0]) and r[1] == -INF
    assert r[2:].tolist() == [-1.0, -1.0, -1.0, 0.0, 0.0, 0.0, 1.0, 2.0, 3.0, 254.0, 255.0,
                              256.0, float(F(1e30))]
    assert orc.rust_as(r, np.uint8).tolist() == [0, 0, 0, 0, 0, 0, 0, 0, 1, 2, 3, 254, 255,
                                                  255, 255]
    # f64 halves near the last bit: 0.49999999999999994 rounds to 0, not 1
    d = np.array([0.49999999999999994, -0.49999999999999994, 2.5, -2.5], dtype=np.float64)
    assert orc.rust_round(d).tolist() == [0.0, -0.0, 3.0, -3.0]
    # a zero result keeps the sign of its input: round(-0.3) is -0.0
    z = orc.rust_round(np.array([-0.3, -0.0, 0.3, -0.49999997], dtype=F))
    assert np.signbit(z).tolist() == [True, True, False, True] and (z == 0).all()


def test_float_to_signed(orc):
    x = np.array([-32768.5, 32767.9, 40000.0, -40000.0, -1.9, NAN], dtype=F)
    assert orc.rust_as(x, np.int16).tolist() == [-32768, 32767, 32767, -32768, -1, 0]
    y = np.array([2147483520.0, 2147483648.0, -2147483648.0, -3e9], dtype=F)
    assert orc.rust_as(y, np.int32).tolist() == [2147483520, 2147483647, -2147483648,
                                                 -2147483648]


def test_integer_casts(orc):
    assert orc.rust_as(np.array([256, 257, -1, 300], dtype=np.int32), np.uint8).tolist() == \
        [0, 1, 255, 44]
    assert orc.rust_as(np.array([-1, -32768], dtype=np.int16), np.int32).tolist() == [-1, -32768]
    assert orc.rust_as(np.array([65535], dtype=np.uint16), np.int32).tolist() == [65535]
    assert orc.rust_as(np.array([65535], dtype=np.uint16), np.int16).tolist() == [-1]
    # to float: nearest even
    assert orc.rust_as(np.array([16777217, 16777219], dtype=np.int32), F).tolist() == \
        [16777216.0, 16777220.0]
    assert orc.rust_as(np.array([255, 0], dtype=np.uint8), F).tolist() == [255.0, 0.0]


def test_blur_png_chain_constant_and_impulse(orc):
    # a constant image stays constant through the chain
    c = np.full((20, 30), 77, dtype=np.uint8)
    assert (orc.blur_png_chain(c, 2.0) == 77).all()
    # an interior impulse of 255: the centre value is round(255 * w_r^2)
    imp = np.zeros((41, 41), dtype=np.uint8)
    imp[20, 20] = 255
    out = orc.blur_png_chain(imp, 1.5)
    w = orc.build_kernel(1.5, orc.kernel_radius(1.5, 1.0)).astype(np.float32)
    centre = F(F(F(255.0) * w[5]) * w[5])
    assert out[20, 20] == int(np.floor(float(centre) + 0.5))
    assert out.sum() > 0 and out[0, 0] == 0
