"""Generates canny_edges.inc: the exact bin-edge thresholds of the direction
binning in CannyEdgeDetectionFilter's non-maximum suppression
(src/filters/edges.rs:602, 612-620):

    angle = gy.atan2(gx).to_degrees().rem_euclid(180.0)
    bin   = 0 if angle < 22.5 or angle >= 157.5, 1 below 67.5, 2 below 112.5, else 3

With r = atan2(gy, gx) rounded to the nearest double, `to_degrees` one
rounded multiply by the double C nearest 180/pi, and `rem_euclid` (exact
for |deg| < 180) one rounded add of 180 for negative angles, the angle is a
non-decreasing function of r on each half-plane.  So for each bin edge e and
half-plane there is a smallest double r* with angle(r*) >= e, and because r
is the rounding of the true angle theta, angle >= e exactly when theta lies
above the midpoint m between r* and its predecessor.  theta > m has the sign
of gy cos(m) - gx sin(m), which the kernel evaluates in double-double
arithmetic when the ordinary computation lands within 1e-9 degrees of an
edge.  The file holds cos(m) and sin(m) as (hi, lo) pairs for the 4 edges on
the upper half-plane (r > 0) and the 4 on the lower (r < 0).

    python gen_canny_edges.py [out.inc]      (default: canny_edges.inc beside it)
"""
import math
import os
import sys
from decimal import Decimal, getcontext
from fractions import Fraction

EDGES = (22.5, 67.5, 112.5, 157.5)
C = 180.0 / math.pi  # Rust: 180.0f64 / consts::PI, one rounded division


def angle_of(r: float) -> float:
    deg = r * C
    return deg + 180.0 if deg < 0.0 else deg


def smallest_reaching(e: float, upper: bool) -> float:
    """The smallest double r on the half-plane with angle_of(r) >= e."""
    r = e / C if upper else (e - 180.0) / C
    while angle_of(r) < e:
        r = math.nextafter(r, math.inf)
    while angle_of(math.nextafter(r, -math.inf)) >= e:
        r = math.nextafter(r, -math.inf)
    return r


def _sin_cos(x: Fraction, digits: int = 60):
    getcontext().prec = digits + 10
    d = Decimal(x.numerator) / Decimal(x.denominator)
    s, c, term_s, term_c = Decimal(0), Decimal(0), d, Decimal(1)
    n = 0
    while True:
        s += term_s
        c += term_c
        term_s = -term_s * d * d / Decimal((2 * n + 2) * (2 * n + 3))
        term_c = -term_c * d * d / Decimal((2 * n + 1) * (2 * n + 2))
        n += 1
        if abs(term_s) < Decimal(10) ** -(digits + 5) and abs(term_c) < Decimal(10) ** -(digits + 5):
            return s, c


def _split(v: Decimal):
    hi = float(v)
    lo = float(v - Decimal(hi))
    return hi, lo


def thresholds():
    """[(edge, upper, r_star, (cos_hi, cos_lo), (sin_hi, sin_lo))] for the 8 edges."""
    out = []
    for upper in (True, False):
        for e in EDGES:
            r = smallest_reaching(e, upper)
            m = (Fraction(r) + Fraction(math.nextafter(r, -math.inf))) / 2
            s, c = _sin_cos(m)
            out.append((e, upper, r, _split(c), _split(s)))
    return out


def render() -> str:
    lines = ["// Generated by gen_canny_edges.py; do not edit.",
             "// {cos_hi, cos_lo, sin_hi, sin_lo} of the bin-edge midpoint m, edges",
             "// 22.5, 67.5, 112.5, 157.5 degrees; row 0: atan2 > 0, row 1: atan2 < 0.",
             "constexpr double kCannyEdgeTrig[2][4][4] = {"]
    rows = thresholds()
    for h in range(2):
        lines.append("    {")
        for k in range(4):
            _, _, r, (ch, cl), (sh, sl) = rows[4 * h + k]
            lines.append(f"        {{{ch.hex()}, {cl.hex()}, {sh.hex()}, {sl.hex()}}},"
                         f"  // r* = {r.hex()}")
        lines.append("    },")
    lines.append("};")
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    path = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(
        os.path.abspath(__file__)), "canny_edges.inc")
    with open(path, "w") as fh:
        fh.write(render())
