"""The scatter forms the spline parameter t = (r - x_i)/h without a divide:
y = RN(1/h) per segment (host), q0 = RN(n y), t = fma(fma(-h, q0, n), y, q0) with
n = r - x_i.  This is Markstein's final correction step; the delta maps are only
bit-identical to the reference's if it returns RN(n/h) for every r the kernel can
see.  Checked here with exact rational arithmetic (each fma rounded once) on every
distinct segment of a mode table, for random r and the segment's edge values."""
import math
from fractions import Fraction

import numpy as np

import sfb_amd as S


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def recip_quotient(n, h):
    y = 1.0 / h
    q0 = n * y
    return fma(fma(-h, q0, n), y, q0)


def test_markstein_quotient_on_all_segments():
    am = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    pairs = set()
    for q in range(am.nmaps):
        x = am.knot_x[am.knot_off[q]:am.knot_off[q + 1]]
        pairs.update(zip(x[:-1].tolist(), x[1:].tolist()))
    rng = np.random.default_rng(0)
    checked = 0
    for x0, x1 in sorted(pairs):
        h = x1 - x0
        rs = [x0, math.nextafter(x1, 0.0), math.nextafter(x0, math.inf), x0 + 0.5 * h]
        rs += (x0 + rng.random(4) * h).tolist()
        for r in rs:
            n = r - x0
            assert recip_quotient(n, h) == n / h, (x0, x1, r)
            checked += 1
    assert checked > 10000


def test_markstein_quotient_hard_cases():
    """Quotients next to rounding ties and extreme significands of h."""
    rng = np.random.default_rng(1)
    for _ in range(3000):
        h = float(rng.uniform(0.05, 3.0))
        h = math.nextafter(h, math.inf) if rng.random() < 0.5 else h
        t = float(rng.random())
        n = t * h
        for nn in (n, math.nextafter(n, 0.0), math.nextafter(n, math.inf)):
            assert recip_quotient(nn, h) == nn / h
    for h in (1.0, 2.0 - 2.0 ** -52, 1.0 + 2.0 ** -52, 0.75, 1.5 - 2.0 ** -52):
        for n in (h * 0.5, math.nextafter(h, 0.0), 2.0 ** -30, 0.1, 1.0 / 3.0):
            assert recip_quotient(n, h) == n / h
