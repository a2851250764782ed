"""The scatter forms the spline parameter t = (r - x_i)/h without a divide:
y = RN(1/h) per segment (host), q0 = RN(n y), t = fma(fma(-h, q0, n), y, q0) with
n = r - x_i.  This is Markstein's final correction step (P. Markstein, IBM J. Res. Dev.
34 (1990) 111-119: with y the correctly rounded reciprocal and q0 within one ulp of n/h,
the corrected quotient is the correctly rounded n/h); the delta maps are only
bit-identical to the reference's if it returns RN(n/h) for every r the kernel can see.
Checked here with exact rational arithmetic (each fma rounded once) on every distinct
segment of the C1 mode table, and with the correctly rounded C fma on every distinct
segment of the nside-128 (kmax 0.1) and nside-256 (kmax 0.15) tables: each segment's
edges, midpoint and 1020 uniformly drawn radii.  The same check without the correction
step finds mismatches (it is not vacuous)."""
import math
from fractions import Fraction

import numpy as np
import pytest

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


@pytest.mark.parametrize("kmax,nside", [(0.1, 128), (0.15, 256)])
def test_markstein_quotient_on_all_segments_c2_c3(kmax, nside):
    from oracle import oracle as O
    am = S.AnlmModes(kmax, 500.0, 1500.0, nside=nside)
    x0, x1 = [], []
    for q in range(am.nmaps):
        x = am.knot_x[am.knot_off[q]:am.knot_off[q + 1]]
        x0.append(x[:-1])
        x1.append(x[1:])
    seg = np.unique(np.stack([np.concatenate(x0), np.concatenate(x1)]), axis=1)
    bad, where = O.markstein_check(seg[0], seg[1], nr=1024, seed=int(kmax * 1000))
    print(f"kmax {kmax}: {seg.shape[1]} distinct segments x 1024 radii, {bad} mismatches")
    assert bad == 0, where
    # control: without the correction step the check does find mismatches
    bad_control, _ = O.markstein_check(seg[0][:2000], seg[1][:2000], nr=1024, variant=1)
    assert bad_control > 0
