"""pixwin on the device (one pixel per ring, pair sums of P_l) against the oracle's
restatement from the pixel means of Y_lm over every pixel (tests/test_pixwin.py pins the
oracle); convergence in the sub-pixel depth at nside 128; pixwin(cmodes) per (l, n1, n2)."""
import numpy as np
import pytest

import sfb_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nside,lmax,depth", [(1, 4, 3), (2, 8, 3), (4, 16, 3)])
def test_pixwin_against_oracle(nside, lmax, depth):
    print(f"pixwin nside {nside} depth {depth}", flush=True)
    w = S.healpix_pixwin(nside, lmax, depth=depth)
    ref = O.pixwin_ylm(nside, lmax, depth)
    print("  max |w - oracle| =", np.abs(w - ref).max(), flush=True)
    assert np.abs(w - ref).max() < 1e-12


def test_pixwin_nside128_depths():
    w5 = S.healpix_pixwin(128, 512, depth=5)
    w4 = S.healpix_pixwin(128, 512, depth=4)
    d = np.abs(w5 - w4)
    print(f"nside 128: w_0 = {w5[0]!r}, w_140 = {w5[140]:.8f}, w_256 = {w5[256]:.8f}, w_512 = {w5[512]:.8f}; "
          f"depth 4 vs 5: max {d.max():.2e} (l <= 256: {d[:257].max():.2e})", flush=True)
    assert abs(w5[0] - 1.0) < 1e-12
    assert np.all(np.diff(w5) < 0)
    assert d[:257].max() < 1e-4 and d.max() < 1e-3


def test_pixwin_cmodes():
    am = S.AnlmModes(0.03, 500.0, 1000.0, nside=32)
    cm = S.ClnnModes(am, Δnmax=1)
    pw = S.pixwin(cm)
    w = S.healpix_pixwin(32, am.lmax)
    assert pw.shape == (S.getlnnsize(cm),)
    for i in range(1, S.getlnnsize(cm) + 1):
        l, n1, n2 = S.getlnn(cm, i)
        assert pw[i - 1] == w[l]
