"""nside = 256, kmax = 0.15 (lmax 214, nmax 48, 5645 maps): N = 4e6 galaxies in the
northern hemisphere with a Gaussian radial profile, FP32-valued weights, and the
matching separable window (mask = northern pixels, phi(r) = the Gaussian).  The
spline tables (up to ~3600 knots) and the nside-256 rings exercise the largest
configuration one GPU runs.  A sample of maps (low and high l) is checked against
the oracle, and the transform must be invariant to the catalogue order."""
import numpy as np
import pytest

import sfb_amd as S
from helpers import gaussian_shell_catalog, map_entries
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c3():
    am = S.AnlmModes(0.15, 500.0, 1500.0, nside=256)
    say("C3 modes", am.lmax, am.nmax, am.nmaps)
    nr = 1000
    wm = S.ConfigurationSpaceModes(500.0, 1500.0, nr, 256)
    phi = np.exp(-(wm.r - 1000.0) ** 2 / (2 * 300.0 ** 2))
    th, _ = O.pix2ang_ring(256, np.arange(12 * 256 ** 2))
    mask = (th < np.pi / 2).astype(np.float64)
    sep = S.SeparableArray(phi, mask, name1="phi", name2="mask")
    win = S.win_rhat_ln(sep, wm, am)
    cat, w = gaussian_shell_catalog(4_000_000, seed=3)
    nbar = cat.shape[1] / S.integrate_window(sep, wm)
    return am, win, cat, w, nbar


def test_c3_sizes(c3):
    am = c3[0]
    assert (am.lmax, am.nmax, am.nmaps, S.getnlmsize(am)) == (214, 48, 5645, 390799)


def test_c3_sample_against_oracle(c3):
    am, win, cat, w, nbar = c3
    say("C3 transform + oracle sample")
    plan = S.get_plan(am)
    a = plan.cat2amln(cat, nbar, win, w)
    assert np.all(np.isfinite(a))
    om = O.OracleModes.__new__(O.OracleModes)
    om.nside, om.lmax, om.nmax, om.lmax_n = am.nside, am.lmax, am.nmax, am.lmax_n
    om.knot_off, om.knot_x, om.knot_y, om.knot_ypp = am.knot_off, am.knot_x, am.knot_y, am.knot_ypp
    wln = win.w_ln.reshape(-1, order="F")
    for q0, q1 in ((0, 3), (212, 215)):
        ref = O.cat2amln(om, cat, nbar, 1, win.mask, wln, w=w, q_range=(q0, q1))
        sel = map_entries(am, q0, q1)
        d = np.abs(a[sel] - ref[sel])
        say(f"maps {q0}..{q1 - 1}: max |d| / max |ref| = {d.max() / np.abs(ref[sel]).max():.2e}")
        assert d.max() < 1e-10 * np.abs(ref[sel]).max()


def test_c3_permutation_invariance(c3):
    am, win, cat, w, nbar = c3
    say("C3 permutation invariance")
    plan = S.get_plan(am)
    a = plan.cat2amln(cat, nbar, win, w)
    perm = np.random.default_rng(4).permutation(cat.shape[1])
    assert np.array_equal(a, plan.cat2amln(cat[:, perm], nbar, win, w[perm]))
