"""Adjoint of cat2amln (sfb_cat2amln_vjp) on the GPU.

L(a) = Re<c, a> for a random cotangent c.  Pins:

* against the oracle: a is affine in the weights, so dL/dw_i = Re<c, a_i> with
  a_i the oracle transform of galaxy i alone (w = 1, zero window); dL/dr_i is the
  central difference of that single-galaxy oracle transform in r, times w_i;
* dL/dnbar against a central difference of the GPU forward in nbar;
* at the headline size (N = 1e6, kmax 0.1, nside 128), the dot-product identity
  Re<c, a(w + dw) - a(w)> = <dL/dw, dw> and dL/dr against central differences of
  the forward for L = sum |a - a_target|^2 (a_target from a second catalogue).
"""
import numpy as np
import pytest

import sfb_amd as S
from helpers import separable_unit_window, shell_volume, uniform_shell_catalog
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def oracle_view(am):
    om = O.OracleModes.__new__(O.OracleModes)
    om.nside, om.lmax, om.nmax, om.lmax_n = am.nside, am.lmax, am.nmax, am.lmax_n
    om.knot_off, om.knot_x, om.knot_y, om.knot_ypp = am.knot_off, am.knot_x, am.knot_y, am.knot_ypp
    return om


def rdot(c, a):
    return float(np.sum(c.real * a.real + c.imag * a.imag))


@pytest.fixture(scope="module")
def small():
    am = S.AnlmModes(0.03, 500.0, 1000.0, nside=16)
    win, _ = separable_unit_window(S, am, nr=100)
    cat = uniform_shell_catalog(3000, 500.0, 1000.0, seed=5)
    w = np.random.default_rng(6).uniform(0.5, 1.5, cat.shape[1])
    nbar = cat.shape[1] / shell_volume(500.0, 1000.0)
    c = np.random.default_rng(7).standard_normal(S.getnlmsize(am)) * (1 + 0j)
    c += 1j * np.random.default_rng(8).standard_normal(c.size)
    return am, win, cat, w, nbar, c


def test_vjp_against_oracle_single_galaxy(small):
    am, win, cat, w, nbar, c = small
    say("vjp vs oracle single-galaxy transforms", am.nmaps, "maps")
    dr, dw, _ = S.cat2amln_vjp(cat, am, nbar, c, w)
    om = oracle_view(am)
    zero_mask = np.zeros(12 * am.nside ** 2)
    zero_wln = np.zeros((am.lmax + 1) * am.nmax)
    h = 1e-3
    for i in (0, 17, 1234, 2999):
        g = cat[:, i:i + 1].copy()
        a_i = O.cat2amln(om, g, nbar, 1, zero_mask, zero_wln, w=np.ones(1))
        assert dw[i] == pytest.approx(rdot(c, a_i), rel=1e-10, abs=1e-12 * np.abs(dw).max())
        gp, gm = g.copy(), g.copy()
        gp[0, 0] += h
        gm[0, 0] -= h
        fd = w[i] * (rdot(c, O.cat2amln(om, gp, nbar, 1, zero_mask, zero_wln, w=np.ones(1))) -
                     rdot(c, O.cat2amln(om, gm, nbar, 1, zero_mask, zero_wln, w=np.ones(1)))) / (2 * h)
        assert dr[i] == pytest.approx(fd, rel=1e-5, abs=1e-9 * np.abs(dr).max())


def test_vjp_nbar_and_unit_weights(small):
    am, win, cat, w, nbar, c = small
    say("vjp dL/dnbar and unit weights")
    _, dw, dnbar = S.cat2amln_vjp(cat, am, nbar, c, w)
    hb = 1e-6 * nbar
    fp = rdot(c, S.cat2amln(cat, am, nbar + hb, win, w))
    fm = rdot(c, S.cat2amln(cat, am, nbar - hb, win, w))
    assert dnbar == pytest.approx((fp - fm) / (2 * hb), rel=1e-6)
    # w = NULL is ones(N); dL/dw does not depend on w, dL/dr scales with w
    dr1, dw1, _ = S.cat2amln_vjp(cat, am, nbar, c, None)
    dr2, dw2, _ = S.cat2amln_vjp(cat, am, nbar, c, np.ones(cat.shape[1]))
    assert np.array_equal(dr1, dr2) and np.array_equal(dw1, dw2)
    assert np.allclose(dw1, dw, rtol=1e-13, atol=0)


def test_vjp_niter0_is_plain_transpose(small):
    """niter = 0: a = A delta, so <c, a(w)> is exactly linear; check with a random dw."""
    am, win, cat, w, nbar, c = small
    say("vjp niter=0 dot-product identity")
    p = S.Plan(am, niter=0)
    _, dw, _ = p.cat2amln_vjp(cat, nbar, c, w)
    d = np.random.default_rng(9).standard_normal(w.size)
    lhs = rdot(c, p.cat2amln(cat, nbar, win, w + d) - p.cat2amln(cat, nbar, win, w))
    assert lhs == pytest.approx(float(dw @ d), rel=1e-10)


@pytest.fixture(scope="module")
def c2():
    am = S.AnlmModes(0.1, 500.0, 1500.0, nside=128)
    win, _ = separable_unit_window(S, am)
    cat = uniform_shell_catalog(1_000_000, 500.0, 1500.0, seed=0)
    nbar = cat.shape[1] / shell_volume(500.0, 1500.0)
    return am, win, cat, nbar, S.get_plan(am)


def test_vjp_c2_dot_product_identity(c2):
    am, win, cat, nbar, plan = c2
    say("C2 vjp dot-product identity")
    rng = np.random.default_rng(10)
    c = rng.standard_normal(S.getnlmsize(am)) + 1j * rng.standard_normal(S.getnlmsize(am))
    w = np.ones(cat.shape[1])
    _, dw, _ = plan.cat2amln_vjp(cat, nbar, c, w)
    d = rng.uniform(-0.5, 0.5, w.size)
    lhs = rdot(c, plan.cat2amln(cat, nbar, win, w + d) - plan.cat2amln(cat, nbar, win, w))
    assert lhs == pytest.approx(float(dw @ d), rel=1e-9)


def test_vjp_c2_radius_finite_differences(c2):
    am, win, cat, nbar, plan = c2
    say("C2 vjp dL/dr vs central differences")
    target = plan.cat2amln(uniform_shell_catalog(1_000_000, 500.0, 1500.0, seed=1), nbar, win)
    a = plan.cat2amln(cat, nbar, win)
    c = 2 * (a - target)  # dL/da for L = sum |a - target|^2
    dr, _, _ = plan.cat2amln_vjp(cat, nbar, c)
    h = 1e-2
    for i in (3, 500_000, 999_999):
        cp, cm = cat.copy(), cat.copy()
        cp[0, i] += h
        cm[0, i] -= h
        # contract the forward difference with dL/da (L's own difference would cancel)
        fd = rdot(c, plan.cat2amln(cp, nbar, win) - plan.cat2amln(cm, nbar, win)) / (2 * h)
        assert dr[i] == pytest.approx(fd, rel=1e-4)


def test_vjp_rejects_nonfinite_angle(small):
    """As the forward call: a NaN angle has no pixel, so the adjoint raises too."""
    am, win, cat, w, nbar, c = small
    bad = cat.copy()
    bad[1, 10] = np.nan
    with pytest.raises(S.SFBError) as e:
        S.cat2amln_vjp(bad, am, nbar, c, w)
    assert "not finite" in str(e.value)
    with pytest.raises(S.SFBError):
        S.cat2amln(bad, am, nbar, win, w)
    # the plan stays usable
    dr, dw, _ = S.cat2amln_vjp(cat, am, nbar, c, w)
    assert np.all(np.isfinite(dr)) and np.all(np.isfinite(dw))
