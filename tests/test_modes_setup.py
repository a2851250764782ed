"""AnlmModes: the scipy oracle pinned by closed forms and the reference's
orthonormality test, and the native setup (libsfb_mi355x.so) checked against it.

Reference: src/SphericalBesselGNLs.jl:219-228,309-336,404-421,558-579;
src/modes.jl:68,122-232; test/test_gnl.jl:11-45 (orthonormality), test/test_modes.jl:13-50.
"""
import math

import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss

import sfb_amd as S
from oracle import oracle as O


@pytest.fixture(scope="module")
def c1_pair():
    return S.AnlmModes(0.05, 500.0, 1000.0, nside=32), O.OracleModes(kmax=0.05, rmin=500.0, rmax=1000.0, nside=32)


def test_closed_form_roots_rmin0():
    om = O.OracleModes(kmax=0.02, rmin=0.0, rmax=1000.0, nside=32)
    n = np.arange(1, om.knl.shape[0] + 1)
    # l=0: J_{-1/2}(k rmax) = 0 -> k rmax = (n - 1/2) pi; l=1: J_{1/2} -> k rmax = n pi
    assert np.allclose(om.knl[:, 0] * 1000 / np.pi, n - 0.5, rtol=1e-9)
    k1 = om.knl[:, 1][np.isfinite(om.knl[:, 1])]
    assert np.allclose(k1 * 1000 / np.pi, np.arange(1, k1.size + 1), rtol=1e-9)


def test_config1_sizes(c1_pair):
    am, om = c1_pair
    for m in (am, om):
        assert (m.nmax, m.lmax, m.nmaps) == (8, 43, 237)
    assert S.getnlmsize(am) == 3963 == om.nlmsize()
    assert 2 * am.nside >= am.lmax  # test_modes.jl:29 holds for the estimate; here nside is given
    assert S.estimate_nside(am.lmax) == 64


@pytest.mark.parametrize("rmin,kmax", [(0.0, 0.01), (500.0, 0.01), (500.0, 0.03)])
def test_orthonormality_gauss_legendre(rmin, kmax):
    """sum dr r^2 g_n g_n' = delta (test/test_gnl.jl:11-45, atol 1e-6), product and oracle."""
    rmax = 2467.10
    om = O.OracleModes(kmax=kmax, rmin=rmin, rmax=rmax)
    am = S.AnlmModes(kmax, rmin, rmax)
    for l in sorted({0, 1, 2, om.lmax // 2, om.lmax}):
        nl = int(om.nmax_l[l])
        for n1 in range(1, nl + 1):
            for n2 in range(n1, min(nl, n1 + 3) + 1):
                nodes, weights = leggauss(16 * (n1 + n2))
                r = (rmin + rmax) / 2 + (rmax - rmin) / 2 * nodes
                dr = weights * (rmax - rmin) / 2
                for g in (om.gnl, am.basisfunctions):
                    v = np.sum(dr * r ** 2 * g(n1, l, r) * g(n2, l, r))
                    assert v == pytest.approx(1.0 if n1 == n2 else 0.0, abs=1e-6), (l, n1, n2)


def test_spline_knots_follow_julia_range():
    # 500:dr:1000+dr/2 with dr = 500/363 has 364 elements, exactly rounded rationals
    x = O.julia_float_range(500.0, 500.0 / 363, 1000.0 + 500.0 / 363 / 2)
    assert x.size == 364 and x[0] == 500.0 and x[-1] == 1000.0
    assert x[1] == float(500 + 500 / 363) or abs(x[1] - (500 + 500 / 363)) <= 1.2e-13
    x0 = O.julia_float_range(0.0, 0.25, 2.0)
    assert x0.tolist() == [0.25 * i for i in range(9)]


def test_native_setup_matches_oracle(c1_pair):
    am, om = c1_pair
    assert np.array_equal(am.lmax_n, om.lmax_n)
    assert np.array_equal(am.nmax_l, om.nmax_l)
    fin = np.isfinite(om.knl)
    assert np.array_equal(np.isfinite(am.knl), fin)
    assert np.max(np.abs(am.knl[fin] / om.knl[fin] - 1)) < 1e-13
    assert np.max(np.abs(am.cnl[fin] / om.cnl[fin] - 1)) < 1e-11
    assert np.array_equal(am.knot_off, om.knot_off)
    assert np.array_equal(am.knot_x, om.knot_x)  # knots are bit-identical
    assert np.max(np.abs(am.knot_y - om.knot_y)) < 1e-12 * np.max(np.abs(om.knot_y))
    # the spline second derivatives: same restated arithmetic on the same knots
    for q in (0, 100, am.nmaps - 1):
        o0, o1 = am.knot_off[q], am.knot_off[q + 1]
        ypp_o = O.spline_build(am.knot_x[o0:o1], am.knot_y[o0:o1])
        assert np.array_equal(am.knot_ypp[o0:o1], ypp_o)
        r = np.random.default_rng(q).uniform(499.0, 1001.0, 500)
        sp = S.Spline1D(am.knot_x[o0:o1], am.knot_y[o0:o1], am.knot_ypp[o0:o1])
        assert np.array_equal(sp(r), O.spline_eval(am.knot_x[o0:o1], am.knot_y[o0:o1], ypp_o, r))
        assert np.array_equal(sp.derivative(r), O.spline_deriv(am.knot_x[o0:o1], am.knot_y[o0:o1], ypp_o, r))


def test_native_fixed_nl_modes():
    am = S.AnlmModes.from_nl(3, 5, 500.0, 1000.0, nside=16)
    om = O.OracleModes(nmax=3, lmax=5, rmin=500.0, rmax=1000.0, nside=16)
    assert np.max(np.abs(am.knl / om.knl - 1)) < 1e-13
    assert np.array_equal(am.knot_x, om.knot_x)


def test_index_roundtrip(c1_pair):
    """getidx <-> getnlm (test/test_modes.jl:30-37)."""
    am, _ = c1_pair
    i = 1
    for n in range(1, am.nmax + 1):
        for l in range(int(am.lmax_n[n - 1]) + 1):
            for m in range(l + 1):
                assert S.getidx(am, n, l, m) == i
                assert S.getnlm(am, i) == (n, l, m)
                i += 1
    assert i - 1 == S.getnlmsize(am)
    assert max(am.lmax_n) == am.lmax and max(am.nmax_l) == am.nmax
