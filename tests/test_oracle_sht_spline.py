"""The oracle's spline and spherical-harmonic transform, pinned by mathematics.

Spline1D{3} (src/Splines.jl:174-377): interpolates its knots, reproduces a
quadratic exactly (the yderiv_k2 end slopes are exact for a quadratic), is C1 and
C2 at the knots, and its second derivatives solve the reference's tridiagonal
system, checked here by a dense numpy solve of the same equations.

Ring SHT (the map2alm behind src/healpix_helpers.jl:50-56): lambda_lm and the
synthesis against scipy's Y_lm at the pixel centres; the analysis against a
brute-force O(npix nlm) sum; niter refinement converging to the band-limited
input; a single-pixel map giving Omega_p Y*_lm (src/cat2anlm.jl:153-196 pins this
uniform quadrature).
"""
import numpy as np
import pytest
from scipy import special

from oracle import oracle as O


def knots(n, a=2.0, b=7.0, seed=0, jitter=0.0):
    x = np.linspace(a, b, n)
    if jitter:
        x[1:-1] += np.random.default_rng(seed).uniform(-jitter, jitter, n - 2) * (b - a) / n
    return x


def test_spline_interpolates_and_is_smooth():
    x = knots(40, jitter=0.3)
    y = np.sin(x) * np.exp(0.1 * x)
    ypp = O.spline_build(x, y)
    assert np.allclose(O.spline_eval(x, y, ypp, x[:-1]), y[:-1], rtol=0, atol=1e-14)
    # C1 and C2 across interior knots
    eps = 1e-7
    for i in range(1, x.size - 1):
        dl = O.spline_deriv(x, y, ypp, [x[i] - eps])[0]
        dr = O.spline_deriv(x, y, ypp, [x[i] + eps])[0]
        assert abs(dl - dr) < 1e-5
    # ypp is the second derivative: finite differences of the value
    xm = 0.5 * (x[5] + x[6])
    h = 1e-4
    f = O.spline_eval(x, y, ypp, [xm - h, xm, xm + h])
    fpp = (f[0] - 2 * f[1] + f[2]) / h ** 2
    t = (xm - x[5]) / (x[6] - x[5])
    assert fpp == pytest.approx(ypp[5] * (1 - t) + ypp[6] * t, rel=1e-5)


def test_spline_reproduces_quadratic():
    x = knots(25, jitter=0.2, seed=3)
    y = 0.7 * x ** 2 - 3 * x + 2
    ypp = O.spline_build(x, y)
    assert np.allclose(ypp, 1.4, atol=1e-10)
    r = np.random.default_rng(1).uniform(x[0], x[-1], 1000)
    assert np.allclose(O.spline_eval(x, y, ypp, r), 0.7 * r ** 2 - 3 * r + 2, atol=1e-11)
    # boundary extrapolation returns the end values (Splines.jl:72-78)
    assert O.spline_eval(x, y, ypp, [x[0] - 1, x[-1] + 1]).tolist() == [y[0], y[-1]]


def test_spline_matches_dense_solve_of_reference_equations():
    x = knots(12, jitter=0.3, seed=5)
    y = np.cos(1.3 * x)
    ypp = O.spline_build(x, y)
    n = x.size
    A = np.zeros((n, n))
    rhs = np.zeros(n)

    def yd(i, j, k):
        return ((y[k] - y[i]) * (x[j] - x[i]) ** 2 - (y[j] - y[i]) * (x[k] - x[i]) ** 2) / (
            (x[k] - x[i]) * (x[j] - x[k]) * (x[j] - x[i]))

    # interior: h_{i-1} M_{i-1} + 2(h_{i-1}+h_i) M_i + h_i M_{i+1} = 6(d_i - d_{i-1})
    for i in range(1, n - 1):
        A[i, i - 1] = x[i] - x[i - 1]
        A[i, i] = 2 * (x[i + 1] - x[i - 1])
        A[i, i + 1] = x[i + 1] - x[i]
        rhs[i] = 6 * (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - 6 * (y[i] - y[i - 1]) / (x[i] - x[i - 1])
    # clamped ends with slopes from the 3-point quadratic: s'(x0) = yp1, s'(xN) = ypN
    h0, hN = x[1] - x[0], x[-1] - x[-2]
    A[0, 0], A[0, 1] = 2 * h0, h0
    rhs[0] = 6 * ((y[1] - y[0]) / h0 - yd(0, 1, 2))
    A[-1, -1], A[-1, -2] = 2 * hN, hN
    rhs[-1] = 6 * (yd(n - 1, n - 2, n - 3) - (y[-1] - y[-2]) / hN)
    M = np.linalg.solve(A, rhs)
    assert np.allclose(ypp, M, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("nside,lmax", [(4, 8), (8, 12)])
def test_lambda_and_synthesis_vs_scipy(nside, lmax):
    lam = O.lambda_table(nside, lmax)
    for ring in (1, nside, 2 * nside, 3 * nside + 1, 4 * nside - 1):
        st, npr, z, phi0 = O.ring_info(nside, ring)
        th = np.arccos(z)
        for l in range(lmax + 1):
            for m in range(l + 1):
                ref = special.sph_harm_y(l, m, th, 0.0).real
                assert lam[ring - 1, O.alm_index(lmax, l, m)] == pytest.approx(ref, abs=1e-13)
    rng = np.random.default_rng(nside)
    nalm = (lmax + 1) * (lmax + 2) // 2
    alm = rng.normal(size=nalm) + 1j * rng.normal(size=nalm)
    alm[[O.alm_index(lmax, l, 0) for l in range(lmax + 1)]] = alm[[O.alm_index(lmax, l, 0) for l in range(lmax + 1)]].real
    th, ph = O.pix2ang_ring(nside, np.arange(12 * nside * nside))
    mb = np.zeros(th.size)
    for l in range(lmax + 1):
        for m in range(l + 1):
            mb += ((alm[O.alm_index(lmax, l, m)] * special.sph_harm_y(l, m, th, ph)).real * (1 if m == 0 else 2))
    assert np.max(np.abs(O.alm2map(nside, lmax, alm) - mb)) < 1e-12 * np.max(np.abs(mb))


def test_analysis_vs_bruteforce_quadrature():
    nside, lmax = 8, 20  # lmax > 2 nside: short polar rings alias
    npix = 12 * nside * nside
    m = np.random.default_rng(2).normal(size=npix)
    th, ph = O.pix2ang_ring(nside, np.arange(npix))
    got = O.map2alm(nside, lmax, m, niter=0)
    for l in range(lmax + 1):
        for mm in range(l + 1):
            ref = 4 * np.pi / npix * np.sum(m * np.conj(special.sph_harm_y(l, mm, th, ph)))
            assert abs(got[O.alm_index(lmax, l, mm)] - ref) < 1e-12


def test_niter_converges_on_bandlimited_input():
    nside, lmax = 16, 24
    nalm = (lmax + 1) * (lmax + 2) // 2
    rng = np.random.default_rng(9)
    alm = rng.normal(size=nalm) + 1j * rng.normal(size=nalm)
    alm[[O.alm_index(lmax, l, 0) for l in range(lmax + 1)]] = alm[[O.alm_index(lmax, l, 0) for l in range(lmax + 1)]].real
    mp = O.alm2map(nside, lmax, alm)
    errs = [np.max(np.abs(O.map2alm(nside, lmax, mp, niter=k) - alm)) for k in (0, 1, 2, 3)]
    assert errs[0] > 1e-4
    assert errs[1] < errs[0] * 0.2 and errs[2] < errs[1] * 0.2 and errs[3] < errs[2] * 0.2


def test_single_pixel_map_is_quadrature_weight_times_conj_ylm():
    nside, lmax = 256, 3
    npix = 12 * nside * nside
    for j in (1, 3 * nside ** 2 + 7, npix // 2 + 5):
        mp = np.zeros(npix)
        mp[j] = 1.0
        th, ph = O.pix2ang_ring(nside, [j])
        got = O.map2alm(nside, lmax, mp, niter=3)
        for l in range(lmax + 1):
            for mm in range(l + 1):
                ref = 4 * np.pi / npix * np.conj(special.sph_harm_y(l, mm, th[0], ph[0]))
                assert got[O.alm_index(lmax, l, mm)] == pytest.approx(ref, rel=1e-5, abs=1e-12)
