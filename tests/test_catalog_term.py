"""The catalogue term of cat2amln, pinned by mathematics (src/cat2anlm.jl:284-291,
derivations/sfb.tex:225-298: delta = n/nbar - W).

* One galaxy of weight w at radius r in pixel p, zero window, niter = 0: the map is
  w g_nl(r)/(nbar dOmega) on pixel p and the ring quadrature of map2alm multiplies by
  dOmega, so a_nlm = w g_nl(r) conj(Y_lm(theta_p, phi_p))/nbar exactly (Y from scipy, the
  pixel centre from the ring-scheme pix2ang; g from the spline to rounding and from the
  spherical Bessel functions to the spline's interpolation error).  A missing dOmega, a
  wrong nbar power, Y instead of conj(Y), another pixel's centre or a wrong (n, l, m)
  order each fail it.  With niter = 3 and lmax << nside the iteration moves it by less
  than 1e-5 (the tolerance of the reference's own single-voxel test).
* Galaxies sharing a pixel add, and a is linear in the weights.
* Shot noise: N galaxies uniform in the shell with the matching window give
  <nbar |a_nlm|^2> = 1 for l >= 1 (orthonormal g_nl r^2 dr and Y_lm dOmega, Poisson
  sampling), to the statistical error of the mode count.

The oracle runs on the CPU; the gpu-marked tests run the same checks through the HIP path.
"""
import numpy as np
import pytest
from scipy import special

import sfb_amd as S
from helpers import shell_volume, uniform_shell_catalog
from oracle import oracle as O


def say(*a):
    print(*a, flush=True)


def oracle_view(am):
    om = O.OracleModes.__new__(O.OracleModes)
    om.nside, om.lmax, om.nmax, om.lmax_n = am.nside, am.lmax, am.nmax, am.lmax_n
    om.knot_off, om.knot_x, om.knot_y, om.knot_ypp = am.knot_off, am.knot_x, am.knot_y, am.knot_ypp
    return om


def closed_form(om, r, th_p, ph_p, w, nbar, exact_g=False):
    """w g_nl(r) conj(Y_lm(theta_p, phi_p)) / nbar in a_nlm order."""
    out = np.zeros(om.nlmsize(), dtype=complex)
    for n in range(1, om.nmax + 1):
        for l in range(int(om.lmax_n[n - 1]) + 1):
            g = float((om.gnl_exact if exact_g else om.gnl)(n, l, np.array([r]))[0])
            for m in range(l + 1):
                y = special.sph_harm_y(l, m, th_p, ph_p)
                out[om.getidx(n, l, m) - 1] = w * g * np.conj(y) / nbar
    return out


def zero_window(om):
    return np.zeros(12 * om.nside ** 2), np.zeros((om.lmax + 1) * om.nmax)


GALAXIES = [  # (r, theta, phi, w): polar cap, both belts, the equator, a knot, near the edges
    (611.3, 0.05, 0.3, 1.0),
    (740.0, 1.2, 4.0, 0.7),
    (500.0 + 1e-9, np.pi / 2, 1.0, 2.5),
    (999.999, 2.9, 6.1, 1.3),
    (833.25, 2.0, 0.0, 0.9),
]


@pytest.fixture(scope="module")
def om():
    return O.OracleModes(kmax=0.03, rmin=500.0, rmax=1000.0, nside=16)


def test_oracle_single_galaxy_closed_form(om):
    nbar = 3e-4
    za, zb = zero_window(om)
    for r, th, ph, w in GALAXIES:
        rtp = np.array([[r], [th], [ph]])
        a = O.cat2amln(om, rtp, nbar, 1, za, zb, w=np.array([w]), niter=0)
        p = O.ang2pix_ring(om.nside, np.array([th]), np.array([ph]))[0]
        thp, php = O.pix2ang_ring(om.nside, np.array([p]))
        ref = closed_form(om, r, thp[0], php[0], w, nbar)
        say(f"galaxy r={r} pixel {p}: max|a - closed form| / max|a| = "
            f"{np.abs(a - ref).max() / np.abs(ref).max():.2e}")
        assert np.abs(a - ref).max() <= 1e-12 * np.abs(ref).max()
        # g from the spherical Bessel functions: the spline's interpolation error only
        ref_b = closed_form(om, r, thp[0], php[0], w, nbar, exact_g=True)
        assert np.abs(a - ref_b).max() <= 1e-6 * np.abs(ref_b).max()
        # the galaxy's own angles instead of the pixel centre are a different number
        ref_own = closed_form(om, r, th, ph, w, nbar)
        assert np.abs(a - ref_own).max() > 1e-4 * np.abs(ref_own).max()


def test_oracle_single_galaxy_niter3():
    om = O.OracleModes(nmax=2, lmax=4, rmin=500.0, rmax=1000.0, nside=256)
    nbar = 1e-3
    za, zb = zero_window(om)
    r, th, ph, w = 700.0, 1.0, 2.0, 1.0
    a = O.cat2amln(om, np.array([[r], [th], [ph]]), nbar, 1, za, zb, w=np.array([w]), niter=3)
    p = O.ang2pix_ring(om.nside, np.array([th]), np.array([ph]))[0]
    thp, php = O.pix2ang_ring(om.nside, np.array([p]))
    ref = closed_form(om, r, thp[0], php[0], w, nbar)
    assert np.linalg.norm(a - ref) <= 1e-5 * np.linalg.norm(ref)


def test_oracle_pixel_sums_and_linearity(om):
    nbar = 3e-4
    za, zb = zero_window(om)
    # three galaxies in one pixel and one elsewhere, then each alone
    cat = np.array([[600.0, 650.5, 777.7, 900.0], [1.0, 1.0, 1.0, 2.5], [2.0, 2.0, 2.0, 5.0]])
    w = np.array([0.5, 1.5, 2.0, 1.0])
    a = O.cat2amln(om, cat, nbar, 1, za, zb, w=w, niter=0)
    parts = sum(O.cat2amln(om, cat[:, i:i + 1], nbar, 1, za, zb, w=w[i:i + 1], niter=0) for i in range(4))
    assert np.abs(a - parts).max() <= 1e-13 * np.abs(a).max()
    a2 = O.cat2amln(om, cat, nbar, 1, za, zb, w=3 * w, niter=3)
    a1 = O.cat2amln(om, cat, nbar, 1, za, zb, w=w, niter=3)
    assert np.abs(a2 - 3 * a1).max() <= 1e-13 * np.abs(a2).max()


def shot_noise_mean(a, amodes_like, nbar):
    """Mean of nbar |a_nlm|^2 over the modes with l >= 1, and its statistical error."""
    vals = []
    i = 0
    for n in range(1, amodes_like.nmax + 1):
        for l in range(int(amodes_like.lmax_n[n - 1]) + 1):
            if l >= 1:
                vals.append(nbar * np.abs(a[i:i + l + 1]) ** 2)
            i += l + 1
    v = np.concatenate(vals)
    return float(v.mean()), float(v.std() / np.sqrt(v.size)), v.size


def test_oracle_shot_noise_c1():
    am = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    wmodes = S.ConfigurationSpaceModes(500.0, 1000.0, 1000, am.nside)
    win = S.win_rhat_ln(S.SeparableArray(np.ones(1000), np.ones(12 * 32 ** 2), "phi", "mask"), wmodes, am)
    cat = uniform_shell_catalog(100_000, 500.0, 1000.0, seed=11)
    nbar = cat.shape[1] / shell_volume(500.0, 1000.0)
    a = O.cat2amln(oracle_view(am), cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"))
    mean, err, nm = shot_noise_mean(a, am, nbar)
    say(f"C1 oracle: <nbar |a|^2> over {nm} modes (l >= 1) = {mean:.4f} +- {err:.4f}")
    assert abs(mean - 1.0) <= max(4 * err, 0.02)


@pytest.mark.gpu
def test_gpu_single_galaxy_closed_form():
    am = S.AnlmModes(0.03, 500.0, 1000.0, nside=16)
    om = oracle_view(am)
    nbar = 3e-4
    za, zb = zero_window(om)
    win = S.SeparableArray(za, zb.reshape(am.lmax + 1, am.nmax, order="F"), "mask", "w_ln")
    for prec in ("fp64", "mixed"):
        plan = S.Plan(am, niter=0, precision=prec)
        for r, th, ph, w in GALAXIES:
            say(prec, "galaxy", r, th, ph)
            a = plan.cat2amln(np.array([[r], [th], [ph]]), nbar, win, np.array([w]))
            p = O.ang2pix_ring(am.nside, np.array([th]), np.array([ph]))[0]
            thp, php = O.pix2ang_ring(am.nside, np.array([p]))
            ref = closed_form(om, r, thp[0], php[0], w, nbar)
            tol = 1e-12 if prec == "fp64" else 1e-6
            assert np.abs(a - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.gpu
def test_gpu_shot_noise_c2():
    am = S.AnlmModes(0.1, 500.0, 1500.0, nside=128)
    wmodes = S.ConfigurationSpaceModes(500.0, 1500.0, 1000, am.nside)
    win = S.win_rhat_ln(S.SeparableArray(np.ones(1000), np.ones(12 * 128 ** 2), "phi", "mask"), wmodes, am)
    cat = uniform_shell_catalog(1_000_000, 500.0, 1500.0, seed=12)
    nbar = cat.shape[1] / shell_volume(500.0, 1500.0)
    for prec in ("fp64", "mixed"):
        a = S.cat2amln(cat, am, nbar, win, precision=prec)
        mean, err, nm = shot_noise_mean(a, am, nbar)
        say(f"C2 {prec}: <nbar |a|^2> over {nm} modes (l >= 1) = {mean:.4f} +- {err:.4f}")
        assert abs(mean - 1.0) <= max(4 * err, 0.01)
