"""Adjoint at the nside-128 configuration (N = 1e6 uniform in 500-1500 Mpc/h, kmax 0.1):
L = sum |a - a_target|^2 with a_target from an independent catalogue (seed 1), the mixed
(FP32) forward and adjoint, checked on 1000 random galaxies against FP64 central finite
differences with step 1e-4 (Mpc/h in r; exact in w, where a is affine).

The finite differences are of the FP64 forward restricted to the moved galaxy: a is affine
in the catalogue (a = sum_i w_i a_i(r_i) - the window term), so moving galaxy i changes a
by w_i (a_i(r_i + h) - a_i(r_i - h)) exactly, with a_i the FP64 transform of that galaxy
alone (unit weight, zero window).  Differencing the full 1e6-galaxy transform instead
would lose the change (~1e-9 of |a|) in the FP64 rounding of a itself.  The error is
measured against the RMS of the 1000 finite differences (a single galaxy's derivative can
be near zero by cancellation over the modes); the bar is 1e-5.
"""
import numpy as np
import pytest

import sfb_amd as S
from helpers import separable_unit_window, shell_volume, uniform_shell_catalog

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rdot(c, a):
    return float(np.sum(c.real * a.real + c.imag * a.imag))


def test_c5_mixed_adjoint_against_fp64_finite_differences():
    am = S.AnlmModes(0.1, 500.0, 1500.0, nside=128)
    win, _ = separable_unit_window(S, am)
    cat = uniform_shell_catalog(1_000_000, 500.0, 1500.0, seed=0)
    N = cat.shape[1]
    nbar = N / shell_volume(500.0, 1500.0)
    w = np.random.default_rng(4).uniform(0.5, 1.5, N)
    p64 = S.Plan(am, precision="fp64")
    target = p64.cat2amln(uniform_shell_catalog(1_000_000, 500.0, 1500.0, seed=1), nbar, win)
    pmx = S.Plan(am, precision="mixed")
    a = pmx.cat2amln(cat, nbar, win, w)
    c = 2 * (a - target)
    dr, dw, dnbar = pmx.cat2amln_vjp(cat, nbar, c, w)
    dr64, dw64, _ = p64.cat2amln_vjp(cat, nbar, c, w)
    zero = S.SeparableArray(np.zeros(12 * am.nside ** 2), np.zeros((am.lmax + 1, am.nmax)), "mask", "w_ln")
    idx = np.random.default_rng(5).choice(N, 1000, replace=False)
    h = 1e-4
    fd_w = np.empty(idx.size)
    fd_r = np.empty(idx.size)
    one = np.ones(1)
    for k, i in enumerate(idx):
        if k % 100 == 0:
            say(f"finite differences: galaxy {k} of {idx.size}")
        g = cat[:, i:i + 1].copy()
        fd_w[k] = rdot(c, p64.cat2amln(g, nbar, zero, one))
        gp, gm = g.copy(), g.copy()
        gp[0, 0] += h
        gm[0, 0] -= h
        fd_r[k] = w[i] * rdot(c, p64.cat2amln(gp, nbar, zero, one) - p64.cat2amln(gm, nbar, zero, one)) / (2 * h)
    for name, got, ref in (("dL/dw mixed", dw[idx], fd_w), ("dL/dr mixed", dr[idx], fd_r),
                           ("dL/dw fp64", dw64[idx], fd_w), ("dL/dr fp64", dr64[idx], fd_r)):
        rms = np.sqrt(np.mean(ref ** 2))
        e = np.abs(got - ref) / rms
        big = np.abs(ref) >= 0.1 * rms
        say(f"{name}: max |adjoint - FD| / RMS(FD) = {e.max():.2e}, median {np.median(e):.2e}, "
            f"max relative where |FD| >= RMS/10: {(np.abs(got - ref) / np.abs(ref))[big].max():.2e}")
        assert e.max() <= 1e-5, name
