"""nside 512, the largest nside a plan accepts: the synthesis direction splits the polar
ring pairs between the radix-4 inverse (rings i <= 256) and the direct sum (i = 257..511).
anlm2field against the oracle's alm2map per shell, and the adjoint's dot-product identity
(a is affine in the weights, so Re<c, a(w + d) - a(w)> = <dL/dw, d> for any niter)."""
import numpy as np
import pytest

import sfb_amd as S
from helpers import separable_unit_window, shell_volume, uniform_shell_catalog
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def big():
    return S.AnlmModes.from_nl(2, 8, 500.0, 1000.0, nside=512)


def test_anlm2field_nside512(big):
    am = big
    wm = S.ConfigurationSpaceModes(500.0, 1000.0, 2, 512)
    L = am.lmax
    rng = np.random.default_rng(51)
    fnlm = rng.standard_normal(S.getnlmsize(am)) + 1j * rng.standard_normal(S.getnlmsize(am))
    say("anlm2field nside 512 vs oracle alm2map")
    out = S.anlm2field(fnlm, wm, am)
    for k in range(wm.nr):
        alm = np.zeros((L + 1) * (L + 2) // 2, dtype=complex)
        for n in range(1, am.nmax + 1):
            for l in range(int(am.lmax_n[n - 1]) + 1):
                g = am.basisfunctions(n, l, wm.r[k])
                for m in range(l + 1):
                    alm[m * (2 * L + 1 - m) // 2 + l] += g * fnlm[S.getidx(am, n, l, m) - 1]
        expect = O.alm2map(512, L, alm)
        assert np.max(np.abs(out[k] - expect)) <= 1e-11 * np.max(np.abs(expect)), k


def test_vjp_nside512_dot_product(big):
    am = big
    win, _ = separable_unit_window(S, am, nr=200)
    cat = uniform_shell_catalog(20000, 500.0, 1000.0, seed=52)
    nbar = cat.shape[1] / shell_volume(500.0, 1000.0)
    rng = np.random.default_rng(53)
    w = rng.uniform(0.5, 1.5, cat.shape[1])
    c = rng.standard_normal(S.getnlmsize(am)) + 1j * rng.standard_normal(S.getnlmsize(am))
    p = S.get_plan(am)
    say("vjp nside 512 dot-product identity")
    _, dw, _ = p.cat2amln_vjp(cat, nbar, c, w)
    d = rng.standard_normal(w.size)
    a1, a0 = p.cat2amln(cat, nbar, win, w + d), p.cat2amln(cat, nbar, win, w)
    lhs = float(np.sum(c.real * (a1 - a0).real + c.imag * (a1 - a0).imag))
    assert lhs == pytest.approx(float(dw @ d), rel=1e-9)
