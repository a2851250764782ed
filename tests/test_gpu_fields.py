"""Window and field transforms on the device: win_rhat_ln (src/windows.jl:244-270),
field2anlm (src/cat2anlm.jl:323-358) and anlm2field (:383-422).

Pins:
* win_rhat_ln on the GPU against the host evaluation (same formula, NumPy), dense and
  separable, NaN pattern included;
* field2anlm against the oracle's map2alm per shell contracted with dr r^2 g_nl(r), and
  anlm2field against the oracle's alm2map per shell;
* the reference's own tests re-expressed: the single-voxel field (test_cat2anlm.jl:153-196,
  rtol 1e-5), field2anlm_v1 (= -cat2amln of an empty catalogue with win_rhat_ln(f))
  against field2anlm_v2 (:235-258, Julia's default rtol sqrt(eps)), and the
  field2anlm / anlm2field round trip (:261-289, rtol 1e-4).
"""
import numpy as np
import pytest
from scipy import special

import sfb_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def julia_isapprox(a, b, rtol):
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


@pytest.fixture(scope="module")
def small():
    am = S.AnlmModes(0.03, 500.0, 1000.0, nside=16)
    wm = S.ConfigurationSpaceModes(500.0, 1000.0, 80, 16)
    return am, wm


def test_win_rhat_ln_dense_and_separable(small):
    am, wm = small
    say("win_rhat_ln GPU vs host")
    rng = np.random.default_rng(30)
    win = rng.uniform(0.0, 1.0, (wm.nr, wm.npix))
    host = S.win_rhat_ln(win, wm, am)
    gpu = S.win_rhat_ln(win, wm, am, device=0)
    assert np.array_equal(np.isnan(host), np.isnan(gpu))
    fin = np.isfinite(host)
    assert np.max(np.abs(host[fin] - gpu[fin])) <= 1e-12 * np.max(np.abs(host[fin]))
    sep = S.SeparableArray(rng.uniform(0.5, 1.5, wm.nr), rng.uniform(0, 1, wm.npix))
    hs, gs = S.win_rhat_ln(sep, wm, am), S.win_rhat_ln(sep, wm, am, device=0)
    assert np.array_equal(hs.mask, gs.mask)
    assert np.array_equal(np.isnan(hs.w_ln), np.isnan(gs.w_ln))
    f = np.isfinite(hs.w_ln)
    assert np.max(np.abs(hs.w_ln[f] - gs.w_ln[f])) <= 1e-12 * np.max(np.abs(hs.w_ln[f]))


def per_shell_gnl(am, r):
    return {(n, l): am.basisfunctions(n, l, r) for n in range(1, am.nmax + 1)
            for l in range(int(am.lmax_n[n - 1]) + 1)}


def test_field_transforms_against_oracle():
    am = S.AnlmModes.from_nl(4, 6, 500.0, 1000.0, nside=8)
    wm = S.ConfigurationSpaceModes(500.0, 1000.0, 12, 8)
    L, npix = am.lmax, wm.npix
    rng = np.random.default_rng(31)
    f = rng.standard_normal((wm.nr, npix))
    g = per_shell_gnl(am, wm.r)
    say("field2anlm vs oracle map2alm per shell")
    ref = np.zeros(S.getnlmsize(am), dtype=complex)
    for k in range(wm.nr):
        alm = O.map2alm(8, L, f[k], niter=3)
        for n in range(1, am.nmax + 1):
            for l in range(int(am.lmax_n[n - 1]) + 1):
                for m in range(l + 1):
                    idx = m * (2 * L + 1 - m) // 2 + l
                    ref[S.getidx(am, n, l, m) - 1] += wm.dr * wm.r[k] ** 2 * g[(n, l)][k] * alm[idx]
    got = S.field2anlm(f, wm, am)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))
    say("anlm2field vs oracle alm2map per shell")
    fnlm = rng.standard_normal(S.getnlmsize(am)) + 1j * rng.standard_normal(S.getnlmsize(am))
    out = S.anlm2field(fnlm, wm, am)
    for k in range(wm.nr):
        alm = np.zeros((L + 1) * (L + 2) // 2, dtype=complex)
        for n in range(1, am.nmax + 1):
            for l in range(int(am.lmax_n[n - 1]) + 1):
                for m in range(l + 1):
                    alm[m * (2 * L + 1 - m) // 2 + l] += g[(n, l)][k] * fnlm[S.getidx(am, n, l, m) - 1]
        expect = O.alm2map(8, L, alm)
        assert np.max(np.abs(out[k] - expect)) <= 1e-11 * np.max(np.abs(expect)), k


def test_field2anlm_single_voxel():
    """test_cat2anlm.jl:153-196 with the device field2anlm."""
    nside, nr = 256, 100
    am = S.AnlmModes.from_nl(2, 3, 500.0, 1000.0, nside=nside)
    wm = S.ConfigurationSpaceModes(500.0, 1000.0, nr, nside)
    npix = 12 * nside ** 2
    f = np.zeros((nr, npix))
    omega = 4 * np.pi / npix
    for j in range(1, npix + 1, 3 * nside ** 2 + 7):
        for i in range(1, nr + 1, 41):
            say("voxel", i, j)
            f[i - 1, j - 1] = 1.0
            a1 = S.field2anlm(f, wm, am)
            th, ph = O.pix2ang_ring(nside, [j - 1])
            r = wm.r[i - 1]
            a0 = np.empty_like(a1)
            for idx in range(a0.size):
                n, l, m = S.getnlm(am, idx + 1)
                y = special.sph_harm_y(l, m, th[0], ph[0])
                a0[idx] = wm.dr * r ** 2 * am.basisfunctions(n, l, r) * np.conj(y) * omega
            assert julia_isapprox(a0, a1, 1e-5)
            f[i - 1, j - 1] = 0.0


@pytest.fixture(scope="module")
def ten():
    am = S.AnlmModes.from_nl(10, 10, 500.0, 1000.0, nside=32)
    wm = S.ConfigurationSpaceModes.from_amodes(am, 100)
    f1 = np.random.default_rng(32).random((wm.nr, wm.npix))
    return am, wm, f1


def test_field2anlm_versions(ten):
    """test_cat2anlm.jl:235-258: v1 = -cat2amln(empty, win_rhat_ln(f)) against v2."""
    am, wm, f1 = ten
    say("field2anlm v1 vs v2")
    v2 = S.field2anlm(f1, wm, am)
    w = S.win_rhat_ln(f1, wm, am, device=0)
    v1 = -S.cat2amln(np.zeros((3, 0)), am, 1.0, np.nan_to_num(w, nan=0.0), [])
    assert julia_isapprox(v1, v2, np.sqrt(np.finfo(float).eps))


def test_anlm2field_roundtrip(ten):
    """test_cat2anlm.jl:261-289."""
    am, wm, f1 = ten
    say("field2anlm / anlm2field round trip")
    f1_nlm = S.field2anlm(f1, wm, am)
    f2 = S.anlm2field(f1_nlm, wm, am)
    f2_nlm = S.field2anlm(f2, wm, am)
    f3 = S.anlm2field(f2_nlm, wm, am)
    f3_nlm = S.field2anlm(f3, wm, am)
    assert julia_isapprox(f1, f2, 1.0)
    assert julia_isapprox(f2, f3, 1e-4)
    assert julia_isapprox(f1_nlm, f2_nlm, 1e-4)
    assert julia_isapprox(f2_nlm, f3_nlm, 1e-4)
