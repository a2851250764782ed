"""The reference's own cat2amln known-answer tests, re-expressed.

* "basis function transform" (test/test_cat2anlm.jl:32-67): a window equal to one
  basis function g_nl(r) Y^R_lm(pixel) goes through win_rhat_ln and cat2amln with
  an empty catalogue; -a gives 1 at that mode (m = 0) or (-1)^m/sqrt(2) (m > 0,
  real harmonics Y^R_lm = (-1)^m sqrt(2) Re Y_lm), and at most 2e-3 elsewhere.
* "field2anlm()" (test/test_cat2anlm.jl:153-196): a single-voxel field gives
  a_nlm = dr r^2 g_nl(r) Omega_p Y*_lm(pixel centre), rtol 1e-5; field2anlm_v1 is
  -cat2amln(empty, win_rhat_ln(f)) (src/cat2anlm.jl:315-321).
* win_rhat_ln of basis windows is ~delta_nn' (test/test_windows.jl:446-478), and
  the dense and separable forms agree (test/test_windows.jl:393-399).

Each runs on the CPU oracle; the GPU variants (marked gpu) run the HIP path.
"""
import numpy as np
import pytest
from scipy import special

import sfb_amd as S
from oracle import oracle as O


def real_ylm(l, m, th, ph):
    y = special.sph_harm_y(l, m, th, ph)
    return y.real if m == 0 else (-1) ** m * np.sqrt(2) * y.real


def oracle_view(am):
    om = O.OracleModes.__new__(O.OracleModes)
    om.nside, om.lmax, om.nmax, om.lmax_n = am.nside, am.lmax, am.nmax, am.lmax_n
    om.knot_off, om.knot_x, om.knot_y, om.knot_ypp = am.knot_off, am.knot_x, am.knot_y, am.knot_ypp
    return om


@pytest.fixture(scope="module")
def basis_setup():
    am = S.AnlmModes(0.02, 0.0, 1000.0, nside=32)
    wmodes = S.ConfigurationSpaceModes(0.0, 1000.0, 200, am.nside)
    npix = 12 * am.nside ** 2
    th, ph = O.pix2ang_ring(am.nside, np.arange(npix))
    return am, wmodes, th, ph


BASIS_CASES = [(1, 0, 0), (2, 3, 1), (1, 7, 4), (3, 5, 5), (4, 2, 2)]


def basis_window(am, wmodes, th, ph, n, l, m):
    g = am.basisfunctions(n, l, wmodes.r)
    return np.multiply.outer(g, real_ylm(l, m, th, ph))  # nr x npix


def check_basis(anlm, am, n, l, m):
    a = -anlm
    idx = S.getidx(am, n, l, m) - 1
    expect = 1.0 if m == 0 else (-1) ** m / np.sqrt(2)
    assert a[idx] == pytest.approx(expect, abs=1e-3)
    rest = np.delete(a, idx)
    assert np.max(np.abs(rest)) <= 2e-3


@pytest.mark.parametrize("n,l,m", BASIS_CASES)
def test_basis_function_transform_oracle(basis_setup, n, l, m):
    am, wmodes, th, ph = basis_setup
    win = S.win_rhat_ln(basis_window(am, wmodes, th, ph, n, l, m), wmodes, am)
    w = np.nan_to_num(win, nan=0.0)
    anlm = O.cat2amln(oracle_view(am), np.zeros((3, 0)), 3e-4, 0, w.reshape(-1, order="F"))
    check_basis(anlm, am, n, l, m)


@pytest.mark.gpu
@pytest.mark.parametrize("n,l,m", BASIS_CASES)
def test_basis_function_transform_gpu(basis_setup, n, l, m):
    am, wmodes, th, ph = basis_setup
    win = np.nan_to_num(S.win_rhat_ln(basis_window(am, wmodes, th, ph, n, l, m), wmodes, am), nan=0.0)
    anlm = S.cat2amln(np.zeros((3, 0)), am, 3e-4, win, [])
    check_basis(anlm, am, n, l, m)
    ref = O.cat2amln(oracle_view(am), np.zeros((3, 0)), 3e-4, 0, win.reshape(-1, order="F"))
    assert np.max(np.abs(anlm - ref)) < 1e-12


@pytest.fixture(scope="module")
def voxel_setup():
    am = S.AnlmModes.from_nl(2, 3, 500.0, 1000.0, nside=256)
    wmodes = S.ConfigurationSpaceModes(500.0, 1000.0, 100, am.nside)
    return am, wmodes


VOXELS = [(0, 0), (41, 3 * 256 ** 2 + 7), (82, 2 * (3 * 256 ** 2 + 7))]


def voxel_expectation(am, wmodes, i, j):
    nside = am.nside
    npix = 12 * nside ** 2
    th, ph = O.pix2ang_ring(nside, [j])
    r = wmodes.r[i]
    out = np.empty(S.getnlmsize(am), dtype=complex)
    for idx in range(out.size):
        n, l, m = S.getnlm(am, idx + 1)
        out[idx] = wmodes.dr * r ** 2 * am.basisfunctions(n, l, r) * np.conj(
            special.sph_harm_y(l, m, th[0], ph[0])) * 4 * np.pi / npix
    return out


def julia_isapprox(a, b, rtol):
    """Julia's isapprox on arrays: norm(a - b) <= rtol * max(norm(a), norm(b))."""
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


def voxel_field(am, wmodes, i, j):
    f = np.zeros((wmodes.nr, 12 * am.nside ** 2))
    f[i, j] = 1.0
    return S.win_rhat_ln(f, wmodes, am)


@pytest.mark.parametrize("i,j", VOXELS)
def test_single_voxel_oracle(voxel_setup, i, j):
    am, wmodes = voxel_setup
    win = voxel_field(am, wmodes, i, j)
    a = -O.cat2amln(oracle_view(am), np.zeros((3, 0)), 1.0, 0, win.reshape(-1, order="F"))
    assert julia_isapprox(a, voxel_expectation(am, wmodes, i, j), 1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("i,j", VOXELS)
def test_single_voxel_gpu(voxel_setup, i, j):
    am, wmodes = voxel_setup
    win = voxel_field(am, wmodes, i, j)
    a = -S.cat2amln(np.zeros((3, 0)), am, 1.0, win, [])
    assert julia_isapprox(a, voxel_expectation(am, wmodes, i, j), 1e-5)


def test_win_rhat_ln_of_basis_is_delta():
    rmin, rmax, nmax, lmax, nr, nside = 500.0, 1000.0, 5, 6, 250, 64
    am = S.AnlmModes.from_nl(nmax, lmax, rmin, rmax, nside=nside)
    wmodes = S.ConfigurationSpaceModes(rmin, rmax, nr, nside)
    win = np.ones((nr, 12 * nside ** 2))
    p = 0
    for n in range(1, nmax + 1):
        for l in range(lmax + 1):
            win[:, p] = am.basisfunctions(n, l, wmodes.r)
            p += 1
    W = S.win_rhat_ln(win, wmodes, am)
    for n in range(1, nmax + 1):
        for l in range(lmax + 1):
            p = 0
            for n2 in range(1, nmax + 1):
                for l2 in range(lmax + 1):
                    if l == l2:
                        assert W[p, l, n - 1] == pytest.approx(1.0 if n == n2 else 0.0, abs=1e-4)
                    p += 1


def test_win_rhat_ln_dense_equals_separable():
    am = S.AnlmModes(0.03, 500.0, 1000.0, nside=8)
    wmodes = S.ConfigurationSpaceModes(500.0, 1000.0, 120, 8)
    rng = np.random.default_rng(0)
    phi = rng.uniform(0.5, 1.5, 120)
    mask = (rng.uniform(size=768) > 0.3).astype(float)
    sep = S.win_rhat_ln(S.SeparableArray(phi, mask), wmodes, am)
    dense = S.win_rhat_ln(np.multiply.outer(phi, mask), wmodes, am)
    assert np.allclose(sep.dense(), dense, rtol=1e-10, equal_nan=True)
    v1 = S.integrate_window(S.SeparableArray(phi, mask), wmodes)
    v2 = S.integrate_window(np.multiply.outer(phi, mask), wmodes)
    assert v1 == pytest.approx(v2, rel=1e-12)
