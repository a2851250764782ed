"""GPU parity of the HIP path against the CPU oracle (C restatement of the reference).

Bars: HEALPix pixel indices and the FP64 delta maps are bit-exact; map2alm and
cat2amln in FP64 agree with the oracle to 1e-10 of the per-(n,l)-block RMS
(the oracle and the GPU sum in different orders, so FP64 rounding differs).
"""
import numpy as np
import pytest

import sfb_amd as S
from helpers import rel_errors, separable_unit_window, shell_volume, uniform_shell_catalog
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c1():
    am = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    win, _ = separable_unit_window(S, am)
    cat = uniform_shell_catalog(100_000, 500.0, 1000.0, seed=0)
    nbar = cat.shape[1] / shell_volume(500.0, 1000.0)
    om = O.OracleModes.__new__(O.OracleModes)
    return am, win, cat, nbar


def _oracle_modes_from(am):
    """Oracle-side view of the product's tables (the same knots go to both sides)."""
    om = O.OracleModes.__new__(O.OracleModes)
    om.nside, om.lmax, om.nmax = am.nside, am.lmax, am.nmax
    om.lmax_n = am.lmax_n
    om.knot_off, om.knot_x, om.knot_y, om.knot_ypp = am.knot_off, am.knot_x, am.knot_y, am.knot_ypp
    return om


def test_ang2pix_bitexact():
    rng = np.random.default_rng(1)
    for nside in (1, 2, 32, 128, 256, 1024):
        say("ang2pix nside", nside)
        n = 1_000_000
        th = np.arccos(rng.uniform(-1, 1, n))
        ph = rng.uniform(-np.pi, 3 * np.pi, n)
        # edges: poles, the |z| = 2/3 belt boundary, phi wrap, pixel centres
        zb = np.arccos(2.0 / 3.0)
        edge_th = np.array([0.0, np.pi, zb, np.pi - zb, np.nextafter(zb, 0), np.nextafter(zb, 4), np.pi / 2])
        edge_ph = np.array([0.0, np.nextafter(2 * np.pi, 0), 2 * np.pi, -1e-300, -np.pi, np.pi, 7 * np.pi])
        eth, eph = np.meshgrid(edge_th, edge_ph)
        npix = 12 * nside * nside
        cp = np.arange(0, npix, max(1, npix // 20000))
        cth, cph = O.pix2ang_ring(nside, cp)
        th = np.concatenate([th, eth.ravel(), cth])
        ph = np.concatenate([ph, eph.ravel(), cph])
        gpu = S.ang2pix_ring(nside, th, ph)
        cpu = O.ang2pix_ring(nside, th, ph)
        bad = np.nonzero(gpu != cpu)[0]
        assert bad.size == 0, f"nside={nside}: {bad.size} mismatches, first {bad[:5]}"
        assert np.array_equal(gpu[-cp.size:], cp)


def test_delta_map_bitexact(c1):
    am, win, cat, nbar = c1
    plan = S.get_plan(am)
    om = _oracle_modes_from(am)
    for q in (0, 1, 17, am.nmaps // 2, am.nmaps - 1):
        say("delta map", q)
        g = plan.delta_map(cat, nbar, win, None, q)
        c = O.delta_map(om, cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"), None, q)
        assert np.array_equal(g, c), f"map {q}: max diff {np.max(np.abs(g - c))}"


def test_delta_map_crowded_pixels_bitexact():
    """Pixels far above the mean count (70000 galaxies in one pixel, 3000 in another),
    past the count clamp of the scatter's pixel ordering, plus out-of-range radii."""
    am = S.AnlmModes(0.03, 500.0, 1000.0, nside=8)
    win, _ = separable_unit_window(S, am, nr=100)
    rng = np.random.default_rng(21)
    base = uniform_shell_catalog(5000, 500.0, 1000.0, seed=22)
    crowd = np.stack([rng.uniform(500.0, 1000.0, 70000), np.full(70000, 0.3), np.full(70000, 1.1)])
    crowd2 = np.stack([rng.uniform(400.0, 1100.0, 3000), np.full(3000, 2.0), np.full(3000, 4.0)])
    cat = np.concatenate([base, crowd, crowd2], axis=1)
    w = rng.uniform(0.5, 1.5, cat.shape[1])
    nbar = cat.shape[1] / shell_volume(500.0, 1000.0)
    plan = S.get_plan(am)
    om = _oracle_modes_from(am)
    for q in (0, am.nmaps // 2, am.nmaps - 1):
        for ww in (w, None):  # given weights, and none (w = 1: no weight loads, padding masked by counts)
            say("crowded delta map", q, "weights" if ww is not None else "unit")
            g = plan.delta_map(cat, nbar, win, ww, q)
            c = O.delta_map(om, cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"), ww, q)
            assert np.array_equal(g, c), f"map {q}: max diff {np.max(np.abs(g - c))}"


def test_delta_map_large_tables_per_pixel_bitexact():
    """Knot tables too large for LDS (6000 knots: the sliced scatter needs (nk + 1) x 32 bytes
    of records per map in LDS, at most 160 KB) take the per-pixel scatter kernel; its delta
    maps are bit-exact too, with and without weights, radii inside and outside the knots."""
    nk, nside, lmax = 6000, 8, 3
    x = np.linspace(500.0, 1000.0, nk)
    scales = 20.0 + 5.0 * np.arange(lmax + 1)
    ys = [np.sin(x / s) for s in scales]
    ypps = [-y / s ** 2 for y, s in zip(ys, scales)]
    am = S.AnlmModes.from_tables(500.0, 1000.0, nside, np.array([lmax]), np.arange(lmax + 2) * nk,
                                 np.tile(x, lmax + 1), np.concatenate(ys), np.concatenate(ypps))
    rng = np.random.default_rng(31)
    win = S.SeparableArray(rng.uniform(0.5, 1.0, 12 * nside * nside), rng.uniform(0.1, 1.0, (lmax + 1, 1)),
                           "mask", "w_ln")
    cat = uniform_shell_catalog(20000, 450.0, 1050.0, seed=32)
    w = rng.uniform(0.5, 1.5, cat.shape[1])
    nbar = cat.shape[1] / shell_volume(450.0, 1050.0)
    plan = S.Plan(am, niter=3)
    om = _oracle_modes_from(am)
    for q in range(lmax + 1):
        for ww in (w, None):
            say("large-table delta map", q, "weights" if ww is not None else "unit")
            g = plan.delta_map(cat, nbar, win, ww, q)
            c = O.delta_map(om, cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"), ww, q)
            assert np.array_equal(g, c), f"map {q}: max diff {np.max(np.abs(g - c))}"


def test_delta_map_tied_radii_and_knot_radii_bitexact(c1):
    """Galaxies that share one radius inside one pixel (their weighted sum runs in input order,
    the order a stable sort by r leaves equal radii in), interleaved in the input with the rest,
    and galaxies exactly on spline knots (segment boundaries, the first and the last knot of a
    map, where the segment search and t = 0 or the boundary record decide): delta maps
    bit-exact against the oracle, with distinct weights so that any reordering would show."""
    am, win, _, nbar = c1
    rng = np.random.default_rng(41)
    q = am.nmaps // 3
    kx = am.knot_x[am.knot_off[q]:am.knot_off[q + 1]]
    parts = []
    for _ in range(40):  # 40 pixels, 50 galaxies each at one radius
        r0, th0, ph0 = rng.uniform(500.0, 1000.0), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)
        parts.append(np.stack([np.full(50, r0), np.full(50, th0), np.full(50, ph0)]))
    nk = kx.size  # every knot of map q, twice, at random angles
    parts.append(np.stack([np.tile(kx, 2), np.arccos(rng.uniform(-1, 1, 2 * nk)), rng.uniform(0, 2 * np.pi, 2 * nk)]))
    parts.append(uniform_shell_catalog(4000, 500.0, 1000.0, seed=42))
    cat = np.concatenate(parts, axis=1)
    cat = cat[:, rng.permutation(cat.shape[1])]
    w = rng.uniform(0.5, 1.5, cat.shape[1])
    plan = S.get_plan(am)
    om = _oracle_modes_from(am)
    for qq in (0, q, am.nmaps - 1):
        say("tied / knot radii delta map", qq)
        g = plan.delta_map(cat, nbar, win, w, qq)
        c = O.delta_map(om, cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"), w, qq)
        assert np.array_equal(g, c), f"map {qq}: max diff {np.max(np.abs(g - c))}"


def test_delta_map_weights_dense_bitexact(c1):
    am, win, cat, nbar = c1
    rng = np.random.default_rng(5)
    w = rng.uniform(0.5, 1.5, cat.shape[1])
    dense = win.dense()
    plan = S.get_plan(am)
    om = _oracle_modes_from(am)
    for q in (3, am.nmaps - 2):
        say("delta map dense+weights", q)
        g = plan.delta_map(cat, nbar, dense, w, q)
        c = O.delta_map(om, cat, nbar, 0, dense.reshape(-1, order="F"), None, w, q)
        assert np.array_equal(g, c)


@pytest.mark.parametrize("nside,lmax,niter", [(8, 16, 0), (8, 16, 3), (8, 32, 3), (16, 40, 3), (32, 43, 3)])
def test_map2alm_vs_oracle(nside, lmax, niter):
    say("map2alm", nside, lmax, niter)
    am = S.AnlmModes.from_nl(2, lmax, 500.0, 1000.0, nside=nside)
    plan = S.Plan(am, niter=niter)
    rng = np.random.default_rng(nside * 100 + lmax)
    maps = rng.normal(size=(3, 12 * nside * nside))
    got = plan.map2alm(maps)
    for i in range(3):
        ref = O.map2alm(nside, lmax, maps[i], niter)
        err = np.max(np.abs(got[i] - ref)) / np.max(np.abs(ref))
        assert err < 1e-12, err


def test_tile_search_leaves_results_unchanged(c1, monkeypatch):
    """The plan's phi-transform tile search picks launch shapes only: a plan built with it
    and one built with SFB_TUNE=0 (fixed defaults) return bit-identical a_nlm, FP64 and mixed."""
    am, win, cat, nbar = c1
    tuned = S.Plan(am, niter=3)
    monkeypatch.setenv("SFB_TUNE", "0")
    fixed = S.Plan(am, niter=3)
    say("tiles", tuned.tiles(), fixed.tiles())
    for prec in ("fp64", "mixed"):
        tuned.set_precision(prec)
        fixed.set_precision(prec)
        a = tuned.cat2amln(cat, nbar, win, None)
        b = fixed.cat2amln(cat, nbar, win, None)
        assert np.array_equal(a, b), prec


def test_cat2amln_c1_fp64_vs_oracle(c1):
    am, win, cat, nbar = c1
    say("cat2amln C1 fp64")
    got = S.cat2amln(cat, am, nbar, win)
    om = _oracle_modes_from(am)
    ref = O.cat2amln(om, cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"))
    e = rel_errors(got, ref, am)
    say(e)
    assert e["max_blk"] < 1e-10
