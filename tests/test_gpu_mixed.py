"""Mixed-precision path (SFB_PREC_MIXED) against the FP64 oracle and the FP64 GPU path.

The path: FP32 spline values with FP64 pixel sums, FP32 delta maps and phi-transform,
the first Legendre analysis on FP64 MFMA over FP32 ring coefficients, FP32 MFMA for the
Jacobi corrections, and the monopoles of the l = 0 maps (which cancel to the radial
sampling noise) corrected with their FP64 full-sky sums.  The bar is the one the path is
built to: for every (n, l) block of a_nlm, max_m |a - a_ref| <= 1e-6 x the block's RMS
|a_ref|; the corrected monopoles are held to 1e-7 (uncorrected they miss by up to 1e-4).
"""
import numpy as np
import pytest

import sfb_amd as S
from helpers import separable_unit_window, shell_volume, uniform_shell_catalog
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BLOCK_BAR = 1e-6


def say(*a):
    print(*a, flush=True)


def oracle_view(am):
    om = O.OracleModes.__new__(O.OracleModes)
    om.nside, om.lmax, om.nmax, om.lmax_n = am.nside, am.lmax, am.nmax, am.lmax_n
    om.knot_off, om.knot_x, om.knot_y, om.knot_ypp = am.knot_off, am.knot_x, am.knot_y, am.knot_ypp
    return om


def blocks(am):
    """(q, n, l, first, end) of every (n, l) block of a_nlm, in map order."""
    out, i, q = [], 0, 0
    for n in range(1, am.nmax + 1):
        for l in range(int(am.lmax_n[n - 1]) + 1):
            out.append((q, n, l, i, i + l + 1))
            i += l + 1
            q += 1
    return out


def block_errors(a, ref, am, qs=None):
    """Per block: max_m |a - ref| / RMS_m |ref|, for the maps qs (all when None)."""
    e = {}
    for q, n, l, i0, i1 in blocks(am):
        if qs is not None and q not in qs:
            continue
        rms = np.sqrt(np.mean(np.abs(ref[i0:i1]) ** 2))
        e[(n, l)] = float(np.abs(a[i0:i1] - ref[i0:i1]).max() / max(rms, 1e-300))
    return e


def worst(e, k=3):
    return sorted(e.items(), key=lambda kv: -kv[1])[:k]


def test_mixed_c1_against_oracle():
    am = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    win, _ = separable_unit_window(S, am)
    cat = uniform_shell_catalog(100_000, 500.0, 1000.0, seed=0)
    w = np.random.default_rng(3).uniform(0.5, 1.5, cat.shape[1])
    nbar = cat.shape[1] / shell_volume(500.0, 1000.0)
    ref = O.cat2amln(oracle_view(am), cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"), w=w)
    for weights in (w, None):
        r = ref if weights is not None else O.cat2amln(oracle_view(am), cat, nbar, 1, win.mask,
                                                       win.w_ln.reshape(-1, order="F"))
        a = S.Plan(am, precision="mixed").cat2amln(cat, nbar, win, weights)
        assert np.all(np.isfinite(a))
        e = block_errors(a, r, am)
        say("C1 mixed", "weights" if weights is not None else "unit", "worst blocks", worst(e))
        assert max(e.values()) <= BLOCK_BAR
        mono = [v for (n, l), v in e.items() if l == 0]
        say("   l = 0 blocks (FP64-corrected monopoles): worst", max(mono))
        assert max(mono) <= 1e-7


@pytest.fixture(scope="module")
def c2():
    am = S.AnlmModes(0.1, 500.0, 1500.0, nside=128)
    win, _ = separable_unit_window(S, am)
    cat = uniform_shell_catalog(1_000_000, 500.0, 1500.0, seed=0)
    nbar = cat.shape[1] / shell_volume(500.0, 1500.0)
    p = S.Plan(am, precision="fp64")
    a64 = p.cat2amln(cat, nbar, win)
    p.set_precision("mixed")
    amx = p.cat2amln(cat, nbar, win)
    return am, win, cat, nbar, p, a64, amx


def test_mixed_c2_every_block_against_fp64_gpu(c2):
    """All 2514 maps against the FP64 GPU path (itself pinned to the oracle below)."""
    am, win, cat, nbar, p, a64, amx = c2
    e = block_errors(amx, a64, am)
    say("C2 mixed vs FP64 GPU, worst blocks", worst(e, 5))
    assert len(e) == am.nmaps
    assert max(e.values()) <= BLOCK_BAR
    mono = max(v for (n, l), v in e.items() if l == 0)
    say("   l = 0 blocks (FP64-corrected monopoles): worst", mono)
    assert mono <= 1e-7


def test_mixed_c2_sample_against_oracle(c2):
    """Maps at low and high n and l (n = 1 at l = 0..7 and l = 133..140, a middle n at its
    top l, the last two n) against the oracle, both precisions."""
    am, win, cat, nbar, p, a64, amx = c2
    bl = blocks(am)
    first_of = {}
    for q, n, l, _, _ in bl:
        first_of.setdefault(n, q)
    lmax1 = int(am.lmax_n[0])
    nmid = am.nmax // 2
    ranges = [(0, 8), (lmax1 - 7, lmax1 + 1),
              (first_of[nmid] + int(am.lmax_n[nmid - 1]) - 3, first_of[nmid] + int(am.lmax_n[nmid - 1]) + 1),
              (first_of[am.nmax - 1], am.nmaps)]
    om = oracle_view(am)
    qs = set()
    ref = np.full(a64.size, np.nan, dtype=complex)
    for q0, q1 in ranges:
        say(f"oracle maps {q0}..{q1 - 1}")
        r = O.cat2amln(om, cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"), q_range=(q0, q1))
        for q, n, l, i0, i1 in bl:
            if q0 <= q < q1:
                ref[i0:i1] = r[i0:i1]
                qs.add(q)
    e64 = block_errors(a64, ref, am, qs)
    emx = block_errors(amx, ref, am, qs)
    say(f"C2 oracle sample: {len(qs)} maps; FP64 worst {worst(e64)}; mixed worst {worst(emx)}")
    assert max(e64.values()) <= 1e-10
    assert max(emx.values()) <= BLOCK_BAR


def test_mixed_switch_back_to_fp64(c2):
    """The FP64 plan switched to mixed and back gives the FP64 result again."""
    am, win, cat, nbar, p, a64, amx = c2
    q = S.Plan(am)
    f1 = q.cat2amln(cat[:, :1000], nbar, win)
    q.set_precision("mixed")
    q.cat2amln(cat[:, :1000], nbar, win)
    q.set_precision("fp64")
    assert np.array_equal(f1, q.cat2amln(cat[:, :1000], nbar, win))


def test_mixed_c1_masked_dense_window():
    """A dense window (hemisphere mask times a radial profile, as win_rhat_ln's
    (npix, lmax+1, nmax) array) on a northern-hemisphere catalogue: the mixed path keeps the
    remainder of each spline record's knot value (the window's pattern would otherwise
    carry the records' FP32 rounding into the low-l modes), against the oracle."""
    am = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    wm = S.ConfigurationSpaceModes(500.0, 1000.0, 200, 32)
    th, _ = O.pix2ang_ring(32, np.arange(12 * 32 ** 2))
    mask = (th < np.pi / 2).astype(np.float64)
    phi = np.exp(-(wm.r - 750.0) ** 2 / (2 * 150.0 ** 2))
    dense = S.win_rhat_ln(phi[:, None] * mask[None, :], wm, am, device=0)
    dense = np.nan_to_num(dense, nan=0.0)
    cat = uniform_shell_catalog(100_000, 500.0, 1000.0, seed=21, theta_max=np.pi / 2)
    nbar = cat.shape[1] / shell_volume(500.0, 1000.0) * 2
    ref = O.cat2amln(oracle_view(am), cat, nbar, 0, dense.reshape(-1, order="F"))
    a = S.Plan(am, precision="mixed").cat2amln(cat, nbar, dense)
    e = block_errors(a, ref, am)
    say("C1 mixed, masked dense window, worst blocks", worst(e))
    assert max(e.values()) <= BLOCK_BAR


def test_mixed_delta_maps_uniform_knots_against_oracle(monkeypatch):
    """The mixed scatter finds each galaxy's spline segment and t on the plan's uniform knots
    itself (no per-group search arrays).  Its FP32 delta maps against the oracle's FP64 maps
    and against the same path with the per-group segment search (SFB_SEG_SEARCH=1), with
    radii on every knot of two maps, exactly at rmin and rmax, outside the knots on both
    sides and at +-inf (the boundary values), plus given weights: per map,
    max |delta - ref| <= 2e-6 max |ref|.  The kernel with its record-table stride as a
    constant equals the runtime-stride one (SFB_SCATTER_STRIDE=0) bit for bit, maps and
    a_nlm.  A NaN radius still gives the non-finite error."""
    am = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    win, _ = separable_unit_window(S, am)
    rng = np.random.default_rng(77)
    qa, qb = am.nmaps // 3, am.nmaps - 1
    radii = [am.knot_x[am.knot_off[q]:am.knot_off[q + 1]] for q in (qa, qb)]
    edge = np.concatenate(radii + [np.array([500.0, 1000.0, 499.999, 1000.001, 100.0, 3000.0, np.inf, -np.inf])])
    edge = np.repeat(edge, 3)
    parts = [np.stack([edge, np.arccos(rng.uniform(-1, 1, edge.size)), rng.uniform(0, 2 * np.pi, edge.size)]),
             uniform_shell_catalog(60_000, 500.0, 1000.0, seed=78)]
    cat = np.concatenate(parts, axis=1)
    cat = cat[:, rng.permutation(cat.shape[1])]
    nbar = cat.shape[1] / shell_volume(500.0, 1000.0)
    om = oracle_view(am)
    onfly = S.Plan(am, precision="mixed")
    monkeypatch.setenv("SFB_SEG_SEARCH", "1")
    search = S.Plan(am, precision="mixed")
    monkeypatch.delenv("SFB_SEG_SEARCH")
    monkeypatch.setenv("SFB_SCATTER_STRIDE", "0")
    rstride = S.Plan(am, precision="mixed")
    monkeypatch.delenv("SFB_SCATTER_STRIDE")
    dense = win.dense()
    for w in (None, rng.uniform(0.5, 1.5, cat.shape[1])):
        for q in (0, qa, qb):
            ref = O.delta_map(om, cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"), w, q)
            g = onfly.delta_map(cat, nbar, win, w, q)
            s = search.delta_map(cat, nbar, win, w, q)
            scale = np.max(np.abs(ref))
            e_ref, e_s = np.max(np.abs(g - ref)) / scale, np.max(np.abs(g - s)) / scale
            say(f"map {q} {'weights' if w is not None else 'unit'}: vs oracle {e_ref:.2e}, vs search {e_s:.2e}")
            assert e_ref <= 2e-6 and e_s <= 2e-6
            assert np.array_equal(g, rstride.delta_map(cat, nbar, win, w, q))
            # a dense window: the split-record kernel, same segments and t
            ref_d = O.delta_map(om, cat, nbar, 0, dense.reshape(-1, order="F"), None, w, q)
            e_d = np.max(np.abs(onfly.delta_map(cat, nbar, dense, w, q) - ref_d)) / np.max(np.abs(ref_d))
            say(f"map {q} dense window: vs oracle {e_d:.2e}")
            assert e_d <= 2e-6
        assert np.array_equal(onfly.cat2amln(cat, nbar, win, w), rstride.cat2amln(cat, nbar, win, w))
    bad = cat.copy()
    bad[0, 11] = np.nan
    with pytest.raises(AssertionError):
        onfly.cat2amln(bad, nbar, win)


def test_mixed_input_order():
    """The mixed path sorts galaxies by pixel only (input order inside a pixel: its FP64
    pixel sums need no r order), so a permuted catalogue changes its result only where an
    FP64 sum's last bit moves an FP32 delta value: per block within 1e-7 of the block RMS,
    and the same input twice gives bit-identical a_nlm."""
    am = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    win, _ = separable_unit_window(S, am)
    cat = uniform_shell_catalog(100_000, 500.0, 1000.0, seed=91)
    nbar = cat.shape[1] / shell_volume(500.0, 1000.0)
    w = np.random.default_rng(92).uniform(0.5, 1.5, cat.shape[1])
    plan = S.Plan(am, precision="mixed")
    a = plan.cat2amln(cat, nbar, win, w)
    assert np.array_equal(a, plan.cat2amln(cat, nbar, win, w))
    perm = np.random.default_rng(93).permutation(cat.shape[1])
    e = block_errors(plan.cat2amln(cat[:, perm], nbar, win, w[perm]), a, am)
    say("mixed, permuted input, worst blocks", worst(e))
    assert max(e.values()) <= 1e-7
