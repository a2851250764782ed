"""Mixed-precision path (SFB_PREC_MIXED): FP32 spline values, delta maps and
phi-transform, FP32 MFMA for every Legendre pass; FP64 setup, segment search,
lambda tables and GEMM epilogues.  Its error against the FP64 oracle is measured
per mode and relative to the largest coefficient; the bars below are the FP32 level
the path is built for (per-mode errors of small coefficients are larger)."""
import numpy as np
import pytest

import sfb_amd as S
from helpers import map_entries, rel_errors, separable_unit_window, shell_volume, uniform_shell_catalog
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def oracle_view(am):
    om = O.OracleModes.__new__(O.OracleModes)
    om.nside, om.lmax, om.nmax, om.lmax_n = am.nside, am.lmax, am.nmax, am.lmax_n
    om.knot_off, om.knot_x, om.knot_y, om.knot_ypp = am.knot_off, am.knot_x, am.knot_y, am.knot_ypp
    return om


def test_mixed_c1_against_oracle():
    am = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    win, _ = separable_unit_window(S, am)
    cat = uniform_shell_catalog(100_000, 500.0, 1000.0, seed=0)
    w = np.random.default_rng(3).uniform(0.5, 1.5, cat.shape[1])
    nbar = cat.shape[1] / shell_volume(500.0, 1000.0)
    ref = O.cat2amln(oracle_view(am), cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"), w=w)
    a = S.Plan(am, precision="mixed").cat2amln(cat, nbar, win, w)
    e = rel_errors(a, ref, am)
    say("C1 mixed errors", e)
    assert np.all(np.isfinite(a))
    assert e["max_abs_over_max"] < 5e-6
    assert e["rms_blk"] < 1e-5


def test_mixed_c2_sample_and_dense_window():
    am = S.AnlmModes(0.1, 500.0, 1500.0, nside=128)
    win, _ = separable_unit_window(S, am)
    cat = uniform_shell_catalog(1_000_000, 500.0, 1500.0, seed=0)
    nbar = cat.shape[1] / shell_volume(500.0, 1500.0)
    a = S.Plan(am, precision="mixed").cat2amln(cat, nbar, win)
    ref = O.cat2amln(oracle_view(am), cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"), q_range=(0, 16))
    sel = map_entries(am, 0, 16)
    d = np.abs(a[sel] - ref[sel])
    say(f"C2 mixed maps 0..15: max |d|/max|ref| = {d.max() / np.abs(ref[sel]).max():.2e}")
    assert d.max() < 5e-6 * np.abs(ref[sel]).max()
    # the FP64 plan switched to mixed and back gives the FP64 result again
    p = S.Plan(am)
    f1 = p.cat2amln(cat[:, :1000], nbar, win)
    p.set_precision("mixed")
    p.cat2amln(cat[:, :1000], nbar, win)
    p.set_precision("fp64")
    assert np.array_equal(f1, p.cat2amln(cat[:, :1000], nbar, win))
