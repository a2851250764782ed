"""Batched mocks (sfb_cat2amln_batch): B catalogues of one survey geometry through
one plan, enqueued back to back with no host synchronisation between them.  Every
mock of the batch must equal its own cat2amln call bit for bit (same kernels, same
summation order), including an empty catalogue and a crowded one whose sliced layout
outgrows the slot bound (it is redone alone with exact sizes), and one mock is checked
against the oracle.  An error in one mock names it.  At the nside-128 geometry
(kmax 0.1), eight Poisson mocks with means drawn in [1e5, 1e6], both precisions."""
import numpy as np
import pytest

import sfb_amd as S
from helpers import separable_unit_window, shell_volume, uniform_shell_catalog
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def setup():
    am = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    win, _ = separable_unit_window(S, am, nr=200)
    rng = np.random.default_rng(40)
    sizes = [int(rng.poisson(mu)) for mu in rng.uniform(1e4, 1e5, 4)] + [0]
    cats = [uniform_shell_catalog(n, 500.0, 1000.0, seed=100 + b) if n else np.zeros((3, 0))
            for b, n in enumerate(sizes)]
    # crowded: 30000 galaxies in one pixel (64 slots per galaxy row: far past the slot bound)
    crowd = np.stack([rng.uniform(500.0, 1000.0, 30000), np.full(30000, 1.0), np.full(30000, 2.0)])
    cats.append(crowd)
    weights = [rng.uniform(0.5, 1.5, c.shape[1]) for c in cats]
    nbars = [max(c.shape[1], 1) / shell_volume(500.0, 1000.0) for c in cats]
    return am, win, cats, weights, nbars


def test_batch_equals_single_calls(setup):
    am, win, cats, weights, nbars = setup
    say("batch of", [c.shape[1] for c in cats])
    plan = S.get_plan(am)
    batch = plan.cat2amln_batch(cats, nbars, win, weights)
    assert batch.shape == (len(cats), S.getnlmsize(am))
    for b, c in enumerate(cats):
        one = plan.cat2amln(c, nbars[b], win, weights[b] if c.shape[1] else [])
        assert np.array_equal(batch[b], one), b
    # unit weights
    batch1 = plan.cat2amln_batch(cats[:2], nbars[:2], win)
    assert np.array_equal(batch1[1], plan.cat2amln(cats[1], nbars[1], win))


def test_batch_mock_against_oracle(setup):
    am, win, cats, weights, nbars = setup
    say("batch mock vs oracle")
    batch = S.get_plan(am).cat2amln_batch(cats[:2], nbars[:2], win, weights[:2])
    om = O.OracleModes.__new__(O.OracleModes)
    om.nside, om.lmax, om.nmax, om.lmax_n = am.nside, am.lmax, am.nmax, am.lmax_n
    om.knot_off, om.knot_x, om.knot_y, om.knot_ypp = am.knot_off, am.knot_x, am.knot_y, am.knot_ypp
    ref = O.cat2amln(om, cats[1], nbars[1], 1, win.mask, win.w_ln.reshape(-1, order="F"), w=weights[1])
    assert np.max(np.abs(batch[1] - ref)) < 1e-10 * np.max(np.abs(ref))


def test_batch_error_names_catalogue(setup):
    am, win, cats, weights, nbars = setup
    say("batch error path")
    bad = [w.copy() for w in weights[:3]]
    bad[2][5] = np.nan
    with pytest.raises(AssertionError) as e:
        S.get_plan(am).cat2amln_batch(cats[:3], nbars[:3], win, bad)
    assert "catalogue 2" in str(e.value)


def test_batch_c2_geometry_poisson_mocks():
    am = S.AnlmModes(0.1, 500.0, 1500.0, nside=128)
    win, _ = separable_unit_window(S, am)
    rng = np.random.default_rng(41)
    sizes = [int(rng.poisson(mu)) for mu in rng.uniform(1e5, 1e6, 8)]
    cats = [uniform_shell_catalog(n, 500.0, 1500.0, seed=200 + b) for b, n in enumerate(sizes)]
    nbars = [c.shape[1] / shell_volume(500.0, 1500.0) for c in cats]
    say("C2 batch of", sizes)
    for prec in ("fp64", "mixed"):
        plan = S.get_plan(am, precision=prec)
        batch = plan.cat2amln_batch(cats, nbars, win)
        for b, c in enumerate(cats):
            assert np.array_equal(batch[b], plan.cat2amln(c, nbars[b], win)), (prec, b)
        if prec == "fp64":
            om = O.OracleModes.__new__(O.OracleModes)
            om.nside, om.lmax, om.nmax, om.lmax_n = am.nside, am.lmax, am.nmax, am.lmax_n
            om.knot_off, om.knot_x, om.knot_y, om.knot_ypp = am.knot_off, am.knot_x, am.knot_y, am.knot_ypp
            ref = O.cat2amln(om, cats[3], nbars[3], 1, win.mask, win.w_ln.reshape(-1, order="F"), q_range=(0, 12))
            n12 = int(np.sum(np.arange(1, 13)))  # maps 0..11 are n = 1, l = 0..11
            d = np.abs(batch[3][:n12] - ref[:n12])
            assert d.max() < 1e-10 * np.abs(ref[:n12]).max()


def test_batch_graph_replays_match_eager(setup):
    """The batch captured as one HIP graph: its first launch and two replays (with other
    eager work on the plan in between) return the eager batch bit for bit."""
    am, win, cats, weights, nbars = setup
    plan = S.Plan(am)
    ws = [w if c.shape[1] else [] for c, w in zip(cats, weights)]
    eager = plan.cat2amln_batch(cats, nbars, win, ws)
    plan.set_batch_graph(True)
    small = cats[1][:, :500]
    for rep in range(3):
        say("graph batch, launch", rep)
        got = plan.cat2amln_batch(cats, nbars, win, ws)
        assert np.array_equal(got, eager), rep
        plan.cat2amln(small, nbars[1], win)  # eager work between launches
    assert plan.set_batch_graph(None) > 0.0
    plan.set_batch_graph(False)
    assert np.array_equal(plan.cat2amln_batch(cats, nbars, win, ws), eager)
