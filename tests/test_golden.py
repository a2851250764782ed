"""Configuration-1 golden fixture (tests/golden/c1_anlm.npz, written by the oracle
through tests/golden/make_golden.py): the oracle still reproduces it, and the HIP
path matches it to 1e-10 of each (n,l) block's RMS."""
import hashlib
import os

import numpy as np
import pytest

import sfb_amd as S
from helpers import rel_errors, shell_volume, uniform_shell_catalog
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_anlm.npz")


@pytest.fixture(scope="module")
def gold():
    g = np.load(GOLD, allow_pickle=False)
    cat = uniform_shell_catalog(100_000, 500.0, 1000.0, seed=0)
    assert hashlib.sha256(np.ascontiguousarray(cat).tobytes()).hexdigest() == str(g["catalog_sha256"])
    return g, cat


def test_oracle_reproduces_golden(gold):
    g, cat = gold
    om = O.OracleModes(kmax=0.05, rmin=500.0, rmax=1000.0, nside=32)
    assert np.array_equal(om.lmax_n, g["lmax_n"])
    anlm = O.cat2amln(om, cat, float(g["nbar"]), 1, np.ones(12 * 32 * 32), g["w_ln"].reshape(-1, order="F"))
    assert np.max(np.abs(anlm - g["anlm"])) <= 1e-13 * np.max(np.abs(g["anlm"]))


@pytest.mark.gpu
def test_gpu_matches_golden(gold):
    g, cat = gold
    am = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    win = S.SeparableArray(np.ones(12 * 32 * 32), g["w_ln"], name1="mask", name2="w_ln")
    assert cat.shape[1] / shell_volume(500.0, 1000.0) == float(g["nbar"])
    got = S.cat2amln(cat, am, float(g["nbar"]), win)
    e = rel_errors(got, g["anlm"], am)
    print(e)
    assert e["max_blk"] < 1e-10
