"""Split-BF16 Jacobi GEMMs (SFB_PREC_BF16X3: hi + mid + lo bf16 terms, six products;
SFB_PREC_BF16X2: hi + lo, three products; FP32 accumulation) against the oracle at nside 32
and against the FP64 GPU path over every (n, l) block at nside 128, to the bar of the mixed
path (max_m |a - ref| <= 1e-6 of the block's RMS).  The Jacobi steps only correct the first
analysis (FP64 MFMA), so even the hi + lo form (~2^-16 per product) stays within the bar."""
import numpy as np
import pytest

import sfb_amd as S
from helpers import separable_unit_window, shell_volume, uniform_shell_catalog
from oracle import oracle as O
from test_gpu_mixed import block_errors, oracle_view, worst

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("prec", ["bf16x3", "bf16x2"])
def test_bf16_c1_against_oracle(prec):
    am = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    win, _ = separable_unit_window(S, am)
    cat = uniform_shell_catalog(100_000, 500.0, 1000.0, seed=0)
    w = np.random.default_rng(3).uniform(0.5, 1.5, cat.shape[1])
    nbar = cat.shape[1] / shell_volume(500.0, 1000.0)
    ref = O.cat2amln(oracle_view(am), cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"), w=w)
    a = S.Plan(am, precision=prec).cat2amln(cat, nbar, win, w)
    e = block_errors(a, ref, am)
    say(f"C1 {prec} worst blocks", worst(e))
    assert max(e.values()) <= 1e-6


def test_bf16_c2_every_block():
    am = S.AnlmModes(0.1, 500.0, 1500.0, nside=128)
    win, _ = separable_unit_window(S, am)
    cat = uniform_shell_catalog(1_000_000, 500.0, 1500.0, seed=0)
    nbar = cat.shape[1] / shell_volume(500.0, 1500.0)
    p = S.Plan(am, precision="fp64")
    a64 = p.cat2amln(cat, nbar, win)
    res = {}
    for prec in ("mixed", "bf16x3", "bf16x2"):
        p.set_precision(prec)
        e = block_errors(p.cat2amln(cat, nbar, win), a64, am)
        res[prec] = max(e.values())
        say(f"C2 {prec} vs FP64 GPU over {len(e)} blocks: worst {worst(e)}")
    assert res["bf16x3"] <= 1e-6
    assert res["bf16x2"] <= 1e-6
