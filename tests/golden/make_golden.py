"""Writes tests/golden/c1_anlm.npz: the oracle's cat2amln at configuration 1.

Configuration 1: 1e5 galaxies uniform in volume in 500 <= r <= 1000 Mpc/h
(helpers.uniform_shell_catalog, Philox seed 0), unit weights, nbar = N/V,
AnlmModes(kmax=0.05, rmin=500, rmax=1000; nside=32) from the scipy restatement
(oracle.OracleModes, independent of the product's native setup), full-sky
separable window mask = 1, phi(r) = 1 on nr = 1000 shells, map2alm niter = 3.

The reference holds no golden a_nlm vectors (its weighted-catalogue test only
prints), so this fixture pins the oracle against regressions and gives the GPU
tests a target that does not depend on running the oracle.

    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests")]

from helpers import shell_volume, uniform_shell_catalog  # noqa: E402
from oracle import oracle as O  # noqa: E402


def window_w_ln(om, nr=1000):
    """W_nl = dr sum_r r^2 g_nl(r) (separable full-sky window, src/windows.jl:260-270)."""
    dr = (om.rmax - om.rmin) / nr
    r = np.linspace(om.rmin + dr / 2, om.rmax - dr / 2, nr)
    w = np.full((om.lmax + 1, om.nmax), np.nan)
    for n in range(1, om.nmax + 1):
        for l in range(int(om.lmax_n[n - 1]) + 1):
            w[l, n - 1] = dr * np.sum(r ** 2 * om.gnl(n, l, r))
    return w


def main():
    om = O.OracleModes(kmax=0.05, rmin=500.0, rmax=1000.0, nside=32)
    cat = uniform_shell_catalog(100_000, 500.0, 1000.0, seed=0)
    nbar = cat.shape[1] / shell_volume(500.0, 1000.0)
    w_ln = window_w_ln(om)
    mask = np.ones(12 * om.nside ** 2)
    anlm = O.cat2amln(om, cat, nbar, 1, mask, w_ln.reshape(-1, order="F"))
    np.savez_compressed(os.path.join(HERE, "c1_anlm.npz"), anlm=anlm, w_ln=w_ln, nbar=nbar,
                        lmax_n=om.lmax_n, knl=om.knl,
                        catalog_sha256=hashlib.sha256(np.ascontiguousarray(cat).tobytes()).hexdigest())
    print("wrote", anlm.size, "coefficients")


if __name__ == "__main__":
    main()
