"""Shared workload builders for the tests (synthetic catalogues of the benchmark
configurations) and error metrics."""
from __future__ import annotations

import numpy as np


def uniform_shell_catalog(N, rmin, rmax, seed, theta_max=np.pi):
    """N galaxies uniform in volume in rmin <= r <= rmax: r^3 uniform in
    [rmin^3, rmax^3], cos(theta) uniform in [cos(theta_max), 1], phi uniform in [0, 2pi).
    Returns the reference's 3 x N (r, theta, phi) layout."""
    rng = np.random.Generator(np.random.Philox(seed))
    u = rng.random((3, N))
    r = np.cbrt(u[0] * (rmax ** 3 - rmin ** 3) + rmin ** 3)
    zmin = np.cos(theta_max)
    costh = zmin + (1.0 - zmin) * (1.0 - u[1])
    th = np.arccos(np.clip(costh, -1.0, 1.0))
    ph = 2 * np.pi * u[2]
    ph[ph >= 2 * np.pi] = 0.0
    return np.stack([r, th, ph])


def shell_volume(rmin, rmax):
    return 4 * np.pi / 3 * (rmax ** 3 - rmin ** 3)


def separable_unit_window(sfb, amodes, nr=1000):
    """Full-sky window mask = 1, phi(r) = 1 on nr shells -> win_rhat_ln as a SeparableArray."""
    wmodes = sfb.ConfigurationSpaceModes(amodes.rmin, amodes.rmax, nr, amodes.nside)
    npix = 12 * amodes.nside ** 2
    win = sfb.SeparableArray(np.ones(nr), np.ones(npix), name1="phi", name2="mask")
    return sfb.win_rhat_ln(win, wmodes, amodes), wmodes


def rel_errors(a, ref, amodes):
    """Max/RMS per-mode relative error and per-(n,l)-block-normalised error."""
    a = np.asarray(a)
    ref = np.asarray(ref)
    d = np.abs(a - ref)
    per_mode = d / np.maximum(np.abs(ref), 1e-300)
    blk = np.empty_like(d)
    i = 0
    for n in range(1, amodes.nmax + 1):
        for l in range(int(amodes.lmax_n[n - 1]) + 1):
            s = slice(i, i + l + 1)
            rms = np.sqrt(np.mean(np.abs(ref[s]) ** 2))
            blk[s] = d[s] / max(rms, 1e-300)
            i += l + 1
    return {
        "max_rel": float(per_mode.max()), "rms_rel": float(np.sqrt(np.mean(per_mode ** 2))),
        "max_blk": float(blk.max()), "rms_blk": float(np.sqrt(np.mean(blk ** 2))),
        "max_abs_over_max": float(d.max() / np.abs(ref).max()),
    }


def gaussian_shell_catalog(N, seed, rmin=500.0, rmax=1500.0, r0=1000.0, sigma=300.0):
    """N galaxies in the northern hemisphere (cos(theta) in (0, 1]) with radial density
    proportional to r^2 exp(-(r - r0)^2 / (2 sigma^2)) on [rmin, rmax] (rejection from
    uniform-in-volume), weights uniform in [0.5, 1.5] rounded to FP32 values."""
    rng = np.random.Generator(np.random.Philox(seed))
    r = np.empty(0)
    while r.size < N:
        u = rng.random(2 * (N - r.size) + 1000)
        rr = np.cbrt(u * (rmax ** 3 - rmin ** 3) + rmin ** 3)
        keep = rng.random(rr.size) < np.exp(-(rr - r0) ** 2 / (2 * sigma ** 2))
        r = np.concatenate([r, rr[keep]])
    r = r[:N]
    costh = 1.0 - rng.random(N)
    th = np.arccos(costh)
    ph = 2 * np.pi * rng.random(N)
    w = rng.uniform(0.5, 1.5, N).astype(np.float32).astype(np.float64)
    return np.stack([r, th, ph]), w


def map_entries(amodes, q0, q1):
    """Boolean selector of the a_nlm entries of maps q0..q1-1 (map order)."""
    sel = np.zeros(int(np.sum((amodes.lmax_n + 1) * (amodes.lmax_n + 2) // 2)), dtype=bool)
    i = 0
    q = 0
    for n in range(1, amodes.nmax + 1):
        for l in range(int(amodes.lmax_n[n - 1]) + 1):
            if q0 <= q < q1:
                sel[i:i + l + 1] = True
            i += l + 1
            q += 1
    return sel
