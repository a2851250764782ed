"""Window inputs of cat2amln -- host mirror of the reference's producers.

Reference: src/windows.jl:80-134 (ConfigurationSpaceModes, window_r, get_rbounds),
:219-228 (integrate_window), :244-270 (win_rhat_ln, dense and separable),
src/SeparableArrays.jl:53-122 (SeparableArray with named fields).
"""
from __future__ import annotations

import numpy as np

from .modes import AnlmModes


class SeparableArray:
    """Lazy outer product A[i..., j...] = arr1[i...] * arr2[j...] (SeparableArrays.jl:53-111).

    The two factors are also reachable under the names given at construction,
    e.g. ``win.mask`` and ``win.w_ln`` for the output of ``win_rhat_ln``."""

    def __init__(self, arr1, arr2, name1="arr1", name2="arr2"):
        self.arr1 = np.asarray(arr1)
        self.arr2 = np.asarray(arr2)
        self.name1, self.name2 = name1, name2

    def __getattr__(self, name):
        d = self.__dict__
        if name == d.get("name1"):
            return d["arr1"]
        if name == d.get("name2"):
            return d["arr2"]
        raise AttributeError(name)

    @property
    def shape(self):
        return self.arr1.shape + self.arr2.shape

    def __getitem__(self, idx):
        if idx == (slice(None),) * len(self.shape) or idx is Ellipsis:
            return np.multiply.outer(self.arr1, self.arr2)
        n1 = self.arr1.ndim
        return self.arr1[idx[:n1]] * self.arr2[idx[n1:]]

    def dense(self):
        return np.multiply.outer(self.arr1, self.arr2)


class ConfigurationSpaceModes:
    """Voxelisation: nr radial shells of width dr between rmin and rmax, HEALPix nside
    (src/windows.jl:98-114)."""

    def __init__(self, rmin, rmax, nr, nside):
        self.rmin, self.rmax = float(rmin), float(rmax)
        self.nr = int(nr)
        self.nside = int(nside)
        self.npix = 12 * self.nside ** 2
        self.dr = (self.rmax - self.rmin) / self.nr
        self.r = np.linspace(self.rmin + self.dr / 2, self.rmax - self.dr / 2, self.nr)

    @classmethod
    def from_amodes(cls, amodes: AnlmModes, nr):
        return cls(amodes.rmin, amodes.rmax, nr, amodes.nside)

    Δr = property(lambda self: self.dr)


def window_r(wmodes):
    return wmodes.r, wmodes.dr


def get_rbounds(wmodes):
    return np.linspace(wmodes.rmin, wmodes.rmax, wmodes.nr + 1)


def integrate_window(win, wmodes):
    """V_eff = dr * dOmega * sum_r r^2 sum_p win[r,p] (src/windows.jl:219-228).
    A SeparableArray (phi(r), mask(p)) is integrated without densifying."""
    r, dr = window_r(wmodes)
    if isinstance(win, SeparableArray):
        npix = win.arr2.size
        radial = win.arr1 * np.sum(win.arr2)
    else:
        win = np.asarray(win)
        npix = win.shape[1]
        radial = win.sum(axis=1)
    domega = 4 * np.pi / npix
    return dr * domega * np.sum(radial * r ** 2)


def win_rhat_ln(win, wmodes, amodes: AnlmModes, *, device=None):
    """W[p, l, n-1] = dr sum_r r^2 g_nl(r) win[r, p] (src/windows.jl:244-257), NaN where
    l > lmax_n[n].  A separable window ``SeparableArray(phi[r], mask[p])`` gives
    ``SeparableArray(mask, w_ln[l, n-1], name1="mask", name2="w_ln")`` (:260-270).

    device=None evaluates on the host (NumPy); device=k runs sfb_win_rhat_ln on GPU k
    (one MFMA GEMM per n for a dense window)."""
    if device is not None:
        from .cat2anlm import get_plan
        return get_plan(amodes, device=device).win_rhat_ln(win, wmodes)
    r, dr = window_r(wmodes)
    gnl = amodes.basisfunctions
    if isinstance(win, SeparableArray):
        phi = np.asarray(win.arr1, dtype=float)
        w_ln = np.full((amodes.lmax + 1, amodes.nmax), np.nan)
        for n in range(1, amodes.nmax + 1):
            for l in range(int(amodes.lmax_n[n - 1]) + 1):
                w_ln[l, n - 1] = dr * np.sum(r ** 2 * gnl(n, l, r) * phi)
        return SeparableArray(np.asarray(win.arr2, dtype=float), w_ln, name1="mask", name2="w_ln")
    win = np.asarray(win, dtype=float)
    out = np.full((win.shape[1], amodes.lmax + 1, amodes.nmax), np.nan, order="F")
    for n in range(1, amodes.nmax + 1):
        for l in range(int(amodes.lmax_n[n - 1]) + 1):
            out[:, l, n - 1] = win.T @ (r ** 2 * gnl(n, l, r))
    out *= dr
    return out
