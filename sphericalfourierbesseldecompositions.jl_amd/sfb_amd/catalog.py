"""Catalogue preparation on the device: xyz2rtp
(src/SphericalFourierBesselDecompositions.jl:517-532), apply_window (src/windows.jl:169-216)
and winweights2galweights (src/cat2anlm.jl:218-238)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .windows import get_rbounds


def xyz2rtp(xyz, *, device=0):
    """3 x N Cartesian positions -> 3 x N (r, theta, phi)."""
    xyz = np.asarray(xyz, dtype=np.float64)
    if xyz.ndim != 2 or xyz.shape[0] != 3:
        raise ValueError("xyz must be 3 x N")
    src = L.f64(xyz.T)
    out = np.empty_like(src)
    L.check(L.lib().sfb_xyz2rtp(int(device), src.shape[0], L.ptr(src), L.ptr(out)))
    return out.T.copy()


def winweights2galweights(weights, wmodes, rθϕ, *, rbounds=None, device=0):
    """weights[bin, pix] of every galaxy; weights is nr x npix.  rbounds defaults to
    get_rbounds(wmodes); a Julia caller passes the reference's own bin edges."""
    w = np.asfortranarray(np.asarray(weights, dtype=np.float64))
    if w.shape != (wmodes.nr, wmodes.npix):
        raise ValueError(f"weights must be (nr, npix) = {(wmodes.nr, wmodes.npix)}")
    rb = L.f64(get_rbounds(wmodes) if rbounds is None else rbounds)
    rtp = L.f64(np.asarray(rθϕ, dtype=np.float64).T)
    out = np.empty(rtp.shape[0])
    L.check(L.lib().sfb_winweights2galweights(int(device), wmodes.nside, wmodes.nr, L.ptr(rb),
                                              ctypes.c_void_p(w.ctypes.data), rtp.shape[0], L.ptr(rtp), L.ptr(out)))
    return out


def apply_window(rθϕ, win, wmodes, *, u=None, rng=None, device=0):
    """Keep each galaxy with probability win[r-bin, pixel] / max(win) (and rmin <= r <= rmax).
    u are the uniform draws (one per galaxy); by default drawn from ``rng``."""
    w = np.asfortranarray(np.asarray(win, dtype=np.float64))
    if w.shape != (wmodes.nr, wmodes.npix):
        raise ValueError(f"window must be (nr, npix) = {(wmodes.nr, wmodes.npix)}")
    rtp = L.f64(np.asarray(rθϕ, dtype=np.float64).T)
    N = rtp.shape[0]
    if u is None:
        u = (rng if rng is not None else np.random.default_rng()).random(N)
    u = L.f64(u)
    if u.size != N:
        raise ValueError("u must have one draw per galaxy")
    out = np.empty_like(rtp)
    n_out = ctypes.c_int64()
    L.check(L.lib().sfb_apply_window(int(device), wmodes.nside, wmodes.nr, wmodes.rmin, wmodes.rmax, wmodes.dr,
                                     ctypes.c_void_p(w.ctypes.data), N, L.ptr(rtp), L.ptr(u), L.ptr(out),
                                     ctypes.byref(n_out)))
    return out[:n_out.value].T.copy()
