"""AnlmModes and the radial basis g_nl(r) -- host mirror of the reference API.

Reference: src/modes.jl:92-232 (AnlmModes, getidx/getnlm/getlmsize/getnlmsize,
estimate_nside), src/GNLs.jl:83-110 (GNL call and derivative interface),
src/Splines.jl:20-69,367-377 (Spline1D).  The numerical setup (k_nl roots,
c_nl/d_nl, spline knots and second derivatives) runs natively in
libsfb_mi355x.so (sfb_modes_create_kmax / sfb_modes_create_nl).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib as L


def estimate_nside(lmax):
    """src/modes.jl:68."""
    return 2 ** max(2, math.ceil(math.log2((2 * lmax + 1) / 2)))


class Spline1D:
    """Cubic spline with knots x, values y and second derivatives ypp (Splines.jl:20-28)."""

    def __init__(self, x, y, ypp=None):
        self.x = L.f64(x)
        self.y = L.f64(y)
        if ypp is None:
            ypp = np.empty_like(self.x)
            L.check(L.lib().sfb_spline_build(self.x.size, L.ptr(self.x), L.ptr(self.y), L.ptr(ypp)))
        self.ypp = L.f64(ypp)

    def _eval(self, r, deriv):
        r = L.f64(np.atleast_1d(r))
        out = np.empty_like(r)
        L.check(L.lib().sfb_spline_eval(self.x.size, L.ptr(self.x), L.ptr(self.y), L.ptr(self.ypp), r.size,
                                        L.ptr(r), L.ptr(out), int(deriv)))
        return out

    def __call__(self, r):
        out = self._eval(r, False)
        return out[0] if np.isscalar(r) else out

    def derivative(self, r):
        out = self._eval(r, True)
        return out[0] if np.isscalar(r) else out


class GNL:
    """Radial basis g_nl(r) = c_nl j_l(k_nl r) + d_nl y_l(k_nl r), potential boundary
    condition, evaluated through its cached splines (GNLs.jl:105-110)."""

    def __init__(self, amodes):
        self._am = amodes
        self.knl, self.cnl, self.dnl = amodes.knl, amodes.cnl, amodes.dnl
        self.rmin, self.rmax = amodes.rmin, amodes.rmax

    def spline(self, n, l):
        q = self._am.map_index(n, l)
        o0, o1 = self._am.knot_off[q], self._am.knot_off[q + 1]
        return Spline1D(self._am.knot_x[o0:o1], self._am.knot_y[o0:o1], self._am.knot_ypp[o0:o1])

    @property
    def gnl(self):
        am = self

        class _Idx:
            def __getitem__(self, nl1):
                n, l1 = nl1
                return am.spline(n, l1 - 1)

        return _Idx()

    def __call__(self, n, l, r):
        return self.spline(n, l)(r)

    def derivative(self, n, l, r):
        return self.spline(n, l).derivative(r)


class AnlmModes:
    """AnlmModes(kmax, rmin, rmax; nside) -- src/modes.jl:122-158.

    Use ``AnlmModes.from_nl(nmax, lmax, rmin, rmax; nside)`` for src/modes.jl:161-175
    and ``AnlmModes.from_tables(...)`` to wrap spline tables built elsewhere.
    Modes are ordered n slowest, then l, then m (src/modes.jl:80-82); the
    (n,l) "maps" are ordered n-major, l-minor.
    """

    def __init__(self, kmax, rmin, rmax, nside=None, *, _handle=None):
        if _handle is None:
            h = ctypes.c_void_p()
            L.check(L.lib().sfb_modes_create_kmax(ctypes.byref(h), float(kmax), float(rmin), float(rmax),
                                                  int(nside or 0)))
            _handle = h
        self.rmin, self.rmax = float(rmin), float(rmax)
        self._load(_handle)

    @classmethod
    def from_nl(cls, nmax, lmax, rmin, rmax, nside=None):
        h = ctypes.c_void_p()
        L.check(L.lib().sfb_modes_create_nl(ctypes.byref(h), int(nmax), int(lmax), float(rmin), float(rmax),
                                            int(nside or 0)))
        return cls(None, rmin, rmax, _handle=h)

    @classmethod
    def from_tables(cls, rmin, rmax, nside, lmax_n, knot_off, knot_x, knot_y, knot_ypp, knl=None):
        self = cls.__new__(cls)
        self.rmin, self.rmax = float(rmin), float(rmax)
        self.lmax_n = L.i64(lmax_n)
        self.nmax = int(self.lmax_n.size)
        self.lmax = int(self.lmax_n.max())
        self.nside = int(nside)
        self.nmax_l = np.array([int(np.sum(self.lmax_n >= l)) for l in range(self.lmax + 1)], dtype=np.int64)
        self.knl = np.full((self.nmax, self.lmax + 1), np.nan) if knl is None else np.asarray(knl, dtype=float)
        self.cnl = np.full_like(self.knl, np.nan)
        self.dnl = np.full_like(self.knl, np.nan)
        self.kmax = float(np.nanmax(self.knl)) if np.isfinite(self.knl).any() else float("nan")
        self.knot_off = L.i64(knot_off)
        self.knot_x, self.knot_y, self.knot_ypp = L.f64(knot_x), L.f64(knot_y), L.f64(knot_ypp)
        self.basisfunctions = GNL(self)
        return self

    def _load(self, h):
        lib = L.lib()
        v = [ctypes.c_int64() for _ in range(6)]
        try:
            L.check(lib.sfb_modes_sizes(h, *[ctypes.byref(x) for x in v]))
            nmax, lmax, nside, nmaps, nlm, nknots = (x.value for x in v)
            self.nmax, self.lmax, self.nside = nmax, lmax, nside
            self.lmax_n = np.empty(nmax, dtype=np.int64)
            self.nmax_l = np.empty(lmax + 1, dtype=np.int64)
            shape = (nmax, lmax + 1)
            knl = np.empty(nmax * (lmax + 1))
            cnl = np.empty_like(knl)
            dnl = np.empty_like(knl)
            self.knot_off = np.empty(nmaps + 1, dtype=np.int64)
            self.knot_x = np.empty(nknots)
            self.knot_y = np.empty(nknots)
            self.knot_ypp = np.empty(nknots)
            L.check(lib.sfb_modes_tables(h, L.ptr(self.lmax_n), L.ptr(self.nmax_l), L.ptr(knl), L.ptr(cnl),
                                         L.ptr(dnl), L.ptr(self.knot_off), L.ptr(self.knot_x), L.ptr(self.knot_y),
                                         L.ptr(self.knot_ypp)))
            # column-major (Julia) -> numpy [n-1, l]
            self.knl = knl.reshape(shape, order="F")
            self.cnl = cnl.reshape(shape, order="F")
            self.dnl = dnl.reshape(shape, order="F")
        finally:
            lib.sfb_modes_destroy(h)
        self.kmax = float(np.nanmax(self.knl))
        self.basisfunctions = GNL(self)

    # ---- index helpers (src/modes.jl:178-232) ----------------------------
    @property
    def nmaps(self):
        return int(np.sum(self.lmax_n + 1))

    def map_index(self, n, l):
        """0-based index of the (n,l) pair in map order."""
        return int(np.sum(self.lmax_n[: n - 1] + 1)) + int(l)

    def map_nl(self):
        """(n, l) of every map, in map order."""
        return [(n, l) for n in range(1, self.nmax + 1) for l in range(int(self.lmax_n[n - 1]) + 1)]

    def spline(self, n, l):
        return self.basisfunctions.spline(n, l)

    def __repr__(self):
        return (f"AnlmModes(kmax={self.kmax}, rmin={self.rmin}, rmax={self.rmax}, nmax={self.nmax}, "
                f"lmax={self.lmax}, nside={self.nside}, num_modes={int(np.isfinite(self.knl).sum())})")


def getlmsize(lmax):
    return lmax * (lmax + 1) // 2 + lmax + 1


def getnlmsize(amodes, nmax=None):
    nmax = amodes.nmax if nmax is None else nmax
    return int(sum(getlmsize(int(l)) for l in amodes.lmax_n[:nmax]))


def getidx(amodes, n, l, m):
    """1-based index of (n,l,m) (src/modes.jl:223-232)."""
    assert n >= 1 and l >= 0 and m >= 0
    return 1 + getnlmsize(amodes, n - 1) + getlmsize(l - 1) + m


def getnlm(amodes, idx):
    """Inverse of getidx (src/modes.jl:203-220)."""
    n = 1
    nmodes = getlmsize(int(amodes.lmax_n[n - 1]))
    while idx > nmodes:
        idx -= nmodes
        n += 1
        nmodes = getlmsize(int(amodes.lmax_n[n - 1]))
    l = 0
    nmodes = l + 1
    while idx > nmodes:
        idx -= nmodes
        l += 1
        nmodes = l + 1
    return n, l, idx - 1


def isvalidnlm(amodes, n, l, m):
    if n > amodes.nmax:
        return False
    if l > amodes.lmax_n[n - 1]:
        return False
    return abs(m) <= l
