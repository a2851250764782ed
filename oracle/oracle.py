"""CPU oracle for the cat2amln path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module, and only as the checker.  The product library never loads it.

Two halves:

* ``liboracle.so`` (``sfb_oracle.c``): ang2pix/pix2ang, Spline1D{3}, the ring
  spherical-harmonic transform with niter refinement, and the cat2amln loop, in C.
* This file: a numpy/scipy restatement of ``AnlmModes`` for the potential
  boundary condition -- the k_nl roots, c_nl/d_nl and the spline knots -- so the
  product's native setup can be checked against an independent implementation.

Reference files followed (hsgg/SphericalFourierBesselDecompositions.jl 0.5.19):
``src/SphericalBesselGNLs.jl:72-166`` (double secant and zero stepping),
``:197-228`` (zero scan, potential zero function), ``:309-336`` (k_nl table),
``:396-421`` (c_nl, d_nl), ``:558-579`` (spline cache), ``src/modes.jl:68,122-158``
(AnlmModes), ``:178-232`` (index helpers).  Third-party pieces restated from their
published definitions: SpecialFunctions.jl 2.5.1 ``sphericalbesselj/y`` =
sqrt(pi/2x) J/Y_{l+1/2}(x) (AMOS Bessel functions; scipy's ``jv``/``yv`` wrap
the same AMOS routines), and Julia 1.11 float ranges ``a:s:b``
(base/twiceprecision.jl: rational reconstruction, then exactly-rounded elements).
"""
from __future__ import annotations

import ctypes
import math
import os
from fractions import Fraction

import numpy as np
from scipy import special

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        i64, i32, dbl = ctypes.c_int64, ctypes.c_int32, ctypes.c_double
        P = ctypes.c_void_p
        L.orc_ang2pix_ring.restype = i64
        L.orc_ang2pix_ring.argtypes = [i64, dbl, dbl]
        L.orc_ang2pix_ring_many.argtypes = [i64, i64, P, P, P]
        L.orc_pix2ang_ring.argtypes = [i64, i64, P, P]
        L.orc_ring_info.argtypes = [i64, i64, P, P, P, P]
        L.orc_spline_build.restype = ctypes.c_int
        L.orc_spline_build.argtypes = [i64, P, P, P]
        L.orc_spline_eval.restype = dbl
        L.orc_spline_eval.argtypes = [i64, P, P, P, dbl]
        L.orc_spline_deriv.restype = dbl
        L.orc_spline_deriv.argtypes = [i64, P, P, P, dbl]
        L.orc_spline_eval_many.argtypes = [i64, P, P, P, i64, P, P]
        L.orc_lambda_table.argtypes = [i64, i64, P]
        L.orc_map2alm.restype = ctypes.c_int
        L.orc_map2alm.argtypes = [i64, i64, i32, P, P]
        L.orc_alm2map.restype = ctypes.c_int
        L.orc_alm2map.argtypes = [i64, i64, P, P]
        L.orc_delta_map.restype = ctypes.c_int
        L.orc_delta_map.argtypes = [i64, i64, i64, P, P, P, P, P, i64, P, P, dbl, i32, P, P, i64, P]
        L.orc_markstein_check.restype = i64
        L.orc_markstein_check.argtypes = [i64, P, P, i64, ctypes.c_uint64, i32, P, P]
        L.orc_cat2amln.restype = ctypes.c_int
        L.orc_cat2amln.argtypes = [i64, i64, i64, P, P, P, P, P, i64, P, P, dbl, i32, P, P, i32,
                                   i64, i64, i32, P]
        _lib = L
    return _lib


def markstein_check(x0, x1, nr=64, seed=0, variant=0):
    """Mismatches of the divide-free t against IEEE division over the segments (x0, x1)
    (variant 1: without the correction step, the check's control)."""
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    x1 = np.ascontiguousarray(x1, dtype=np.float64)
    b0, br = ctypes.c_double(), ctypes.c_double()
    n = lib().orc_markstein_check(x0.size, _p(x0), _p(x1), int(nr), int(seed), int(variant), ctypes.byref(b0),
                                 ctypes.byref(br))
    return int(n), (b0.value, br.value)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------
# HEALPix helpers
# --------------------------------------------------------------------------
def ang2pix_ring(nside, theta, phi):
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.empty(theta.shape, dtype=np.int64)
    lib().orc_ang2pix_ring_many(nside, theta.size, _p(theta), _p(phi), _p(out))
    return out


def pix2ang_ring(nside, pix):
    pix = np.atleast_1d(np.asarray(pix, dtype=np.int64))
    th = np.empty(pix.shape)
    ph = np.empty(pix.shape)
    t = ctypes.c_double()
    p = ctypes.c_double()
    for i, q in enumerate(pix):
        lib().orc_pix2ang_ring(nside, int(q), ctypes.byref(t), ctypes.byref(p))
        th[i], ph[i] = t.value, p.value
    return th, ph


def ring_info(nside, ring):
    st, n, z, phi0 = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
    lib().orc_ring_info(nside, ring, ctypes.byref(st), ctypes.byref(n), ctypes.byref(z), ctypes.byref(phi0))
    return st.value, n.value, z.value, phi0.value


# --------------------------------------------------------------------------
# Splines
# --------------------------------------------------------------------------
def spline_build(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    ypp = np.empty_like(x)
    rc = lib().orc_spline_build(x.size, _p(x), _p(y), _p(ypp))
    if rc != 0:
        raise ValueError("spline needs at least 4 knots")
    return ypp


def spline_eval(x, y, ypp, r):
    r = np.ascontiguousarray(np.atleast_1d(r), dtype=np.float64)
    out = np.empty_like(r)
    lib().orc_spline_eval_many(x.size, _p(x), _p(y), _p(ypp), r.size, _p(r), _p(out))
    return out


def spline_deriv(x, y, ypp, r):
    return np.array([lib().orc_spline_deriv(x.size, _p(x), _p(y), _p(ypp), float(v)) for v in np.atleast_1d(r)])


def lambda_table(nside, lmax):
    nalm = (lmax + 1) * (lmax + 2) // 2
    out = np.empty((4 * nside - 1, nalm))
    lib().orc_lambda_table(nside, lmax, _p(out))
    return out


def alm_index(lmax, l, m):
    """Healpix m-major index (LMcalcStructs.jl:13-17), 0-based."""
    return m * (2 * lmax + 1 - m) // 2 + l


def map2alm(nside, lmax, m, niter=3):
    m = np.ascontiguousarray(m, dtype=np.float64)
    nalm = (lmax + 1) * (lmax + 2) // 2
    out = np.empty(2 * nalm)
    rc = lib().orc_map2alm(nside, lmax, niter, _p(m), _p(out))
    if rc != 0:
        raise RuntimeError(f"orc_map2alm failed: {rc}")
    return out[0::2] + 1j * out[1::2]


def alm2map(nside, lmax, alm):
    a = np.empty(2 * alm.size)
    a[0::2] = alm.real
    a[1::2] = alm.imag
    out = np.empty(12 * nside * nside)
    lib().orc_alm2map(nside, lmax, _p(a), _p(out))
    return out


# --------------------------------------------------------------------------
# Julia float ranges (a:s:b), restated for the spline knots
# --------------------------------------------------------------------------
def _rat(x: float):
    y = x
    a = d = 1
    b = c = 0
    m = 16777216  # maxintfloat(Float32, Int)
    while abs(y) <= m:
        f = int(y)  # trunc
        y -= f
        a, c = f * a + c, a
        b, d = f * b + d, b
        if max(abs(a), abs(b)) > m:
            return c, d
        if float(a) / float(b) == x:
            break
        y = math.inf if y == 0.0 else 1.0 / y
    return a, b


def _isbetween(a, b, c):
    return a <= b <= c or c <= b <= a


def julia_float_range(start: float, step: float, stop: float) -> np.ndarray:
    """Elements of Julia's ``start:step:stop`` for Float64 arguments."""
    sn, sd = _rat(step)
    if sd != 0 and float(sn) / float(sd) == step:
        an, ad = _rat(start)
        bn, bd = _rat(stop)
        if ad != 0 and bd != 0 and float(an) / float(ad) == start and float(bn) / float(bd) == stop:
            den = ad * sd // math.gcd(ad, sd)
            mm = 2.0 ** 53
            if den != 0 and abs(start * den) <= mm and abs(step * den) <= mm and den % ad == 0 and den % sd == 0:
                a_n = int(round(start * den))
                s_n = int(round(step * den))
                num = den * bn - bd * a_n + s_n * bd
                dd = s_n * bd
                q = abs(num) // abs(dd)
                if (num < 0) != (dd < 0):
                    q = -q
                length = max(0, q)
                if _isbetween(start, start + (length - 1) * step, stop + step / 2) and \
                        not _isbetween(start, start + length * step, stop):
                    return np.array([float(Fraction(a_n + i * s_n, den)) for i in range(length)])
    lf = (stop - start) / step
    if lf < 0:
        length = 0
    elif lf == 0:
        length = 1
    else:
        length = int(round(lf)) + 1
        stop2 = start + (length - 1) * step
        length -= int(start < stop < stop2) + int(start > stop > stop2)
    fs, fst = Fraction(start), Fraction(step)
    return np.array([float(fs + i * fst) for i in range(length)])


# --------------------------------------------------------------------------
# AnlmModes, potential boundary condition
# --------------------------------------------------------------------------
def sphericalbesselj(l, x):
    x = float(x)
    jv = special.jv(l + 0.5, x)
    if abs(x) <= math.sqrt(np.finfo(float).eps):
        return 1.0 if l == 0 else 0.0
    return math.sqrt(math.pi / (2 * x)) * jv


def sphericalbessely(l, x):
    x = float(x)
    return math.sqrt(math.pi / (2 * x)) * special.yv(l + 0.5, x)


def knl_zero_function_potential(k, l, rmin, rmax):
    scale = k ** 3 / (1 + abs(k) ** 3)
    jlmax = special.jv(l - 0.5, k * rmax)
    if k * rmin == 0:
        return jlmax
    ylmax = special.yv(l - 0.5, k * rmax) * scale
    jlmin = special.jv(l + 1.5, k * rmin)
    ylmin = special.yv(l + 1.5, k * rmin) * scale
    return jlmax * ylmin - jlmin * ylmax


def _sign(v):
    return float(np.sign(v))


def doublesecant_method(fn, a, b, maxevals=1000, xrtol=np.finfo(float).eps):
    fa = fn(a)
    fb = fn(b)
    assert _sign(fa) * _sign(fb) <= 0
    neval = 0
    while a != b and neval < maxevals:
        m = (fb - fa) / (b - a)
        c = b - fb / m
        if abs(c - b) <= xrtol * c or abs(c - a) <= xrtol * c:
            return c
        fc = fn(c)
        d = b - (fb + fc) / m
        if _sign(fa) * _sign(fc) <= 0:
            b, fb = c, fc
        elif _sign(fb) * _sign(fc) <= 0:
            a, fa = c, fc
        else:
            a, fa, b, fb = c, fc, c, fc
        if a <= d <= b:
            fd = fn(d)
            if _sign(fa) * _sign(fd) <= 0:
                b, fb = d, fd
            elif _sign(fb) * _sign(fd) <= 0:
                a, fa = d, fd
            else:
                a, fa, b, fb = d, fd, d, fd
        neval += 1
    return math.nan


def calc_next_fn_zero(func, x, delta, maxevals=1000, xrtol=1e-10):
    f = func(x)
    xnew = x + delta
    fnew = func(xnew)
    while _sign(fnew) == _sign(f):
        x, f = xnew, fnew
        xnew += delta
        fnew = func(xnew)
    z = doublesecant_method(func, x, xnew, maxevals=maxevals, xrtol=xrtol)
    assert math.isfinite(z)
    return z


def calc_zeros(func, xmin, xmax, delta):
    out = []
    while xmin <= xmax:
        x = calc_next_fn_zero(func, xmin, delta)
        if xmin <= x <= xmax:
            out.append(x)
        xmin = x + delta
    return out


def calc_knl_zeros(kmax, rmin, rmax):
    nmax = math.ceil(kmax * rmax / math.pi) + 1
    lmax = math.ceil(kmax * rmax)
    knl = np.full((nmax, lmax + 1), np.nan)
    for l in range(lmax + 1):
        delta = math.pi / rmax / 4
        kmin = (l + 1.5) / rmax
        kn = calc_zeros(lambda k: knl_zero_function_potential(k, l, rmin, rmax), kmin, kmax, delta)
        if nmax < len(kn):
            raise RuntimeError("nmax too small")
        if l == lmax:
            assert len(kn) == 0
        knl[: len(kn), l] = kn
    return knl


def calc_knl_fixed(nmax, lmax, rmin, rmax):
    """AnlmModes(nmax, lmax, ...) roots (SphericalBesselGNLs.jl:294-306, :175-188)."""
    knl = np.full((nmax, lmax + 1), np.nan)
    for l in range(lmax + 1):
        delta = math.pi / rmax / 4
        xmin = (l + 1.5) / rmax
        func = lambda k: knl_zero_function_potential(k, l, rmin, rmax)  # noqa: E731
        for n in range(nmax):
            knl[n, l] = calc_next_fn_zero(func, xmin, delta)
            xmin = knl[n, l] + delta
    return knl


def calc_sphbes_gnl(q, l, c, d):
    jl = sphericalbesselj(l, q)
    if d == 0:
        return c * jl
    return c * jl + d * sphericalbessely(l, q)


def calc_cnl_dnl_potential(k, n, l, rmin, rmax):
    x = k * rmin
    ymin = -math.inf if x == 0 else special.yv(l + 1.5, x)
    dc = -special.jv(l + 1.5, x) / ymin
    g_rmin = calc_sphbes_gnl(k * rmin, l, 1.0, dc)
    g_rmax = calc_sphbes_gnl(k * rmax, l, 1.0, dc)
    num_one = (rmax * rmax * rmax * (g_rmax * g_rmax) - rmin * rmin * rmin * (g_rmin * g_rmin)) / 2
    assert num_one >= 0
    expo = n + (1 if l > 0 else 0) * (1 if n > 1 else 0)
    c = (-1.0) ** expo / math.sqrt(num_one)
    return c, dc * c


def estimate_nside(lmax):
    return 2 ** max(2, math.ceil(math.log2((2 * lmax + 1) / 2)))


class OracleModes:
    """AnlmModes(kmax, rmin, rmax; nside) or AnlmModes(nmax, lmax, rmin, rmax; nside)."""

    def __init__(self, kmax=None, rmin=0.0, rmax=1.0, nside=None, nmax=None, lmax=None):
        rmin, rmax = float(rmin), float(rmax)
        if kmax is not None:
            knl_full = calc_knl_zeros(float(kmax), rmin, rmax)
        else:
            knl_full = calc_knl_fixed(int(nmax), int(lmax), rmin, rmax)
        self.rmin, self.rmax = rmin, rmax
        self.knl_full = knl_full
        nmax_full, lmaxp1 = knl_full.shape
        cnl = np.full(knl_full.shape, np.nan)
        dnl = np.full(knl_full.shape, np.nan)
        for n in range(1, nmax_full + 1):
            for l in range(lmaxp1):
                k = knl_full[n - 1, l]
                if np.isfinite(k):
                    cnl[n - 1, l], dnl[n - 1, l] = calc_cnl_dnl_potential(k, n, l, rmin, rmax)
        self.cnl, self.dnl = cnl, dnl
        if kmax is not None:
            modes = np.nan_to_num(knl_full, nan=np.inf) <= float(kmax)
            self.nmax = int(np.nonzero(modes[:, 0])[0][-1]) + 1
            self.lmax_n = np.array([int(np.nonzero(modes[n])[0][-1]) if modes[n].any() else -1
                                    for n in range(self.nmax)], dtype=np.int64)
            self.lmax = int(np.nonzero(modes[0])[0][-1])
            self.nmax_l = np.array([int(np.nonzero(modes[:, l])[0][-1]) + 1 if modes[:, l].any() else -1
                                    for l in range(self.lmax + 1)], dtype=np.int64)
        else:
            self.nmax, self.lmax = int(nmax), int(lmax)
            self.lmax_n = np.full(self.nmax, self.lmax, dtype=np.int64)
            self.nmax_l = np.full(self.lmax + 1, self.nmax, dtype=np.int64)
        self.nside = int(nside) if nside is not None else estimate_nside(self.lmax)
        self.knl = knl_full[: self.nmax, : self.lmax + 1].copy()
        self.cnl_full, self.dnl_full = cnl, dnl
        self.cnl = cnl[: self.nmax, : self.lmax + 1].copy()
        self.dnl = dnl[: self.nmax, : self.lmax + 1].copy()
        self.kmax = float(np.nanmax(knl_full))
        # spline cache (gen_gnl_cache, SphericalBesselGNLs.jl:558-579), in map order
        xs, ys, yps, offs = [], [], [], [0]
        for n in range(1, self.nmax + 1):
            for l in range(self.lmax_n[n - 1] + 1):
                x, y = self.knots(n, l)
                xs.append(x)
                ys.append(y)
                yps.append(spline_build(x, y))
                offs.append(offs[-1] + x.size)
        self.knot_off = np.array(offs, dtype=np.int64)
        self.knot_x = np.concatenate(xs)
        self.knot_y = np.concatenate(ys)
        self.knot_ypp = np.concatenate(yps)

    def knots(self, n, l):
        k = self.knl_full[n - 1, l]
        rmin, rmax = self.rmin, self.rmax
        nr = 3 + 12 * math.ceil(2 * k * (rmax - rmin))
        dr = (rmax - rmin) / nr
        x = julia_float_range(rmin, dr, rmax + dr / 2)
        c, d = self.cnl_full[n - 1, l], self.dnl_full[n - 1, l]
        y = np.array([calc_sphbes_gnl(k * r, l, c, d) for r in x])
        return x, y

    def gnl_exact(self, n, l, r):
        k = self.knl_full[n - 1, l]
        c, d = self.cnl_full[n - 1, l], self.dnl_full[n - 1, l]
        return np.array([calc_sphbes_gnl(k * float(v), l, c, d) for v in np.atleast_1d(r)])

    def gnl(self, n, l, r):
        q = self.q_of(n, l)
        o0, o1 = self.knot_off[q], self.knot_off[q + 1]
        return spline_eval(self.knot_x[o0:o1], self.knot_y[o0:o1], self.knot_ypp[o0:o1], r)

    @property
    def nmaps(self):
        return int(np.sum(self.lmax_n + 1))

    def q_of(self, n, l):
        return int(np.sum(self.lmax_n[: n - 1] + 1)) + l

    def nlmsize(self):
        return int(sum((L + 1) * (L + 2) // 2 for L in self.lmax_n))

    def getidx(self, n, l, m):
        """1-based, as modes.jl:223-232."""
        return 1 + int(sum((L + 1) * (L + 2) // 2 for L in self.lmax_n[: n - 1])) + l * (l + 1) // 2 + m


def cat2amln(om: OracleModes, rtp, nbar, win_kind, win_a, win_b=None, w=None, niter=3,
             q_range=None, nthreads=0):
    """Oracle cat2amln.  rtp is (3,N) as the reference passes it.  Returns complex anlm."""
    rtp = np.asarray(rtp, dtype=np.float64)
    N = rtp.shape[1] if rtp.size else 0
    rtp_c = np.ascontiguousarray(rtp.T)  # N x 3 row-major == 3 x N column-major
    w = None if w is None or len(w) == 0 else np.ascontiguousarray(w, dtype=np.float64)
    win_a = np.ascontiguousarray(win_a, dtype=np.float64)
    win_b = None if win_b is None else np.ascontiguousarray(win_b, dtype=np.float64)
    lmax_n = np.ascontiguousarray(om.lmax_n, dtype=np.int64)
    out = np.full(2 * om.nlmsize(), np.nan)
    q0, q1 = (0, om.nmaps) if q_range is None else q_range
    rc = lib().orc_cat2amln(om.nside, om.lmax, om.nmax, _p(lmax_n), _p(om.knot_off), _p(om.knot_x),
                            _p(om.knot_y), _p(om.knot_ypp), N, _p(rtp_c), _p(w), float(nbar), int(win_kind),
                            _p(win_a), _p(win_b), int(niter), int(q0), int(q1), int(nthreads), _p(out))
    if rc == -2:
        raise RuntimeError("lmax > 4*nside")
    if rc == -3:
        raise AssertionError("non-finite anlm")
    return out[0::2] + 1j * out[1::2]


def delta_map(om: OracleModes, rtp, nbar, win_kind, win_a, win_b, w, q):
    rtp = np.asarray(rtp, dtype=np.float64)
    N = rtp.shape[1] if rtp.size else 0
    rtp_c = np.ascontiguousarray(rtp.T)
    w = None if w is None or len(w) == 0 else np.ascontiguousarray(w, dtype=np.float64)
    win_a = np.ascontiguousarray(win_a, dtype=np.float64)
    win_b = None if win_b is None else np.ascontiguousarray(win_b, dtype=np.float64)
    lmax_n = np.ascontiguousarray(om.lmax_n, dtype=np.int64)
    out = np.empty(12 * om.nside * om.nside)
    lib().orc_delta_map(om.nside, om.lmax, om.nmax, _p(lmax_n), _p(om.knot_off), _p(om.knot_x),
                        _p(om.knot_y), _p(om.knot_ypp), N, _p(rtp_c), _p(w), float(nbar), int(win_kind),
                        _p(win_a), _p(win_b), int(q), _p(out))
    return out


# --------------------------------------------------------------------------
# Pixel window (what src/SphericalFourierBesselDecompositions.jl:158-167 reads from HEALPix's
# tables through Healpix.pixwin), restated from its definition (Gorski et al. 2005):
# w_l^2 = (1/Npix) sum_p (4 pi / (2l + 1)) sum_{m=-l..l} |<Y_lm>_p|^2, <.>_p the mean over
# pixel p.  The means are taken over the centres of every pixel's 4^k sub-pixels at nside 2^k
# (assigned to pixels by ang2pix), for every pixel of the map, and depths k - 1 and k are
# combined as (4 W_k - W_{k-1}) / 3.
# --------------------------------------------------------------------------
def _subpixels(nside, k):
    nh = nside << k
    th, ph = pix2ang_ring(nh, np.arange(12 * nh * nh))
    par = ang2pix_ring(nside, th, ph)
    if not np.all(np.bincount(par, minlength=12 * nside * nside) == 4 ** k):
        raise AssertionError("sub-pixels do not tile the pixels")
    return th, ph, par


def pixwin_ylm(nside, lmax, depth):
    """w_l from the pixel means of Y_lm (scipy sph_harm_y), every pixel."""
    from scipy import special
    npix = 12 * nside * nside

    def W(k):
        th, ph, par = _subpixels(nside, k)
        out = np.zeros(lmax + 1)
        for l in range(lmax + 1):
            s = np.zeros(npix)
            for m in range(-l, l + 1):
                y = special.sph_harm_y(l, m, th, ph)
                mr = np.bincount(par, weights=y.real, minlength=npix) / 4 ** k
                mi = np.bincount(par, weights=y.imag, minlength=npix) / 4 ** k
                s += mr * mr + mi * mi
            out[l] = np.mean(4 * np.pi / (2 * l + 1) * s)
        return out

    return np.sqrt(np.maximum((4 * W(depth) - W(depth - 1)) / 3, 0.0))


def pixwin_pairs(nside, lmax, depth):
    """w_l from the addition theorem: w_l^2(p) = mean of P_l(n . n') over the pairs of
    sub-pixel centres of p (scipy eval_legendre), every pixel."""
    from scipy import special
    npix = 12 * nside * nside

    def W(k):
        th, ph, par = _subpixels(nside, k)
        v = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
        out = np.zeros(lmax + 1)
        for p in range(npix):
            u = v[par == p]
            x = np.clip(u @ u.T, -1.0, 1.0)
            for l in range(lmax + 1):
                out[l] += special.eval_legendre(l, x).mean()
        return out / npix

    return np.sqrt(np.maximum((4 * W(depth) - W(depth - 1)) / 3, 0.0))
