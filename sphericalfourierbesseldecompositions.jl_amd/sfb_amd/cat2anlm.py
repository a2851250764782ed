"""cat2amln on the MI355X -- host mirror of src/cat2anlm.jl:257-312.

``cat2amln(rθϕ, amodes, nbar, win_rhat_ln, weight)`` has the reference's
positional signature and return value (a complex vector of length
getnlmsize(amodes), ordered n, l, m).  Keyword arguments select the map2alm
iteration count (default 3, src/healpix_helpers.jl:50), the Legendre-step
precision and the device.  Plans (device tables, work buffers) are cached per
(amodes, niter, precision, device).
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _lib as L
from .modes import AnlmModes, getnlmsize
from .windows import SeparableArray

_PRECISIONS = {"fp64": L.SFB_PREC_FP64, "mixed": L.SFB_PREC_MIXED, "bf16x3": L.SFB_PREC_BF16X3,
               "bf16x2": L.SFB_PREC_BF16X2}


class Plan:
    """A device plan for one AnlmModes (C ABI sfb_plan)."""

    def __init__(self, amodes: AnlmModes, niter=3, precision="fp64", device=0):
        self.amodes = amodes
        self.niter = int(niter)
        self.precision = precision
        self.device = int(device)
        h = ctypes.c_void_p()
        L.check(L.lib().sfb_plan_create(ctypes.byref(h), amodes.nside, amodes.lmax, amodes.nmax,
                                        L.ptr(amodes.lmax_n), L.ptr(amodes.knot_off), L.ptr(amodes.knot_x),
                                        L.ptr(amodes.knot_y), L.ptr(amodes.knot_ypp), self.niter,
                                        _PRECISIONS[precision], self.device))
        self._h = h
        self._fin = weakref.finalize(self, L.lib().sfb_plan_destroy, h)
        self.nlmsize = getnlmsize(amodes)
        self.npix = 12 * amodes.nside ** 2

    @property
    def handle(self):
        return self._h

    def set_precision(self, precision):
        L.check(L.lib().sfb_plan_set_precision(self._h, _PRECISIONS[precision]))
        self.precision = precision

    def close(self):
        self._fin()

    def profile(self, enable=True):
        """Per-phase HIP-event timing of subsequent calls (clears the totals)."""
        L.check(L.lib().sfb_plan_profile(self._h, int(bool(enable))))

    def phase_times(self):
        """{phase: (total_ms, launches)} and the number of profiled calls."""
        ms = np.zeros(len(L.PHASES))
        n = np.zeros(len(L.PHASES), dtype=np.int64)
        calls = ctypes.c_int64()
        L.check(L.lib().sfb_plan_phase_times(self._h, L.ptr(ms), L.ptr(n), ctypes.byref(calls)))
        return {ph: (float(ms[i]), int(n[i])) for i, ph in enumerate(L.PHASES)}, calls.value

    def set_batch_graph(self, enable=True):
        """Batched mocks as one captured HIP graph (replayed when a batch call repeats);
        returns the wall time (ms) of the last capture and instantiation."""
        ms = ctypes.c_double()
        L.check(L.lib().sfb_plan_set_batch_graph(self._h, -1 if enable is None else int(bool(enable)),
                                                 ctypes.byref(ms)))
        return ms.value

    def tiles(self):
        """The phi-transform tile shapes the plan's tile search chose, and its time (ms)."""
        t = np.zeros(6, dtype=np.int32)
        ms = ctypes.c_double()
        L.check(L.lib().sfb_plan_tiles(self._h, L.ptr(t), ctypes.byref(ms)))
        keys = ("fft_maps_fp64", "fft_threads_fp64", "fft_maps_fp32", "fft_threads_fp32", "polar_maps_fp64",
                "polar_maps_fp32")
        return dict(zip(keys, (int(x) for x in t))), ms.value

    # -- helpers --------------------------------------------------------
    def _window(self, win):
        am = self.amodes
        if isinstance(win, SeparableArray):
            a = L.f64(np.asarray(win.arr1).reshape(-1))
            b = L.f64(np.asarray(win.arr2).reshape(-1, order="F"))
            if a.size != self.npix or b.size != (am.lmax + 1) * am.nmax:
                raise ValueError("separable window must be (mask[npix], w_ln[lmax+1, nmax])")
            return L.SFB_WIN_SEPARABLE, a, b
        w = np.asarray(win, dtype=np.float64)
        if w.shape != (self.npix, am.lmax + 1, am.nmax):
            raise ValueError(f"dense window must have shape {(self.npix, am.lmax + 1, am.nmax)}, got {w.shape}")
        return L.SFB_WIN_DENSE, L.f64(w.reshape(-1, order="F")), None

    @staticmethod
    def _catalogue(rtp, weight):
        rtp = np.asarray(rtp, dtype=np.float64)
        if rtp.size == 0:
            return 0, np.zeros(3), None
        if rtp.ndim != 2 or rtp.shape[0] != 3:
            raise ValueError("rθϕ must be a 3 x N array")
        N = rtp.shape[1]
        host = L.f64(rtp.T)  # N x 3 row-major == 3 x N column-major
        w = None
        if weight is not None and len(weight) > 0:
            w = L.f64(weight)
            if w.size != N:
                raise ValueError("weight must have one entry per galaxy")
        return N, host, w

    def cat2amln(self, rtp, nbar, win_rhat_ln, weight=None):
        N, host, w = self._catalogue(rtp, weight)
        kind, a, b = self._window(win_rhat_ln)
        out = np.empty(2 * self.nlmsize)
        L.check(L.lib().sfb_cat2amln(self._h, N, L.ptr(host), L.ptr(w), float(nbar), kind, L.ptr(a), L.ptr(b),
                                     L.ptr(out)))
        return out.view(np.complex128)  # (re, im) pairs are numpy's complex layout

    def cat2amln_device(self, N, d_rtp, d_w, nbar, win_kind, d_win_a, d_win_b, d_out, stream=None):
        """All arrays are device pointers (ints) or torch CUDA tensors."""
        def dp(t):
            if t is None:
                return None
            return ctypes.c_void_p(t if isinstance(t, int) else t.data_ptr())
        L.check(L.lib().sfb_cat2amln_device(self._h, int(N), dp(d_rtp), dp(d_w), float(nbar), int(win_kind),
                                            dp(d_win_a), dp(d_win_b), dp(d_out),
                                            None if stream is None else ctypes.c_void_p(stream)))

    # -- window and field transforms (src/windows.jl:244-270, src/cat2anlm.jl:323-422)
    def win_rhat_ln(self, win, wmodes):
        """win_rhat_ln on the device: dense (nr, npix) window -> (npix, lmax+1, nmax)
        array, or SeparableArray(phi, mask) -> SeparableArray(mask, w_ln)."""
        am = self.amodes
        r = L.f64(wmodes.r)
        if isinstance(win, SeparableArray):
            phi = L.f64(np.asarray(win.arr1).reshape(-1))
            if phi.size != r.size:
                raise ValueError("phi must have one value per radial bin")
            out = np.empty((am.lmax + 1) * am.nmax)
            L.check(L.lib().sfb_win_rhat_ln(self._h, r.size, L.ptr(r), float(wmodes.dr), L.SFB_WIN_SEPARABLE,
                                            L.ptr(phi), L.ptr(out)))
            return SeparableArray(np.asarray(win.arr2, dtype=float), out.reshape(am.lmax + 1, am.nmax, order="F"),
                                  name1="mask", name2="w_ln")
        w = np.asarray(win, dtype=np.float64)
        if w.shape != (r.size, self.npix):
            raise ValueError(f"dense window must be (nr, npix) = {(r.size, self.npix)}")
        wf = np.asfortranarray(w)
        out = np.empty((self.npix, am.lmax + 1, am.nmax), order="F")
        L.check(L.lib().sfb_win_rhat_ln(self._h, r.size, L.ptr(r), float(wmodes.dr), L.SFB_WIN_DENSE,
                                        ctypes.c_void_p(wf.ctypes.data), ctypes.c_void_p(out.ctypes.data)))
        return out

    def field2anlm(self, f_xyz, wmodes):
        """field2anlm (src/cat2anlm.jl:323-358): nr x npix field -> a_nlm."""
        f = np.asfortranarray(np.asarray(f_xyz, dtype=np.float64))
        if f.shape != (wmodes.nr, self.npix):
            raise ValueError(f"field must be (nr, npix) = {(wmodes.nr, self.npix)}")
        r = L.f64(wmodes.r)
        out = np.empty(2 * self.nlmsize)
        L.check(L.lib().sfb_field2anlm(self._h, r.size, L.ptr(r), float(wmodes.dr), ctypes.c_void_p(f.ctypes.data),
                                       L.ptr(out)))
        return out[0::2] + 1j * out[1::2]

    def anlm2field(self, f_nlm, wmodes):
        """anlm2field (src/cat2anlm.jl:383-422): a_nlm -> nr x npix field."""
        c = np.asarray(f_nlm, dtype=np.complex128).reshape(-1)
        if c.size != self.nlmsize:
            raise ValueError(f"f_nlm must have {self.nlmsize} entries")
        cot = L.f64(np.stack([c.real, c.imag], axis=1).reshape(-1))
        r = L.f64(wmodes.r)
        out = np.empty((r.size, self.npix), order="F")
        L.check(L.lib().sfb_anlm2field(self._h, r.size, L.ptr(r), L.ptr(cot), ctypes.c_void_p(out.ctypes.data)))
        return out

    def cat2amln_batch(self, catalogues, nbars, win_rhat_ln, weights=None):
        """cat2amln of every catalogue in `catalogues` (3 x N_b arrays) with its own nbar
        and the shared window; returns a (B, nlmsize) complex array."""
        B = len(catalogues)
        nbars = L.f64(np.broadcast_to(np.asarray(nbars, dtype=np.float64), (B,)))
        parts = [np.asarray(c, dtype=np.float64).reshape(3, -1) for c in catalogues]
        N_b = np.array([p.shape[1] for p in parts], dtype=np.int64)
        host = L.f64(np.concatenate([p.T for p in parts], axis=0)) if N_b.sum() else np.zeros(3)
        w = None
        if weights is not None:
            w = L.f64(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in weights]))
            if w.size != N_b.sum():
                raise ValueError("weights must have one entry per galaxy of every catalogue")
        kind, a, b = self._window(win_rhat_ln)
        out = np.empty(2 * self.nlmsize * max(B, 1))
        L.check(L.lib().sfb_cat2amln_batch(self._h, B, L.ptr(N_b), L.ptr(host), L.ptr(w), L.ptr(nbars), kind,
                                           L.ptr(a), L.ptr(b), L.ptr(out)))
        out = out[:2 * self.nlmsize * B]
        return (out[0::2] + 1j * out[1::2]).reshape(B, self.nlmsize)

    def cat2amln_batch_device(self, B, N_b, d_rtp, d_w, nbars, win_kind, d_win_a, d_win_b, d_out, stream=None):
        """sfb_cat2amln_batch_device; N_b and nbars are host sequences."""
        def dp(t):
            if t is None:
                return None
            return ctypes.c_void_p(t if isinstance(t, int) else t.data_ptr())
        N_b = np.ascontiguousarray(N_b, dtype=np.int64)
        nb = L.f64(np.asarray(nbars, dtype=np.float64))
        L.check(L.lib().sfb_cat2amln_batch_device(self._h, int(B), L.ptr(N_b), dp(d_rtp), dp(d_w), L.ptr(nb),
                                                  int(win_kind), dp(d_win_a), dp(d_win_b), dp(d_out),
                                                  None if stream is None else ctypes.c_void_p(stream)))

    def cat2amln_vjp(self, rtp, nbar, dL_danlm, weight=None):
        """Vector-Jacobian product of cat2amln at (rθϕ, weight, nbar).

        dL_danlm is the cotangent of the complex a_nlm (dL/dRe a + i dL/dIm a).
        Returns (dL/dr, dL/dw, dL/dnbar); dL/dθ and dL/dφ are zero because the
        pixel assignment is piecewise constant.
        """
        N, host, w = self._catalogue(rtp, weight)
        c = np.asarray(dL_danlm, dtype=np.complex128).reshape(-1)
        if c.size != self.nlmsize:
            raise ValueError(f"dL_danlm must have {self.nlmsize} entries")
        cot = L.f64(np.stack([c.real, c.imag], axis=1).reshape(-1))
        dr = np.zeros(N)
        dw = np.zeros(N)
        dnbar = ctypes.c_double()
        L.check(L.lib().sfb_cat2amln_vjp(self._h, N, L.ptr(host), L.ptr(w), float(nbar), L.ptr(cot), L.ptr(dr),
                                         L.ptr(dw), ctypes.byref(dnbar)))
        return dr, dw, dnbar.value

    def cat2amln_vjp_device(self, N, d_rtp, d_w, nbar, d_cot, d_dr, d_dw, stream=None):
        """sfb_cat2amln_vjp_device: device pointers (ints) or torch CUDA tensors."""
        def dp(t):
            if t is None:
                return None
            return ctypes.c_void_p(t if isinstance(t, int) else t.data_ptr())
        L.check(L.lib().sfb_cat2amln_vjp_device(self._h, int(N), dp(d_rtp), dp(d_w), float(nbar), dp(d_cot),
                                                dp(d_dr), dp(d_dw),
                                                None if stream is None else ctypes.c_void_p(stream)))

    def delta_map(self, rtp, nbar, win_rhat_ln, weight, q):
        N, host, w = self._catalogue(rtp, weight)
        kind, a, b = self._window(win_rhat_ln)
        out = np.empty(self.npix)
        L.check(L.lib().sfb_delta_map(self._h, N, L.ptr(host), L.ptr(w), float(nbar), kind, L.ptr(a), L.ptr(b),
                                      int(q), L.ptr(out)))
        return out

    def map2alm(self, maps):
        maps = L.f64(np.atleast_2d(maps))
        nmap = maps.shape[0]
        nalm = (self.amodes.lmax + 1) * (self.amodes.lmax + 2) // 2
        out = np.empty(2 * nalm * nmap)
        L.check(L.lib().sfb_map2alm(self._h, nmap, L.ptr(maps), L.ptr(out)))
        return (out[0::2] + 1j * out[1::2]).reshape(nmap, nalm)


_plans: dict = {}


def get_plan(amodes, niter=3, precision="fp64", device=0) -> Plan:
    key = (id(amodes), int(niter), precision, int(device))
    p = _plans.get(key)
    if p is None or p.amodes is not amodes:
        p = Plan(amodes, niter, precision, device)
        _plans[key] = p
    return p


def field2anlm(f_xyz, wmodes, amodes, *, niter=3, device=0):
    """field2anlm(f_xyz, wmodes, amodes) (src/cat2anlm.jl:361-370) on the device."""
    return get_plan(amodes, niter, "fp64", device).field2anlm(f_xyz, wmodes)


def anlm2field(f_nlm, wmodes, amodes, *, device=0):
    """anlm2field(f_nlm, wmodes, amodes) (src/cat2anlm.jl:373-422) on the device."""
    return get_plan(amodes, 3, "fp64", device).anlm2field(f_nlm, wmodes)


def cat2amln_vjp(rθϕ, amodes, nbar, dL_danlm, weight=None, *, niter=3, precision="fp64", device=0):
    """(dL/dr, dL/dw, dL/dnbar) of cat2amln for the cotangent dL_danlm."""
    return get_plan(amodes, niter, precision, device).cat2amln_vjp(rθϕ, nbar, dL_danlm, weight)


def cat2amln(rθϕ, amodes, nbar, win_rhat_ln, weight=None, *, niter=3, precision="fp64", device=0):
    """Catalogue -> SFB coefficients a_nlm (src/cat2anlm.jl:257-310).

    rθϕ: 3 x N (r, θ, ϕ); weight: N weights or None/empty for unit weights;
    win_rhat_ln: dense (npix, lmax+1, nmax) array or SeparableArray(mask, w_ln).
    Raises AssertionError on a non-finite result, like the reference."""
    return get_plan(amodes, niter, precision, device).cat2amln(rθϕ, nbar, win_rhat_ln, weight)


cat2nlm = cat2amln


def ang2pix_ring(nside, theta, phi, device=0):
    """0-based ring-scheme pixel of (θ, ϕ), computed on the GPU."""
    th = L.f64(theta).reshape(-1)
    ph = L.f64(phi).reshape(-1)
    out = np.empty(th.size, dtype=np.int32)
    L.check(L.lib().sfb_ang2pix_ring(int(device), int(nside), th.size, L.ptr(th), L.ptr(ph), L.ptr(out)))
    return out
