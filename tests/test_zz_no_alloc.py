"""sfb_apply_window_noalloc (device arrays, static scratch, no allocation) against
sfb_apply_window, through the C ABI: two calls in a row, the second with fewer galaxies,
and N beyond the scratch refused before any device work."""
import ctypes

import numpy as np
import pytest

import sfb_amd as S
from sfb_amd import _lib as L
from oracle import oracle as O

MAX_N = 1 << 22


def say(*a):
    print(*a, flush=True)


def test_noalloc_rejects_n_beyond_scratch():
    n = ctypes.c_int64()
    rc = L.lib().sfb_apply_window_noalloc(0, 8, 4, 500.0, 1000.0, 125.0, ctypes.c_void_p(8), MAX_N + 1,
                                          ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8),
                                          ctypes.c_void_p(8), None)
    assert rc == L.SFB_ERR_ARG
    assert "scratch" in L.lib().sfb_last_error().decode()


@pytest.mark.gpu
def test_noalloc_matches_apply_window():
    """Device memory and a stream from the HIP runtime the library links (no second runtime)."""
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    H2D, D2H = 1, 2
    ptrs, stream = [], ctypes.c_void_p()

    def dmalloc(nbytes):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        ptrs.append(p)
        return p.value

    def copy(dst, src, nbytes, kind):
        assert hip.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), ctypes.c_size_t(nbytes), kind) == 0

    wm = S.ConfigurationSpaceModes(500.0, 1000.0, 37, 16)
    rng = np.random.default_rng(61)
    win = rng.uniform(0.0, 1.0, (wm.nr, wm.npix))
    win_f = np.asfortranarray(win).reshape(-1, order="F").copy()
    try:
        assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
        d_win = dmalloc(win_f.nbytes)
        copy(d_win, win_f.ctypes.data, win_f.nbytes, H2D)
        NMAX = 300_000
        d_rtp, d_u, d_out, d_n = dmalloc(24 * NMAX), dmalloc(8 * NMAX), dmalloc(24 * NMAX), dmalloc(8)
        for N, st in ((NMAX, None), (70_001, stream.value)):  # the second call smaller, on a caller stream
            rtp = np.stack([rng.uniform(480.0, 1020.0, N), np.arccos(rng.uniform(-1, 1, N)),
                            rng.uniform(0, 2 * np.pi, N)])
            u = rng.random(N)
            say("apply_window_noalloc N =", N)
            ref = S.apply_window(rtp, win, wm, u=u)
            rtp_c = np.ascontiguousarray(rtp.T)
            copy(d_rtp, rtp_c.ctypes.data, rtp_c.nbytes, H2D)
            copy(d_u, u.ctypes.data, u.nbytes, H2D)
            poison = np.full((N, 3), np.nan)
            copy(d_out, poison.ctypes.data, poison.nbytes, H2D)
            L.check(L.lib().sfb_apply_window_noalloc(0, wm.nside, wm.nr, wm.rmin, wm.rmax, wm.dr,
                                                     ctypes.c_void_p(d_win), N, ctypes.c_void_p(d_rtp),
                                                     ctypes.c_void_p(d_u), ctypes.c_void_p(d_out),
                                                     ctypes.c_void_p(d_n), None if st is None else ctypes.c_void_p(st)))
            assert hip.hipDeviceSynchronize() == 0
            n = np.zeros(1, dtype=np.int64)
            copy(n.ctypes.data, d_n, 8, D2H)
            out = np.empty((N, 3))
            copy(out.ctypes.data, d_out, out.nbytes, D2H)
            got = out[: int(n[0])].T
            assert int(n[0]) == ref.shape[1]
            assert np.array_equal(got, ref)
            # the keep rule itself, from the oracle's pixel index
            pix = O.ang2pix_ring(16, rtp[1], rtp[2])
            idx_r = np.ceil((rtp[0] - 500.0) / wm.dr).astype(np.int64)
            idx_r[idx_r == 0] = 1
            ok = (rtp[0] >= 500.0) & (rtp[0] <= 1000.0) & (idx_r <= wm.nr)
            keep = np.zeros(N, dtype=bool)
            keep[ok] = u[ok] <= win[idx_r[ok] - 1, pix[ok]] * (1.0 / win.max())
            assert np.array_equal(got, rtp[:, keep])
    finally:
        if stream.value:
            hip.hipStreamDestroy(stream)
        for p in ptrs:
            hip.hipFree(p)
