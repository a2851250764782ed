"""Size-independent properties of the HIP cat2amln at the headline size
(N = 1e6, kmax = 0.1, nside = 128: 2514 maps), where the oracle is too slow to
run in full, plus the reference's error behaviour.

* Permutation invariance: galaxies are summed per pixel in (r, input index)
  order, so shuffling a catalogue with distinct radii gives bit-identical a_nlm.
* Affine structure: cat2amln(w) = A(S(w) - W) with S linear, so
  f(2w) - 2 f(w) + f(empty catalogue) = 0 up to rounding.
* The device entry point (inputs in HBM) against the host one: bit-identical.
* A bounded oracle sample: the first maps of the same call against the oracle.
* Errors: a NaN weight gives AssertionError (src/cat2anlm.jl:308), a NaN angle is
  rejected (Healpix.jl's floor(Int, NaN) throws), lmax > 4 nside is refused.
"""
import numpy as np
import pytest

import sfb_amd as S
from helpers import rel_errors, separable_unit_window, shell_volume, uniform_shell_catalog
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c2():
    am = S.AnlmModes(0.1, 500.0, 1500.0, nside=128)
    win, _ = separable_unit_window(S, am)
    cat = uniform_shell_catalog(1_000_000, 500.0, 1500.0, seed=0)
    nbar = cat.shape[1] / shell_volume(500.0, 1500.0)
    plan = S.get_plan(am)
    return am, win, cat, nbar, plan


def test_permutation_invariance_bitexact(c2):
    am, win, cat, nbar, plan = c2
    say("C2 permutation invariance")
    a = plan.cat2amln(cat, nbar, win)
    perm = np.random.default_rng(11).permutation(cat.shape[1])
    b = plan.cat2amln(cat[:, perm], nbar, win)
    assert np.array_equal(a, b)


def test_affine_in_weights(c2):
    am, win, cat, nbar, plan = c2
    say("C2 affine structure")
    w = np.random.default_rng(12).uniform(0.5, 1.5, cat.shape[1])
    f1 = plan.cat2amln(cat, nbar, win, w)
    f2 = plan.cat2amln(cat, nbar, win, 2 * w)
    f0 = plan.cat2amln(np.zeros((3, 0)), nbar, win, [])
    resid = f2 - 2 * f1 + f0
    assert np.max(np.abs(resid)) < 1e-11 * np.max(np.abs(f2))


def test_device_entry_matches_host_entry(c2):
    """sfb_cat2amln_device (catalogue, window and output already in HBM: the call bench.py
    times) returns the a_nlm of sfb_cat2amln from host arrays bit for bit, on both paths, with
    unit and given weights, on the plan's stream and on a caller's stream.  Device memory and
    the stream come from the HIP runtime the library itself links (no second runtime)."""
    import ctypes
    from sfb_amd import _lib as L
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    H2D, D2H = 1, 2
    am, win, cat, nbar, plan = c2
    N = cat.shape[1]
    w = np.random.default_rng(13).uniform(0.5, 1.5, N)
    host_in = [np.ascontiguousarray(cat.T), w, np.ascontiguousarray(win.mask, dtype=np.float64),
               np.ascontiguousarray(win.w_ln.reshape(-1, order="F"))]
    nout = 2 * S.getnlmsize(am)
    ptrs, stream = [], ctypes.c_void_p()

    def dmalloc(nbytes):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        ptrs.append(p)
        return p.value

    def copy(dst, src, nbytes, kind):
        assert hip.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), ctypes.c_size_t(nbytes), kind) == 0

    try:
        d_rtp, d_w, d_mask, d_wln = (dmalloc(a.nbytes) for a in host_in)
        for d, a in zip((d_rtp, d_w, d_mask, d_wln), host_in):
            copy(d, a.ctypes.data, a.nbytes, H2D)
        d_out = dmalloc(8 * nout)
        assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
        for prec in ("fp64", "mixed"):
            plan.set_precision(prec)
            for ww, dw in ((None, None), (w, d_w)):
                host = plan.cat2amln(cat, nbar, win, ww)
                for st in (None, stream.value):
                    say("device entry", prec, "unit" if ww is None else "weights",
                        "plan stream" if st is None else "caller stream")
                    out = np.full(nout, np.nan)
                    copy(d_out, out.ctypes.data, out.nbytes, H2D)
                    plan.cat2amln_device(N, d_rtp, dw, nbar, L.SFB_WIN_SEPARABLE, d_mask, d_wln, d_out, stream=st)
                    copy(out.ctypes.data, d_out, out.nbytes, D2H)
                    assert np.array_equal(out[0::2] + 1j * out[1::2], host), (prec, ww is None, st)
    finally:
        plan.set_precision("fp64")
        if stream.value:
            hip.hipStreamDestroy(stream)
        for p in ptrs:
            hip.hipFree(p)



def test_host_staging_edges_match_device_entry():
    """The host-array entry point stages the caller's arrays through pinned memory in 4 MiB
    chunks (each chunk's transfer enqueued as soon as a thread pool has copied it); the result
    must be the device entry point's bit for bit at the sizes around the chunking: no galaxy,
    one, a catalogue of 4 MiB + 24 bytes (a second chunk of one galaxy), unit and given weights,
    separable and dense windows, FP64 and mixed."""
    import ctypes
    from sfb_amd import _lib as L
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    H2D, D2H = 1, 2
    am = S.AnlmModes(0.03, 500.0, 1000.0, nside=16)
    win, _ = separable_unit_window(S, am, nr=100)
    dense = win.dense()
    plan = S.Plan(am)
    nout = 2 * S.getnlmsize(am)
    ptrs = []

    def dmalloc(nbytes):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 8))) == 0
        ptrs.append(p)
        return p.value

    def copy(dst, src, nbytes, kind):
        if nbytes:
            assert hip.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), ctypes.c_size_t(nbytes), kind) == 0

    try:
        d_out = dmalloc(8 * nout)
        for N in (0, 1, (4 << 20) // 24 + 1):
            cat = uniform_shell_catalog(N, 500.0, 1000.0, seed=N + 3) if N else np.zeros((3, 0))
            nbar = max(N, 1) / shell_volume(500.0, 1000.0)
            w = np.random.default_rng(N).uniform(0.5, 1.5, N)
            rtp = np.ascontiguousarray(cat.T)
            d_rtp, d_w = dmalloc(rtp.nbytes), dmalloc(w.nbytes)
            copy(d_rtp, rtp.ctypes.data, rtp.nbytes, H2D)
            copy(d_w, w.ctypes.data, w.nbytes, H2D)
            for kind, wwin in (("separable", win), ("dense", dense)):
                if kind == "separable":
                    wa = np.ascontiguousarray(win.mask, dtype=np.float64)
                    wb = np.ascontiguousarray(win.w_ln.reshape(-1, order="F"))
                    code = L.SFB_WIN_SEPARABLE
                else:
                    wa, wb, code = np.ascontiguousarray(dense.reshape(-1, order="F")), None, L.SFB_WIN_DENSE
                d_wa = dmalloc(wa.nbytes)
                copy(d_wa, wa.ctypes.data, wa.nbytes, H2D)
                d_wb = None
                if wb is not None:
                    d_wb = dmalloc(wb.nbytes)
                    copy(d_wb, wb.ctypes.data, wb.nbytes, H2D)
                for prec in ("fp64", "mixed"):
                    plan.set_precision(prec)
                    for ww, dw in ((None, None), (w, d_w)):
                        say("host staging", N, kind, prec, "unit" if ww is None else "weights")
                        host = plan.cat2amln(cat, nbar, wwin, ww)
                        out = np.full(nout, np.nan)
                        copy(d_out, out.ctypes.data, out.nbytes, H2D)
                        plan.cat2amln_device(N, d_rtp if N else None, dw if N else None, nbar, code, d_wa, d_wb,
                                             d_out)
                        copy(out.ctypes.data, d_out, out.nbytes, D2H)
                        assert np.array_equal(out[0::2] + 1j * out[1::2], host), (N, kind, prec, ww is None)
    finally:
        plan.close()
        for p in ptrs:
            hip.hipFree(p)

def test_c2_sample_against_oracle(c2):
    am, win, cat, nbar, plan = c2
    say("C2 oracle sample")
    a = plan.cat2amln(cat, nbar, win)
    om = O.OracleModes.__new__(O.OracleModes)
    om.nside, om.lmax, om.nmax, om.lmax_n = am.nside, am.lmax, am.nmax, am.lmax_n
    om.knot_off, om.knot_x, om.knot_y, om.knot_ypp = am.knot_off, am.knot_x, am.knot_y, am.knot_ypp
    q1 = 12
    ref = O.cat2amln(om, cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"), q_range=(0, q1))
    n_out = sum(l + 1 for l in range(q1))  # maps 0..q1-1 are n = 1, l = 0..q1-1
    d = np.abs(a[:n_out] - ref[:n_out])
    assert np.max(d) < 1e-10 * np.max(np.abs(ref[:n_out]))


def test_nan_weight_is_assertion_error():
    am = S.AnlmModes(0.03, 500.0, 1000.0, nside=16)
    win, _ = separable_unit_window(S, am, nr=100)
    cat = uniform_shell_catalog(1000, 500.0, 1000.0, seed=2)
    w = np.ones(1000)
    w[17] = np.nan
    with pytest.raises(AssertionError):
        S.cat2amln(cat, am, 1e-4, win, w)
    cat[1, 5] = np.nan
    with pytest.raises(S.SFBError) as e:
        S.cat2amln(cat, am, 1e-4, win)
    assert "not finite" in str(e.value)
    # the plan is still usable after errors
    assert np.all(np.isfinite(S.cat2amln(cat[:, :4], am, 1e-4, win)))
