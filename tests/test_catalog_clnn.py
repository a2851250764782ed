"""Catalogue preparation (xyz2rtp, winweights2galweights, apply_window) and
amln2clnn on the device.

* xyz2rtp: r is the correctly rounded sqrt of the same sum, so it is bit-identical
  to NumPy's; theta and phi agree to a few ulp (acos/atan2 of the two math libraries).
* winweights2galweights and apply_window are index work (pixel, radial bin,
  acceptance): bit-identical to the same rules evaluated with the oracle's ang2pix and
  np.searchsorted (searchsortedlast) on the host.
* amln2clnn: unit-modulus a_lm (real a_l0) give C_l(n, n) = 1; random a_nlm against
  the defining sum (src/SphericalFourierBesselDecompositions.jl:114-128) evaluated with
  NumPy; the (l, n1, n2) order and getidx follow src/modes.jl:342-362,448-478.
"""
import numpy as np
import pytest

import sfb_amd as S
from sfb_amd import clnn as C
from oracle import oracle as O

@pytest.mark.gpu
def test_xyz2rtp():
    rng = np.random.default_rng(50)
    xyz = rng.uniform(-1500, 1500, (3, 200_000))
    rtp = S.xyz2rtp(xyz)
    r = np.sqrt(xyz[0] * xyz[0] + xyz[1] * xyz[1] + xyz[2] * xyz[2])
    assert np.array_equal(rtp[0], r)
    th = np.arccos(xyz[2] / r)
    ph = np.arctan2(xyz[1] / r, xyz[0] / r)
    assert np.max(np.abs(rtp[1] - th) / np.spacing(th)) <= 4
    assert np.max(np.abs(rtp[2] - ph)) <= 4 * np.spacing(np.pi)


@pytest.mark.gpu
def test_winweights2galweights_and_apply_window():
    wm = S.ConfigurationSpaceModes(500.0, 1000.0, 37, 16)
    rng = np.random.default_rng(51)
    N = 300_000
    rtp = np.stack([rng.uniform(480.0, 1020.0, N), np.arccos(rng.uniform(-1, 1, N)), rng.uniform(0, 2 * np.pi, N)])
    rtp[0, :5] = [500.0, 1000.0, wm.r[3], S.get_rbounds(wm)[7], 750.0]
    pix = O.ang2pix_ring(16, rtp[1], rtp[2])
    rb = S.get_rbounds(wm)
    weights = rng.uniform(0.1, 2.0, (wm.nr, wm.npix))
    inside = (rtp[0] >= 500.0) & (rtp[0] <= 1000.0)
    w = S.winweights2galweights(weights, wm, rtp[:, inside])
    r_in = rtp[0, inside]
    b = np.where(r_in == rb[-1], wm.nr, np.searchsorted(rb, r_in, side="right"))
    assert np.array_equal(w, weights[b - 1, pix[inside]])
    with pytest.raises(S.SFBError):
        S.winweights2galweights(weights, wm, rtp)  # radii outside the bins
    win = rng.uniform(0.0, 1.0, (wm.nr, wm.npix))
    u = rng.random(N)
    kept = S.apply_window(rtp, win, wm, u=u)
    idx_r = np.ceil((rtp[0] - 500.0) / wm.dr).astype(np.int64)
    idx_r[idx_r == 0] = 1
    ok = (rtp[0] >= 500.0) & (rtp[0] <= 1000.0) & (idx_r <= wm.nr)
    keep = np.zeros(N, dtype=bool)
    keep[ok] = u[ok] <= win[idx_r[ok] - 1, pix[ok]] * (1.0 / win.max())
    assert np.array_equal(kept, rtp[:, keep])


def test_clnn_modes_order_and_index():
    am = S.AnlmModes.from_nl(3, 2, 500.0, 1000.0, nside=8)
    cm = S.ClnnModes(am)
    # sorted by l, then n2 - n1, then n1 (sort_lnn)
    assert cm.lnn[:6].tolist() == [[0, 1, 1], [0, 2, 2], [0, 3, 3], [0, 1, 2], [0, 2, 3], [0, 1, 3]]
    for i in range(1, S.getlnnsize(cm) + 1):
        l, n1, n2 = S.getlnn(cm, i)
        assert C.getidx(cm, l, n1, n2) == i == C.getidx(cm, l, n2, n1)
        # ClnnModes{true} closed form (src/modes.jl:448-460)
        idx = int(np.flatnonzero(cm.lnn[:, 0] == l)[0]) + 1
        for dn in range(1, n2 - n1 + 1):
            idx += int(am.nmax_l[l]) - dn + 1
        assert idx + n1 - 1 == i
    assert S.ClnnModes(am, Δnmax=0).lnn[:, 1].tolist() == S.ClnnModes(am, Δnmax=0).lnn[:, 2].tolist()


@pytest.mark.gpu
def test_amln2clnn():
    am = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    cm = S.ClnnModes(am, Δnmax=2)
    nlm = S.getnlmsize(am)
    rng = np.random.default_rng(52)
    ph = rng.uniform(0, 2 * np.pi, nlm)
    a = np.exp(1j * ph)
    for n in range(1, am.nmax + 1):
        for l in range(int(am.lmax_n[n - 1]) + 1):
            a[S.getidx(am, n, l, 0) - 1] = np.sign(np.cos(ph[S.getidx(am, n, l, 0) - 1])) or 1.0
    c = S.amln2clnn(a, cm)
    diag = cm.lnn[:, 1] == cm.lnn[:, 2]
    assert np.allclose(c[diag], 1.0, rtol=0, atol=1e-14)
    a1 = rng.standard_normal(nlm) + 1j * rng.standard_normal(nlm)
    a2 = rng.standard_normal(nlm) + 1j * rng.standard_normal(nlm)
    c12 = S.amln2clnn(a1, a2, cm)
    for i in range(0, S.getlnnsize(cm), 7):
        l, n1, n2 = S.getlnn(cm, i + 1)
        x = a1[S.getidx(am, n1, l, 0) - 1:S.getidx(am, n1, l, l)]
        y = a2[S.getidx(am, n2, l, 0) - 1:S.getidx(am, n2, l, l)]
        expect = (np.real(x[0] * y[0]) + 2 * np.sum(np.real(x[1:] * np.conj(y[1:])))) / (2 * l + 1)
        assert c12[i] == pytest.approx(expect, rel=1e-13, abs=1e-15)


def test_clnn_modes_cross():
    """ClnnModes over two AnlmModes (src/modes.jl:396-399): every (n1 of A, n2 of B) for
    l <= min(lmax_A, lmax_B), sorted by l, n2 - n1, n1; n1 and n2 are not swapped."""
    amA = S.AnlmModes.from_nl(3, 2, 500.0, 1000.0, nside=8)
    amB = S.AnlmModes.from_nl(2, 3, 400.0, 1000.0, nside=8)
    cm = S.ClnnModes(amA, amB)
    assert not cm.symmetric and not hasattr(cm, "amodes")
    lmax = min(amA.lmax, amB.lmax)
    assert S.getlnnsize(cm) == sum(int(amA.nmax_l[l]) * int(amB.nmax_l[l]) for l in range(lmax + 1))
    keys = [(l, n2 - n1, n1) for l, n1, n2 in cm.lnn.tolist()]
    assert keys == sorted(keys) and len(set(map(tuple, cm.lnn.tolist()))) == len(keys)
    for i in range(1, S.getlnnsize(cm) + 1):
        l, n1, n2 = S.getlnn(cm, i)
        assert C.getidx(cm, l, n1, n2) == i
    assert (0, 2, 1) in map(tuple, cm.lnn.tolist()) and (0, 1, 2) in map(tuple, cm.lnn.tolist())
    # over one AnlmModes given twice: every pair, of which the symmetric ClnnModes holds n1 <= n2
    both = set(map(tuple, S.ClnnModes(amA, amA).lnn.tolist()))
    assert set(map(tuple, S.ClnnModes(amA).lnn.tolist())) == {t for t in both if t[1] <= t[2]}
    dk = 0.004
    cmk = S.ClnnModes(amA, amB, Δkmax=dk)
    for l, n1, n2 in cmk.lnn.tolist():
        assert abs(amB.knl[n2 - 1, l] - amA.knl[n1 - 1, l]) <= dk


@pytest.mark.gpu
def test_amln2clnn_cross():
    """Cross spectra between two mode sets against the reference's alm2cl over getidx of each
    (src/SphericalFourierBesselDecompositions.jl:114-150), and the cross form over one
    AnlmModes equal to the symmetric one on its entries."""
    amA = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    amB = S.AnlmModes(0.04, 400.0, 1000.0, nside=32)
    rng = np.random.default_rng(53)
    a1 = rng.normal(size=S.getnlmsize(amA)) + 1j * rng.normal(size=S.getnlmsize(amA))
    a2 = rng.normal(size=S.getnlmsize(amB)) + 1j * rng.normal(size=S.getnlmsize(amB))
    cm = S.ClnnModes(amA, amB, Δnmax=3)
    c = S.amln2clnn(a1, a2, cm)
    ref = np.empty_like(c)
    for i, (l, n1, n2) in enumerate(cm.lnn.tolist()):
        x = a1[[S.getidx(amA, n1, l, m) - 1 for m in range(l + 1)]]
        y = a2[[S.getidx(amB, n2, l, m) - 1 for m in range(l + 1)]]
        ref[i] = (np.real(x[0] * y[0]) + 2 * np.sum(np.real(x[1:] * np.conj(y[1:])))) / (2 * l + 1)
    assert np.abs(c - ref).max() <= 1e-13 * np.abs(ref).max()
    cs, cx = S.ClnnModes(amA), S.ClnnModes(amA, amA)
    b = rng.normal(size=S.getnlmsize(amA)) + 1j * rng.normal(size=S.getnlmsize(amA))
    vs = dict(zip(map(tuple, cs.lnn.tolist()), S.amln2clnn(a1, b, cs)))
    vx = dict(zip(map(tuple, cx.lnn.tolist()), S.amln2clnn(a1, b, cx)))
    assert all(vs[t] == vx[t] for t in vs)
