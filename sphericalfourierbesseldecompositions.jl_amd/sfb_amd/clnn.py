"""ClnnModes and amln2clnn: the pseudo-power-spectrum consumer of a_nlm.

ClnnModes (src/modes.jl:280-447): the (l, n1, n2) triples with |n2 - n1| <= Δnmax and
|k_n2l - k_n1l| <= Δkmax, sorted by l, then Δn = n2 - n1, then n1 (sort_lnn, :338-355).
Over one AnlmModes (ClnnModes{true}, :390-393) only n1 <= n2; over two (ClnnModes{false},
:396-399) every (n1 of A, n2 of B) for l <= min(lmax_A, lmax_B) (calc_lnn, :358-387).
amln2clnn (src/SphericalFourierBesselDecompositions.jl:114-150) runs on the device.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .modes import AnlmModes


class ClnnModes:
    """ClnnModes(amodes; Δkmax, Δnmax) or ClnnModes(amodesA, amodesB; Δkmax, Δnmax)."""

    def __init__(self, amodesA: AnlmModes, amodesB: AnlmModes | None = None, *, Δkmax=np.inf,
                 Δnmax=np.iinfo(np.int64).max):
        self.symmetric = amodesB is None
        amodesB = amodesA if amodesB is None else amodesB
        self.amodesA, self.amodesB = amodesA, amodesB
        if self.symmetric:
            self.amodes = amodesA  # the reference's cmodes.amodes of ClnnModes{true}
        lnn = []
        dk_out, dn_out = -np.inf, 0
        for l in range(min(amodesA.lmax, amodesB.lmax) + 1):
            na, nb = int(amodesA.nmax_l[l]), int(amodesB.nmax_l[l])
            for n1 in range(1, na + 1):
                for n2 in range(n1 if self.symmetric else 1, nb + 1):
                    dk = amodesB.knl[n2 - 1, l] - amodesA.knl[n1 - 1, l]
                    if abs(dk) <= Δkmax and abs(n2 - n1) <= Δnmax:
                        lnn.append((l, n1, n2))
                        dk_out = max(abs(dk), dk_out)
                        dn_out = max(abs(n2 - n1), dn_out)
        lnn.sort(key=lambda t: (t[0], t[2] - t[1], t[1]))
        self.lnn = np.array(lnn, dtype=np.int64).reshape(-1, 3)
        self.Δkmax, self.Δnmax = dk_out, dn_out
        self._index = {tuple(t): i + 1 for i, t in enumerate(lnn)}

    def getlnnsize(self):
        return self.lnn.shape[0]


def getlnnsize(cmodes: ClnnModes):
    return cmodes.getlnnsize()


def getlnn(cmodes: ClnnModes, idx):
    """(l, n1, n2) of the 1-based index idx."""
    l, n1, n2 = cmodes.lnn[idx - 1]
    return int(l), int(n1), int(n2)


def getidx(cmodes: ClnnModes, l, n1, n2):
    """1-based index of (l, n1, n2); over one AnlmModes n1 and n2 are swapped when n1 > n2."""
    if cmodes.symmetric and n1 > n2:
        n1, n2 = n2, n1
    try:
        return cmodes._index[(int(l), int(n1), int(n2))]
    except KeyError:
        raise KeyError(f"Cannot find index of (l, n1, n2) = {(l, n1, n2)}") from None


def amln2clnn(anlm1, *args, device=0):
    """amln2clnn(anlm, cmodes) or amln2clnn(anlm1, anlm2, cmodes) on the device; anlm1 in the
    order of cmodes.amodesA, anlm2 in that of cmodes.amodesB."""
    from .cat2anlm import get_plan
    if len(args) == 1:
        anlm2, cmodes = anlm1, args[0]
    else:
        anlm2, cmodes = args
    pa = get_plan(cmodes.amodesA, device=device)
    pb = pa if cmodes.amodesB is cmodes.amodesA else get_plan(cmodes.amodesB, device=device)

    def inter(a, plan):
        a = np.asarray(a, dtype=np.complex128).reshape(-1)
        if a.size != plan.nlmsize:
            raise ValueError(f"a_nlm must have {plan.nlmsize} entries")
        return L.f64(np.stack([a.real, a.imag], axis=1).reshape(-1))

    a1, a2 = inter(anlm1, pa), inter(anlm2, pb)
    lnn = np.ascontiguousarray(cmodes.lnn, dtype=np.int64)
    out = np.empty(lnn.shape[0])
    L.check(L.lib().sfb_amln2clnn_cross(pa.handle, pb.handle, lnn.shape[0], L.ptr(lnn), L.ptr(a1), L.ptr(a2),
                                        L.ptr(out)))
    return out


anlm2clnn = amln2clnn


def healpix_pixwin(nside, lmax=None, *, depth=0, device=0):
    """The HEALPix pixel window w_l, l = 0..lmax (default 4 nside): what Healpix.pixwin(nside)
    returns from HEALPix's tables, here computed on the device from its definition
    (w_l^2 = mean over pixels of (4 pi/(2l+1)) sum_m |<Y_lm>_pixel|^2, pixel means over the
    4^depth sub-pixel centres, extrapolated from depths depth-1 and depth; depth 0 picks 5 up
    to nside 32, 4 up to nside 256 and 3 above)."""
    nside = int(nside)
    lmax = 4 * nside if lmax is None else int(lmax)
    out = np.empty(lmax + 1)
    L.check(L.lib().sfb_pixwin(nside, lmax, int(depth), int(device), L.ptr(out)))
    return out


def pixwin(cmodes: ClnnModes, *, depth=0, device=0):
    """pixwin(cmodes) (src/SphericalFourierBesselDecompositions.jl:158-167): the angular pixel
    window w_l of every (l, n1, n2) entry of a C_l(n, n') vector."""
    w = healpix_pixwin(cmodes.amodesA.nside, int(cmodes.lnn[:, 0].max()) if len(cmodes.lnn) else 0,
                       depth=depth, device=device)
    return w[cmodes.lnn[:, 0]]
