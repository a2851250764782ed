"""The oracle's HEALPix ring geometry and ang2pix, pinned without the reference.

Healpix.jl is not available, so the ring ang2pix is pinned by: every pixel centre
maps to itself; an independent formulation through the HEALPix projection
(Calabretta & Roukema 2007: (z, phi) -> plane -> rotated face grid -> ring and
in-ring index) agrees on random points away from pixel edges; ring sizes,
offsets and z follow Gorski et al. 2005; north/south mirror symmetry; phi
periodicity.
"""
import numpy as np
import pytest

from oracle import oracle as O


def ring_of_pixel(nside, pix):
    ring = np.empty_like(pix)
    for i, p in enumerate(pix):
        lo, hi = 1, 4 * nside - 1
        while lo < hi:  # first ring whose start is > p, minus one
            mid = (lo + hi + 1) // 2
            st, _, _, _ = O.ring_info(nside, mid)
            if st <= p:
                lo = mid
            else:
                hi = mid - 1
        ring[i] = lo
    return ring


def ang2pix_projection(nside, theta, phi):
    """Ring-scheme pixel via the HEALPix projection (independent of the oracle's formulas),
    for even nside.  Belt: x = phi, y = 3pi/8 z; caps: sigma = sqrt(3(1-|z|)),
    x = phi_f + (phi - phi_f) sigma, y = pi/4 (2 - sigma) sign(z), phi_f the facet centre.
    In u = x nside/(pi/4), v = (pi/2 - y) nside/(pi/4) every pixel is the diamond
    |du| + |dv| <= 1, i.e. a 2x2 square in a = u + v, b = v - u centred on odd (a, b);
    ring = (a_c + b_c)/2, and the centre is unprojected to find the in-ring index."""
    z = np.cos(theta)
    phi = np.mod(phi, 2 * np.pi)
    az = np.abs(z)
    x = phi.copy()
    y = 3 * np.pi / 8 * z
    cap = az > 2.0 / 3.0
    sigma = np.sqrt(3 * (1 - az[cap]))
    phif = (np.floor(phi[cap] / (np.pi / 2)) + 0.5) * (np.pi / 2)
    x[cap] = phif + (phi[cap] - phif) * sigma
    y[cap] = np.pi / 4 * (2 - sigma) * np.sign(z[cap])
    u = x * nside / (np.pi / 4)
    v = (np.pi / 2 - y) * nside / (np.pi / 4)
    a, b = u + v, v - u
    ac = 2 * np.floor(a / 2) + 1
    bc = 2 * np.floor(b / 2) + 1
    ring = ((ac + bc) / 2).astype(np.int64)
    xc = np.mod((ac - bc) / 2 * (np.pi / 4) / nside, 2 * np.pi)
    yc = np.pi / 2 - ring * (np.pi / 4) / nside
    out = np.empty(theta.shape, dtype=np.int64)
    for k in range(theta.size):
        st, npr, zc, phi0 = O.ring_info(nside, int(ring[k]))
        if abs(yc[k]) <= np.pi / 4 + 1e-12:
            phic = xc[k]
        else:
            sg = 2 - abs(yc[k]) / (np.pi / 4)
            pf = (np.floor(xc[k] / (np.pi / 2)) + 0.5) * (np.pi / 2)
            phic = pf + (xc[k] - pf) / sg
        j = int(np.round((phic - phi0) / (2 * np.pi / npr))) % npr
        out[k] = st + j
    return out, a, b


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16, 64])
def test_pixel_centres_roundtrip(nside):
    npix = 12 * nside * nside
    pix = np.arange(npix)
    th, ph = O.pix2ang_ring(nside, pix)
    assert np.array_equal(O.ang2pix_ring(nside, th, ph), pix)
    # centres moved by a small fraction of a pixel stay in their pixel
    eps = 0.05 * np.sqrt(4 * np.pi / npix)
    rng = np.random.default_rng(nside)
    dth = rng.uniform(-eps, eps, npix) * 0.3
    dph = rng.uniform(-eps, eps, npix) * 0.3 / np.maximum(np.sin(th), 0.3)
    assert np.array_equal(O.ang2pix_ring(nside, np.clip(th + dth, 0, np.pi), ph + dph), pix)


@pytest.mark.parametrize("nside", [1, 4, 32])
def test_ring_geometry_gorski(nside):
    npix = 12 * nside * nside
    tot = 0
    for i in range(1, 4 * nside):
        st, npr, z, phi0 = O.ring_info(nside, i)
        north = min(i, 4 * nside - i)
        assert st == tot
        if north < nside:
            assert npr == 4 * north
            assert z == pytest.approx(np.sign(2 * nside - i) * (1 - north ** 2 / (3 * nside ** 2)), abs=1e-15)
            assert phi0 == pytest.approx(np.pi / (4 * north))
        else:
            assert npr == 4 * nside
            assert z == pytest.approx(4 / 3 - 2 * i / (3 * nside), abs=1e-15)
            assert phi0 == pytest.approx(0.0 if (i - nside) % 2 else np.pi / (4 * nside))
        tot += npr
    assert tot == npix


def test_projection_agrees_away_from_edges():
    rng = np.random.default_rng(3)
    for nside in (4, 16, 32):
        n = 20000
        th = np.arccos(rng.uniform(-1, 1, n))
        ph = rng.uniform(0, 2 * np.pi, n)
        ref, a, b = ang2pix_projection(nside, th, ph)
        got = O.ang2pix_ring(nside, th, ph)
        ha, hb = a / 2, b / 2  # pixel edges where a or b is an even integer
        frac_a = np.minimum(ha - np.floor(ha), np.ceil(ha) - ha)
        frac_b = np.minimum(hb - np.floor(hb), np.ceil(hb) - hb)
        safe = (frac_a > 1e-6) & (frac_b > 1e-6)
        assert safe.mean() > 0.99
        assert np.array_equal(got[safe], ref[safe])


def test_mirror_and_periodicity():
    rng = np.random.default_rng(4)
    nside = 8
    npix = 12 * nside * nside
    th = np.arccos(rng.uniform(-1, 1, 5000))
    ph = rng.uniform(0, 2 * np.pi, 5000)
    p = O.ang2pix_ring(nside, th, ph)
    # phi periodicity and negative phi (xyz2rtp gives phi in (-pi, pi])
    assert np.array_equal(O.ang2pix_ring(nside, th, ph + 2 * np.pi), p)
    assert np.array_equal(O.ang2pix_ring(nside, th, ph - 2 * np.pi), p)
    # north/south mirror: ring i <-> ring 4 nside - i, same in-ring index
    ps = O.ang2pix_ring(nside, np.pi - th, ph)
    rn = ring_of_pixel(nside, p[:200])
    rs = ring_of_pixel(nside, ps[:200])
    assert np.array_equal(rn + rs, np.full(200, 4 * nside))
    for k in range(200):
        stn = O.ring_info(nside, int(rn[k]))[0]
        sts = O.ring_info(nside, int(rs[k]))[0]
        assert p[k] - stn == ps[k] - sts
    assert p.min() >= 0 and p.max() < npix
