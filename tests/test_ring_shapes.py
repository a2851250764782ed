"""Launch shapes of the ring transforms (host arithmetic only, no GPU): every polar ring
pair of every accepted nside is covered by a kernel whose block fits in LDS, with at
least one map per block; nside above 512 is refused before any device work."""
import ctypes

import numpy as np
import pytest

import sfb_amd as S
from sfb_amd import _lib as L

LDS = 160 * 1024  # MI355X LDS per workgroup


def shapes(nside, lmax, lds=LDS):
    out = np.zeros(8, dtype=np.int64)
    rc = L.lib().sfb_ring_kernel_shapes(nside, lmax, lds, L.ptr(out))
    return rc, out


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16, 32, 64, 128, 256, 512])
@pytest.mark.parametrize("lfac", [1, 2, 3, 4])
def test_every_polar_ring_has_a_fitting_kernel(nside, lfac):
    lmax = lfac * nside
    rc, o = shapes(nside, lmax)
    assert rc == L.SFB_OK
    npolar = nside - 1
    assert o[0] == npolar                      # forward: radix-4 kernel on every polar pair
    assert o[1] <= LDS and o[2] >= 1           # ... within LDS, >= 1 map per block
    assert (o[1] > 0) == (npolar > 0)
    assert o[3] + o[4] == npolar or (o[7] == 0 and o[3] + o[4] == 2 * nside)
    assert o[3] <= 256                         # i <= 256: at most 128 output pairs per map
    if o[3] > 0:
        assert 0 < o[5] <= LDS and o[6] >= 1
    assert o[7] >= 1                           # the inverse FFT fits for nside <= 512


def test_nside_512_splits_the_inverse():
    rc, o = shapes(512, 1500)
    assert rc == L.SFB_OK
    assert (o[3], o[4]) == (256, 255)          # rings i <= 256 radix-4, i = 257..511 direct sum


def test_nside_1024_refused():
    rc, _ = shapes(1024, 3000)
    assert rc == L.SFB_ERR_ARG
    am = S.AnlmModes.from_nl(1, 3, 500.0, 1000.0, nside=8)
    h = ctypes.c_void_p()
    rc = L.lib().sfb_plan_create(ctypes.byref(h), 1024, am.lmax, am.nmax, L.ptr(am.lmax_n), L.ptr(am.knot_off),
                                 L.ptr(am.knot_x), L.ptr(am.knot_y), L.ptr(am.knot_ypp), 3, 0, 0)
    assert rc == L.SFB_ERR_ARG
    assert "512" in L.lib().sfb_last_error().decode()
