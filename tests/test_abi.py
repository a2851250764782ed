"""The C ABI: every function include/sfb.h declares is exported by the library
(no GPU needed), and the argument checks that precede any device work return the
documented status codes."""
import ctypes
import os
import re

import numpy as np
import pytest

import sfb_amd as S
from sfb_amd import _lib as L

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "sfb.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(sfb_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_header_symbols_exported_and_bound():
    names = declared_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(L.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), f"{n} declared in sfb.h but not exported"
        assert n in L.SIGNATURES, f"{n} has no ctypes binding"
    assert L.lib().sfb_version().decode().startswith("sfb_mi355x")


def plan_args(lmax_n=(3,), nside=2, lmax=3):
    am = S.AnlmModes.from_nl(1, 3, 500.0, 1000.0, nside=8)
    return am


def test_plan_rejects_lmax_above_4_nside():
    am = S.AnlmModes.from_nl(1, 9, 500.0, 1000.0, nside=2)
    h = ctypes.c_void_p()
    rc = L.lib().sfb_plan_create(ctypes.byref(h), 2, am.lmax, am.nmax, L.ptr(am.lmax_n), L.ptr(am.knot_off),
                                 L.ptr(am.knot_x), L.ptr(am.knot_y), L.ptr(am.knot_ypp), 3, 0, 0)
    assert rc == L.SFB_ERR_LMAX
    assert "4*nside" in L.lib().sfb_last_error().decode()
    with pytest.raises(S.SFBError) as e:
        S.Plan(am)
    assert e.value.code == L.SFB_ERR_LMAX


def test_plan_rejects_bad_arguments():
    am = S.AnlmModes.from_nl(1, 3, 500.0, 1000.0, nside=8)
    lib = L.lib()
    h = ctypes.c_void_p()
    args = [8, am.lmax, am.nmax, L.ptr(am.lmax_n), L.ptr(am.knot_off), L.ptr(am.knot_x), L.ptr(am.knot_y),
            L.ptr(am.knot_ypp), 3, 0, 0]
    assert lib.sfb_plan_create(None, *args) == L.SFB_ERR_ARG
    bad = list(args)
    bad[0] = 12  # nside not a power of two
    assert lib.sfb_plan_create(ctypes.byref(h), *bad) == L.SFB_ERR_ARG
    bad = list(args)
    bad[9] = 7  # unknown precision
    assert lib.sfb_plan_create(ctypes.byref(h), *bad) == L.SFB_ERR_ARG
    bad = list(args)
    bad[3] = None  # lmax_n
    assert lib.sfb_plan_create(ctypes.byref(h), *bad) == L.SFB_ERR_ARG
    assert lib.sfb_cat2amln(None, 0, None, None, 1.0, 1, None, None, None) == L.SFB_ERR_ARG


def test_modes_and_spline_argument_checks():
    lib = L.lib()
    h = ctypes.c_void_p()
    assert lib.sfb_modes_create_kmax(ctypes.byref(h), 0.05, 1000.0, 500.0, 0) == L.SFB_ERR_ARG
    assert lib.sfb_modes_create_nl(ctypes.byref(h), 0, 3, 0.0, 1.0, 0) == L.SFB_ERR_ARG
    x = np.arange(3.0)
    assert lib.sfb_spline_build(3, L.ptr(x), L.ptr(x), L.ptr(x.copy())) == L.SFB_ERR_ARG
    with pytest.raises(S.SFBError):
        S.Spline1D([0.0, 1.0, 2.0], [1.0, 2.0, 3.0])


def test_window_shape_checks():
    am = S.AnlmModes.from_nl(1, 3, 500.0, 1000.0, nside=8)
    p = S.Plan.__new__(S.Plan)
    p.amodes, p.npix = am, 12 * 64
    with pytest.raises(ValueError):
        p._window(np.zeros((10, 4, 1)))
    with pytest.raises(ValueError):
        p._window(S.SeparableArray(np.ones(5), np.ones((4, 1))))
    with pytest.raises(ValueError):
        p._catalogue(np.zeros((2, 5)), None)
