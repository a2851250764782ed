"""Error paths forced from the host (SFB_FAULT): the root finder's maxevals exhaustion
(the reference's doublesecant_method returns NaN after maxevals and AnlmModes fails,
src/SphericalBesselGNLs.jl:126-127) gives SFB_ERR_SETUP on the CPU; a device
allocation failure gives SFB_ERR_DEVICE at plan creation and at a call that must grow a
buffer, and the plan stays usable once the fault is lifted (gpu)."""
import ctypes
import os

import numpy as np
import pytest

import sfb_amd as S
from sfb_amd import _lib as L


@pytest.fixture
def fault(monkeypatch):
    def set_(what):
        monkeypatch.setenv("SFB_FAULT", what)
    yield set_
    monkeypatch.delenv("SFB_FAULT", raising=False)


def test_root_finder_exhaustion_is_setup_error(fault):
    fault("maxevals")
    h = ctypes.c_void_p()
    rc = L.lib().sfb_modes_create_kmax(ctypes.byref(h), 0.05, 500.0, 1000.0, 32)
    assert rc == L.SFB_ERR_SETUP
    assert "root finding failed" in L.lib().sfb_last_error().decode()
    with pytest.raises(S.SFBError) as e:
        S.AnlmModes(0.05, 500.0, 1000.0, nside=32)
    assert e.value.code == L.SFB_ERR_SETUP
    os.environ.pop("SFB_FAULT")
    am = S.AnlmModes(0.05, 500.0, 1000.0, nside=32)  # lifted: the same call succeeds
    assert am.nmaps == 237


def test_unknown_fault_value_changes_nothing(fault):
    fault("nothing-of-that-name")
    assert S.AnlmModes(0.05, 500.0, 1000.0, nside=32).nmaps == 237


@pytest.mark.gpu
def test_device_allocation_failure(fault):
    from helpers import separable_unit_window, shell_volume, uniform_shell_catalog
    am = S.AnlmModes(0.03, 500.0, 1000.0, nside=16)
    win, _ = separable_unit_window(S, am, nr=100)
    fault("alloc")
    with pytest.raises(S.SFBError) as e:
        S.Plan(am)
    assert e.value.code == L.SFB_ERR_DEVICE and "SFB_FAULT" in str(e.value)
    os.environ.pop("SFB_FAULT")
    p = S.Plan(am)
    small = uniform_shell_catalog(100, 500.0, 1000.0, seed=1)
    big = uniform_shell_catalog(20000, 500.0, 1000.0, seed=2)
    nbar = 1e-4
    a0 = p.cat2amln(small, nbar, win)
    fault("alloc")
    with pytest.raises(S.SFBError) as e:  # the catalogue buffers must grow: allocation fails
        p.cat2amln(big, nbar, win)
    assert e.value.code == L.SFB_ERR_DEVICE
    os.environ.pop("SFB_FAULT")
    assert np.array_equal(p.cat2amln(small, nbar, win), a0)  # the plan is still usable
    ref = S.Plan(am).cat2amln(big, nbar, win)
    assert np.array_equal(p.cat2amln(big, nbar, win), ref)
