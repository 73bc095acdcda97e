"""apo_nb_calc_force_noalloc against apo_nb_calc_force through the C ABI:
the same gradient, energies and virial bit for bit, on two calls in a row
(the 3,000-atom water box, then a smaller 1,029-atom box that reuses the
shared static partials array), with and without a list rebuild; sizes beyond
the array and a system whose list was never built are refused."""
import ctypes

import numpy as np
import pytest

from apocharmm_amd import _lib, systems

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _bufs(fm):
    dev = torch.device("cuda")
    return (torch.zeros(3 * fm.stride, dtype=torch.int64, device=dev),
            torch.zeros(_lib.NUM_ENERGY, dtype=torch.float64, device=dev),
            torch.zeros(9, dtype=torch.float64, device=dev))


def _call(entry, fm, x, flags):
    f, e, v = _bufs(fm)
    # fill with garbage: every output word must be written by the call
    f.fill_(123456789)
    e.fill_(7.0)
    v.fill_(-3.0)
    _lib.check(entry(fm._handle, ctypes.c_void_p(x.data_ptr()), flags,
                     ctypes.c_void_p(f.data_ptr()), fm.stride, ctypes.c_void_p(e.data_ptr()),
                     ctypes.c_void_p(v.data_ptr()), None))
    torch.cuda.synchronize()
    return f.cpu().numpy(), e.cpu().numpy(), v.cpu().numpy()


def test_noalloc_equals_allocating_entry():
    from apocharmm_amd.force_manager import ForceManager
    L = _lib.lib()
    L.apo_nb_calc_force_noalloc.argtypes = L.apo_nb_calc_force.argtypes
    L.apo_nb_calc_force_noalloc.restype = ctypes.c_int
    big = systems.config1()
    small = systems.tip3p_box(343, 21.7, 24, seed=7, cutnb=10.0, ctofnb=9.0, ctonnb=8.0)
    EV = _lib.APO_CALC_ENERGY | _lib.APO_CALC_VIRIAL
    for name, sp in (("3000 atoms", big), ("1029 atoms", small)):
        fm = ForceManager(sp).initialize()
        x = torch.tensor(systems.xyzq_of(sp), device="cuda")
        ref = _call(L.apo_nb_calc_force, fm, x, EV | _lib.APO_RESET_LIST)
        for flags in (EV | _lib.APO_RESET_LIST, EV, _lib.APO_CALC_ENERGY):
            got = _call(L.apo_nb_calc_force_noalloc, fm, x, flags)
            want = _call(L.apo_nb_calc_force, fm, x, flags)
            print("[noalloc %s flags %d] epot %.6f (allocating entry %.6f)" % (
                name, flags, got[1][-1], want[1][-1]), flush=True)
            assert np.array_equal(got[0], want[0])
            assert np.array_equal(got[0], ref[0])
            assert np.array_equal(got[1], want[1])
            if flags & _lib.APO_CALC_VIRIAL:
                assert np.array_equal(got[2], want[2])
        fm.close()


def test_noalloc_refuses_what_it_cannot_hold():
    from apocharmm_amd.force_manager import ForceManager
    L = _lib.lib()
    L.apo_nb_calc_force_noalloc.argtypes = L.apo_nb_calc_force.argtypes
    L.apo_nb_calc_force_noalloc.restype = ctypes.c_int
    sp = systems.tip3p_box(343, 21.7, 24, seed=7, cutnb=10.0, ctofnb=9.0, ctonnb=8.0)
    fm = ForceManager(sp).initialize()
    x = torch.tensor(systems.xyzq_of(sp), device="cuda")
    f, e, v = _bufs(fm)
    args = (ctypes.c_void_p(x.data_ptr()), _lib.APO_CALC_ENERGY, ctypes.c_void_p(f.data_ptr()),
            fm.stride, ctypes.c_void_p(e.data_ptr()), None, None)
    # no list built yet: refused before any work is queued
    assert L.apo_nb_calc_force_noalloc(fm._handle, *args) == _lib.APO_INVALID_ARGUMENT
    assert b"no capacity" in L.apo_last_error()
    fm.close()
    # a 200^3 PME grid needs 200 * 101 * 7 reciprocal partials: beyond the
    # array, refused before anything runs on the device
    sp = dict(sp)
    sp["fft_grid"] = np.array([200, 200, 200], np.int32)
    fm = ForceManager(sp).initialize()
    assert L.apo_nb_calc_force_noalloc(fm._handle, *args) == _lib.APO_INVALID_ARGUMENT
    assert b"holds at most" in L.apo_last_error()
    fm.close()
