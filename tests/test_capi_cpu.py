"""C ABI checks that need no GPU: the library loads, exports every entry point
declared in include/*.h, validates arguments before touching the
device, and derives the same exclusion topology as the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

from apocharmm_amd import _lib, systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    inc = os.path.join(ROOT, "include")
    hdr = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc))
                  if f.endswith(".h"))
    return sorted(set(re.findall(r"APO_API\s+[^;(]*?\b(apo_\w+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n


def test_version_string():
    assert "gfx950" in _lib.version()


def _exclusions_capi(n, bonds):
    L = _lib.lib()
    b = np.ascontiguousarray(np.asarray(bonds, np.int32).reshape(-1))
    counts = (ctypes.c_int32 * 3)()
    _lib.check(L.apo_topology_exclusions(n, len(b) // 2, b.ctypes.data, counts, None, None, None, None))
    ptr = np.zeros(n + 1, np.int32)
    idx = np.zeros(max(counts[0], 1), np.int32)
    in14 = np.zeros(max(2 * counts[1], 1), np.int32)
    ex14 = np.zeros(max(2 * counts[2], 1), np.int32)
    _lib.check(L.apo_topology_exclusions(n, len(b) // 2, b.ctypes.data, counts, ptr.ctypes.data,
                                         idx.ctypes.data, in14.ctypes.data, ex14.ctypes.data))
    return ptr, idx[: counts[0]], in14[: 2 * counts[1]].reshape(-1, 2), ex14[: 2 * counts[2]].reshape(-1, 2)


@pytest.mark.parametrize("which", ["water", "chain"])
def test_exclusion_topology_matches_oracle(oracle, which):
    if which == "water":
        spec = systems.tip3p_box(27, 9.3, 8, seed=1)
    else:
        spec = systems.solvated_chain(n_chain=60, n_waters=40, L=24.0, fft_grid=24)
    n = len(spec["charges"])
    got = _exclusions_capi(n, spec["bonds"])
    want = oracle.exclusion_lists(n, spec["bonds"])
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    if which == "chain":
        assert len(got[2]) == 57  # one 1-4 pair per chain dihedral


def test_invalid_arguments_rejected_before_device_use():
    L = _lib.lib()
    p = _lib.NbParams()
    t = _lib.NbTopology()
    h = ctypes.c_void_p()
    p.box[0] = p.box[1] = p.box[2] = 20.0
    p.kappa, p.cutnb, p.ctofnb, p.ctonnb = 0.34, 14.0, 12.0, 10.0   # ctofnb > box/2
    t.natom = 3
    st = L.apo_nb_create(ctypes.byref(p), ctypes.byref(t), ctypes.byref(h))
    assert st == _lib.APO_INVALID_ARGUMENT
    assert b"half the box" in L.apo_last_error()
    assert L.apo_nb_calc_force(None, None, 0, None, 0, None, None, None) == _lib.APO_INVALID_ARGUMENT
    assert b"null" in L.apo_last_error()
