"""CHARMM PSF / CRD / parameter readers against the reference's own test data.

Pins: the atom, bond and angle counts asserted by unittest-psf.cpp (waterbox
:92-100), its per-atom golden files (charges, masses, types, names) and its
1-4 exclusion golden files (iblo14 / inb14, compared there with
compareAbsoluteFromFile); the c47b1 fixture the force tests use
(tests/golden/waterbox.npz, converted independently by make_fixtures.py);
the literal parameter lines of the files for the JAC DHFR system; and that the
oracle evaluates the JAC system with finite energies whose terms the force
field fixes in sign.
"""
import gzip
import os

import numpy as np
import pytest

import golden_util
from apocharmm_amd import charmm_io as C

D = os.path.join(golden_util.GOLDEN, "charmm")


def f(name):
    return os.path.join(D, name + ".gz")


def txt(name):
    with gzip.open(f(name), "rt") as h:
        return [l.strip() for l in h.read().splitlines() if l.strip()]


@pytest.fixture(scope="module")
def waterbox():
    psf = C.CharmmPSF(f("waterbox.psf"))
    prm = C.CharmmParameters([f("toppar_water_ions.str")])
    return psf, prm


def test_waterbox_psf_counts_and_atoms(waterbox):
    psf, _ = waterbox
    assert psf.getNumAtoms() == 11748
    assert psf.getNumBonds() == 11748
    assert psf.getNumAngles() == 3916
    assert psf.getNumDihedrals() == 0 and psf.getNumImpropers() == 0
    assert np.allclose(psf.getCharges(), [float(v) for v in txt("waterbox_atomCharges.txt")])
    assert np.allclose(psf.getMasses(), [float(v) for v in txt("waterbox_atomMasses.txt")])
    assert psf.getAtomTypes() == txt("waterbox_atomTypes.txt")
    assert psf.getAtomNames() == txt("waterbox_atomNames.txt")
    assert len(psf.getWaterMolecules()) == 3916


def _iblo14_inb14(n, bonds):
    """The reference's iblo14/inb14 layout (src/CharmmPSF.cu:363-382) from the
    exclusion derivation the product uses (apo_topology_exclusions)."""
    from apocharmm_amd import _lib
    import ctypes
    L = _lib.lib()
    b = np.ascontiguousarray(bonds, np.int32)
    counts = (ctypes.c_int32 * 3)()
    _lib.check(L.apo_topology_exclusions(n, len(b), b.ctypes.data_as(ctypes.c_void_p), counts,
                                         None, None, None, None))
    ptr = np.zeros(n + 1, np.int32)
    idx = np.zeros(counts[0], np.int32)
    in14 = np.zeros(2 * counts[1], np.int32)
    ex14 = np.zeros(2 * counts[2], np.int32)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    _lib.check(L.apo_topology_exclusions(n, len(b), vp(b), counts, vp(ptr), vp(idx), vp(in14),
                                         vp(ex14)))
    partners = [set() for _ in range(n)]
    for i in range(n):
        for j in idx[ptr[i]:ptr[i + 1]]:
            partners[i].add(int(j))
    for a, c in in14.reshape(-1, 2):
        partners[a].add(int(c))
        partners[c].add(int(a))
    iblo, inb = [], []
    for i in range(n):
        inb += [j + 1 for j in sorted(partners[i]) if j > i]
        iblo.append(len(inb))
    return iblo, inb


def test_waterbox_1_4_exclusion_lists_match_golden(waterbox):
    psf, _ = waterbox
    iblo, inb = _iblo14_inb14(psf.getNumAtoms(), psf.getBonds())
    assert iblo == [int(v) for v in txt("waterbox_iblo14.txt")]
    assert inb == [int(v) for v in txt("waterbox_inb14.txt")]


def test_waterbox_spec_equals_c47b1_fixture(waterbox):
    psf, prm = waterbox
    crd = C.CharmmCrd(f("waterbox.crd"))
    sp = C.charmm_system(psf, prm, crd, (50.0, 50.0, 50.0), 48, cutnb=10.0, ctofnb=8.0,
                         ctonnb=7.0)
    ref, xyzq, _ = golden_util.load("waterbox")
    assert np.array_equal(sp["xyz"], xyzq[:, :3])
    assert np.allclose(sp["charges"], ref["charges"])
    assert np.array_equal(sp["vdwtype"], ref["vdwtype"])
    assert np.allclose(sp["eps"], ref["eps"]) and np.allclose(sp["rmin_half"], ref["rmin_half"])
    assert np.array_equal(sp["bonds"], ref["bonds"])
    assert np.allclose(sp["bond_par"], ref["bond_par"])
    assert np.array_equal(sp["angles"], ref["angles"])
    assert np.allclose(sp["angle_par"], ref["angle_par"])
    # no NBFIX here: the explicit tables equal the combination rule
    from apocharmm_amd.systems import lj_table
    assert np.array_equal(sp["vdwtab"], lj_table(ref["eps"], ref["rmin_half"]))


def test_small_fixtures_read():
    d = C.charmm_system(f("waterDimer.psf"), [f("toppar_water_ions.str")], f("waterDimer.crd"),
                        (520.0,) * 3, 48)
    assert len(d["charges"]) == 6 and len(d["waters"]) == 2
    a = C.charmm_system(f("argon_10.psf"), [f("argon.prm")], f("argon_10.crd"), (50.0,) * 3, 48)
    assert len(a["charges"]) == 10 and len(a["bonds"]) == 0
    assert a["eps"].shape == (1,)


@pytest.fixture(scope="module")
def jac():
    psf = C.CharmmPSF(f("jac_5dhfr.psf"))
    prm = C.CharmmParameters([f("par_all22_prot.prm"), f("toppar_water_ions.str")])
    return psf, prm


def test_jac_counts_and_parameter_lines(jac):
    psf, prm = jac
    # the PSF's own section headers
    assert (psf.getNumAtoms(), psf.getNumBonds(), psf.getNumAngles(), psf.getNumDihedrals(),
            psf.getNumImpropers(), psf.getNumCrossTerms()) == (23558, 23592, 11584, 6701, 418, 0)
    assert len(psf.getWaterMolecules()) == 7023
    assert abs(psf.getCharges().sum() + 11.0) < 1e-9
    # literal lines of par_all22_prot.prm
    assert prm.bonds[("C", "CT1")] == (1.49, 250.0)
    # "H NH1 CT1 35.000 117.0000": outer types ordered, no Urey-Bradley term
    th0, k = prm.angles[("CT1", "NH1", "H")]
    assert k == 35.0 and abs(th0 - np.deg2rad(117.0)) < 1e-12
    assert prm.ureyb[("CT1", "NH1", "H")] == (0.0, 0.0)
    # "HA CT1 C 33.000 109.50 30.00 2.16300": with a Urey-Bradley term
    th0, k = prm.angles[("C", "CT1", "HA")]
    assert k == 33.0 and abs(th0 - np.deg2rad(109.5)) < 1e-12
    assert prm.ureyb[("C", "CT1", "HA")] == (2.163, 30.0)
    terms = prm._dihedral_terms("H", "NH1", "CT1", "C")
    assert all(t[0] >= 1 for t in terms)
    assert prm.vdw["CT1"] == (-0.02, 2.275) and prm.vdw14["CT1"] == (-0.01, 1.9)
    assert prm.vdw["OT"] == (-0.1521, 1.7682)


def test_jac_system_builds_and_oracle_is_finite(jac):
    from oracle import oracle as O
    psf, prm = jac
    sp = C.charmm_system(psf, prm, f("jac_5dhfr.crd"), (62.23,) * 3, 64)
    n = len(sp["charges"])
    assert len(sp["dihe"]) >= psf.getNumDihedrals()          # multi-term series add rows
    assert np.all(sp["dihe_par"][:, 0] >= 1)
    assert len(sp["imp"]) == 418 and np.all(sp["imp_par"][:, 0] == 0)
    assert len(sp["ureyb"]) > 0 and np.all(sp["ub_par"][:, 1] > 0.01)
    # 1-4 table uses the 14 parameters where the file gives them
    t = sp["type_names"].index("CT1")
    row = t * (t + 1) // 2 + t
    assert sp["vdwtab14"][2 * row + 1] != sp["vdwtab"][2 * row + 1]
    from apocharmm_amd import systems
    r = O.evaluate(O.prepare(sp), systems.xyzq_of(sp), nthreads=8)
    e = r["energy"]
    assert all(np.isfinite(v) for v in e.values())
    assert e["bond"] > 0 and e["angle"] > 0 and e["ureyb"] >= 0 and e["dihe"] > 0
    assert e["ewse"] < 0 and e["elec"] < 0
    assert np.all(np.isfinite(r["grad"])) and r["grad"].shape == (n, 3)
