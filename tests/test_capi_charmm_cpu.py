"""The reference's existing C ABI (include/apocharmm_amd_capi.h) without a
GPU: PSF, CRD and parameter handles read the reference's own test files
(test/data, committed under tests/golden/charmm/) and return what the Python
readers return (which the c47b1 fixtures and unittest-psf.cpp's counts pin);
the force manager is created and configured; errors follow the reference's
status / apo_last_error conventions."""
import gzip
import os
import shutil

import numpy as np
import pytest

import golden_util
from apocharmm_amd import _lib
from apocharmm_amd import apocharmm_c as A
from apocharmm_amd import charmm_io as C

D = os.path.join(golden_util.GOLDEN, "charmm")


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    """The reference's files, uncompressed (the C readers take plain text)."""
    out = tmp_path_factory.mktemp("charmm")
    paths = {}
    for name in ("waterbox.psf", "waterbox.crd", "toppar_water_ions.str", "jac_5dhfr.psf",
                 "jac_5dhfr.crd", "par_all22_prot.prm"):
        dst = out / name
        with gzip.open(os.path.join(D, name + ".gz"), "rb") as fi, open(dst, "wb") as fo:
            shutil.copyfileobj(fi, fo)
        paths[name] = str(dst)
    return paths


@pytest.mark.parametrize("name", ["waterbox", "jac_5dhfr"])
def test_psf_handle_matches_python_reader(files, name):
    path = files[name + ".psf"]
    ref = C.CharmmPSF(path)
    with A.CharmmPsf(path) as psf:
        assert psf.getNumAtoms() == ref.getNumAtoms()
        assert psf.getNumBonds() == ref.getNumBonds()
        assert psf.getNumAngles() == ref.getNumAngles()
        assert psf.getNumDihedrals() == ref.getNumDihedrals()
        assert psf.getNumImpropers() == ref.getNumImpropers()
        assert psf.getNumCrossTerms() == ref.getNumCrossTerms()
        assert psf.getAtomNames() == ref.getAtomNames()
        assert psf.getAtomTypes() == [t for t in ref.getAtomTypes()]
        assert psf.getResidueNames() == ref.getResidueNames()
        assert psf.getSegmentIdentifiers() == ref.getSegmentIdentifiers()
        assert psf.getResidueIdentifiers() == ref.getResidueIdentifiers()
        np.testing.assert_array_equal(psf.getCharges(), ref.getCharges())
        np.testing.assert_array_equal(psf.getMasses(), ref.getMasses())
        assert psf.getNetCharge() == pytest.approx(ref.getCharges().sum(), abs=1e-9)
        assert psf.getTotalMass() == pytest.approx(ref.getMasses().sum(), rel=1e-12)
        assert psf.getFileName() == path


def test_waterbox_counts_of_unittest_psf(files):
    """The counts test/unittests/unittest-psf.cpp asserts for waterbox.psf."""
    with A.CharmmPsf(files["waterbox.psf"]) as psf:
        assert psf.getNumAtoms() == 11748
        assert psf.getNumBonds() == 11748
        assert psf.getNumAngles() == 3916


@pytest.mark.parametrize("name", ["waterbox", "jac_5dhfr"])
def test_crd_handle_matches_python_reader(files, name):
    path = files[name + ".crd"]
    with A.CharmmCrd(path) as crd:
        x = crd.getCoordinates()
    np.testing.assert_array_equal(x, C.CharmmCrd(path).getCoordinates())


def test_force_manager_created_from_files(files):
    psf = A.CharmmPsf(files["jac_5dhfr.psf"])
    prm = A.CharmmParameters([files["par_all22_prot.prm"], files["toppar_water_ions.str"]])
    fm = A.ForceManager(psf, prm)
    fm.setBoxDimensions((62.23, 62.23, 62.23))
    fm.setFFTGrid(64, 64, 64)
    fm.setCutoff(12.0)
    with pytest.raises(_lib.ApoInvalidArgument, match="box dimensions"):
        fm.setBoxDimensions((62.23, -1.0, 62.23))
    with pytest.raises(_lib.ApoError):  # no GPU side before a context exists
        fm.nbSystem()
    # the handles can go in any order: the manager keeps what it uses
    psf.close()
    prm.close()
    fm.close()


def test_errors_follow_the_reference_conventions(files, tmp_path):
    with pytest.raises(_lib.ApoError, match="could not open"):
        A.CharmmPsf(str(tmp_path / "missing.psf"))
    L = A._lib_c()
    import ctypes
    h = ctypes.c_void_p()
    assert L.apo_charmm_psf_create(ctypes.byref(h), None) == _lib.APO_INVALID_ARGUMENT
    assert b"NULL or empty" in L.apo_last_error()
    with A.CharmmPsf(files["waterbox.psf"]) as psf:
        buf = ctypes.create_string_buffer(8)
        assert (L.apo_charmm_psf_get_atom_names(buf, 8, psf.handle) ==
                _lib.APO_INVALID_ARGUMENT)
        assert b"too small" in L.apo_last_error()
    bad = tmp_path / "bad.prm"
    bad.write_text("NONBONDED\nOT 0.0 -0.1521\n")
    with pytest.raises(_lib.ApoError, match="NONBONDED"):
        A.CharmmParameters(str(bad))
