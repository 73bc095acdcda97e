"""Pins the FP64 oracle to the reference's CHARMM c47b1 golden values.

Tolerances are the reference's own (test/unittests/unittest-eneForces.cpp):
waterDimer energies margin 1e-5, forces/virial 1e-4; argon10 energy Approx
(1.2e-5 rel), forces 1e-5, virial 1e-6; waterbox energies Approx, forces 0.01,
virial 1e-4*|W00|.  The .dat files hold gradients dU/dx (the reference's
"forces" are gradients), so a sign error fails here.
"""
import numpy as np
import pytest

import golden_util


@pytest.mark.parametrize("name", golden_util.NAMES)
def test_oracle_matches_c47b1(oracle, name):
    spec, xyzq, ref = golden_util.load(name)
    r = oracle.evaluate(oracle.prepare(spec), xyzq)
    for key, val in ref["energy"].items():
        tol = max(ref["energy_atol"], ref["energy_rtol"] * abs(val))
        assert abs(r["energy"][key] - val) <= tol, (key, r["energy"][key], val)
    assert np.abs(r["grad"] - ref["grad"]).max() <= ref["force_atol"]
    assert np.abs(r["virial"].reshape(-1) - ref["virial"]).max() <= ref["virial_atol"]


def test_waterbox_gradient_is_tight(oracle):
    """The oracle reproduces the printed c47b1 gradients to their 5-decimal
    rounding, far inside the reference's 0.01 tolerance."""
    spec, xyzq, ref = golden_util.load("waterbox")
    r = oracle.evaluate(oracle.prepare(spec), xyzq)
    assert np.abs(r["grad"] - ref["grad"]).max() < 2e-5
