"""The reference's existing C ABI on the GPU (include/apocharmm_amd_capi.h),
driven through the project's ctypes mirror of python/apocharmm
(apocharmm_amd/apocharmm_c.py):
- the c47b1 waterbox read by the C++ readers meets the reference's golden
  values (test/unittests/unittest-eneForces.cpp:313-341);
- the JAC system built by the C++ resolver gives bit for bit the gradient of
  the Python-built system;
- CharmmContext + CudaLangevinThermostatIntegrator run SETTLE + SHAKE
  dynamics with restart and DCD subscribers, and a run resumed from a restart
  file continues the trajectory it was written from."""
import ctypes
import gzip
import os
import shutil
import struct

import numpy as np
import pytest

import golden_util
from apocharmm_amd import _lib, systems
from apocharmm_amd import apocharmm_c as A
from apocharmm_amd import charmm_io as C

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

D = os.path.join(golden_util.GOLDEN, "charmm")


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    out = tmp_path_factory.mktemp("charmm")
    paths = {}
    for name in ("waterbox.psf", "waterbox.crd", "toppar_water_ions.str", "jac_5dhfr.psf",
                 "jac_5dhfr.crd", "par_all22_prot.prm"):
        with gzip.open(os.path.join(D, name + ".gz"), "rb") as fi, open(out / name, "wb") as fo:
            shutil.copyfileobj(fi, fo)
        paths[name] = str(out / name)
    return paths


def _eval_nb(fm, xyz, charges):
    """Energies, virial and gradient of the manager's force path at xyz."""
    nb = fm.nbSystem()
    L = _lib.lib()
    n = len(xyz)
    st = ctypes.c_int32()
    _lib.check(L.apo_nb_force_stride(nb, ctypes.byref(st)))
    x = torch.tensor(np.hstack([xyz, np.asarray(charges)[:, None]]), dtype=torch.float32,
                     device="cuda")
    f = torch.zeros(3 * st.value, dtype=torch.int64, device="cuda")
    e = torch.zeros(_lib.NUM_ENERGY, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, dtype=torch.float64, device="cuda")
    flags = _lib.APO_RESET_LIST | _lib.APO_CALC_ENERGY | _lib.APO_CALC_VIRIAL
    _lib.check(L.apo_nb_calc_force(nb, ctypes.c_void_p(x.data_ptr()), flags,
                                   ctypes.c_void_p(f.data_ptr()), st.value,
                                   ctypes.c_void_p(e.data_ptr()), ctypes.c_void_p(v.data_ptr()),
                                   None))
    torch.cuda.synchronize()
    g = f.view(3, st.value)[:, :n].t().cpu().numpy()
    return dict(grad_fixed=g, grad=g * 2.0 ** -40,
                energy=dict(zip(_lib.ENERGY_NAMES, e.cpu().numpy())), virial=v.cpu().numpy())


def test_waterbox_through_reference_capi_matches_c47b1(files):
    psf = A.CharmmPsf(files["waterbox.psf"])
    prm = A.CharmmParameters([files["toppar_water_ions.str"]])
    fm = A.ForceManager(psf, prm)
    fm.setBoxDimensions((50.0, 50.0, 50.0))
    fm.setFFTGrid(48, 48, 48)
    fm.setCutoff(10.0)
    fm.setCtofnb(8.0)
    fm.setCtonnb(7.0)
    ctx = A.CharmmContext(fm)
    crd = A.CharmmCrd(files["waterbox.crd"])
    got = _eval_nb(fm, crd.getCoordinates(), psf.getCharges())
    _, _, ref = golden_util.load("waterbox")
    print("\n[capi waterbox]", {k: round(float(v), 5) for k, v in got["energy"].items()},
          flush=True)
    for k, v in ref["energy"].items():
        assert abs(got["energy"][k] - v) <= ref["energy_rtol"] * abs(v) + ref["energy_atol"], k
    assert np.abs(got["grad"] - ref["grad"]).max() <= ref["force_atol"]
    assert np.abs(got["virial"] - ref["virial"]).max() <= ref["virial_atol"]
    ctx.close()
    fm.close()


def test_jac_capi_gradient_bitwise_equals_python_built(files):
    psf = A.CharmmPsf(files["jac_5dhfr.psf"])
    prm = A.CharmmParameters([files["par_all22_prot.prm"], files["toppar_water_ions.str"]])
    fm = A.ForceManager(psf, prm)
    fm.setBoxDimensions((62.23,) * 3)
    fm.setFFTGrid(64, 64, 64)
    fm.setCtonnb(10.0)
    fm.setCtofnb(12.0)
    ctx = A.CharmmContext(fm)
    xyz = A.CharmmCrd(files["jac_5dhfr.crd"]).getCoordinates()
    got = _eval_nb(fm, xyz, psf.getCharges())
    from apocharmm_amd.force_manager import ForceManager
    spec = C.charmm_system(files["jac_5dhfr.psf"],
                           [files["par_all22_prot.prm"], files["toppar_water_ions.str"]],
                           files["jac_5dhfr.crd"], (62.23,) * 3, 64)
    pfm = ForceManager(spec).initialize()
    x = torch.tensor(systems.xyzq_of(spec), device="cuda")
    pfm.calcForce(x, reset=True, calcEnergy=True)
    torch.cuda.synchronize()
    ref = pfm.getForcesFixed().cpu().numpy()
    e = pfm.getEnergyComponents()
    print("\n[capi jac] epot %.6f (python-built %.6f)" % (got["energy"]["epot"], e["epot"]),
          flush=True)
    assert np.array_equal(got["grad_fixed"], ref)
    for k in _lib.ENERGY_NAMES:
        assert got["energy"][k] == e[k], k
    pfm.close()
    ctx.close()
    fm.close()


def _jac_context(files, seed):
    psf = A.CharmmPsf(files["jac_5dhfr.psf"])
    prm = A.CharmmParameters([files["par_all22_prot.prm"], files["toppar_water_ions.str"]])
    fm = A.ForceManager(psf, prm)
    fm.setBoxDimensions((62.23,) * 3)
    fm.setFFTGrid(64, 64, 64)
    fm.setCtonnb(10.0)
    fm.setCtofnb(12.0)
    ctx = A.CharmmContext(fm)
    ctx.setCoordinates(A.CharmmCrd(files["jac_5dhfr.crd"]))
    ctx.setRandomSeedForVelocities(seed)
    ctx.assignVelocitiesAtTemperature(300.0)
    return psf, fm, ctx


def test_context_integrator_subscribers_and_restart(files, tmp_path):
    psf, fm, ctx = _jac_context(files, 42)
    n = psf.getNumAtoms()
    t0 = ctx.computeTemperature()
    ndof = 3 * n - 3 - 3 * 7023 - 1221
    # draws of 3N unconstrained components, reported per constrained dof
    print("\n[capi ctx] assigned velocities: T = %.1f K (ndof %d)" % (t0, ndof), flush=True)
    assert abs(t0 / (300.0 * 3 * n / ndof) - 1) < 0.03
    _, _, ctx2 = _jac_context(files, 42)
    assert ctx2.computeTemperature() == t0  # the seed fixes the draws
    lt = A.CudaLangevinThermostatIntegrator(0.002, friction=1.0, temperature=300.0, seed=9)
    lt.setCharmmContext(ctx)
    rst = A.RestartSubscriber(str(tmp_path / "jac.rst"), 20)
    dcd = A.DcdSubscriber(str(tmp_path / "jac.dcd"), 10)
    lt.subscribe(rst)
    lt.subscribe(dcd)
    lt.propagate(40)
    a40 = lt.getState()
    assert a40["total_steps"] == 40
    shutil.copy(tmp_path / "jac.rst", tmp_path / "jac40.rst")  # the next writes replace it
    # constraints: water geometry and SHAKE bonds
    spec = C.charmm_system(files["jac_5dhfr.psf"],
                           [files["par_all22_prot.prm"], files["toppar_water_ions.str"]],
                           files["jac_5dhfr.crd"], (62.23,) * 3, 64)
    x = a40["xyz"]
    w = spec["waters"]
    assert np.abs(np.linalg.norm(x[w[:, 0]] - x[w[:, 1]], axis=1) - 0.9572).max() < 1e-9
    g, r2 = spec["shake_groups"], spec["shake_r2"]
    for k in range(1, 4):
        m = g[:, k] >= 0
        d2 = np.sum((x[g[m, 0]] - x[g[m, k]]) ** 2, 1)
        assert np.abs(d2 / r2[m] - 1).max() < 1e-12
    t40 = ctx.computeTemperature()
    print("[capi md] 40 steps at 2 fs with SETTLE + SHAKE: T = %.1f K" % t40, flush=True)
    assert 200.0 < t40 < 400.0
    # DCD: header with 4 frames (every 10 steps), one unit cell + 3 coordinate records each
    raw = open(tmp_path / "jac.dcd", "rb").read()
    assert struct.unpack("<i4s", raw[:8]) == (84, b"CORD")
    nfile, = struct.unpack("<i", raw[8:12])
    assert nfile == 4
    frame = (4 + 48 + 4) + 3 * (4 + 4 * n + 4)
    header = 4 + 84 + 4 + 4 + 164 + 4 + 4 + 4 + 4
    assert len(raw) == header + 4 * frame
    last = np.frombuffer(raw[len(raw) - (4 + 4 * n + 4) + 4:len(raw) - 4], np.float32)
    assert np.array_equal(last, x[:, 2].astype(np.float32))
    # resume: a new context and integrator read the restart written at step 40
    lt.propagate(20)
    a60 = lt.getState()
    _, _, ctxb = _jac_context(files, 7)
    lb = A.CudaLangevinThermostatIntegrator(0.002, friction=1.0, temperature=300.0, seed=123)
    lb.setCharmmContext(ctxb)
    lb.initializeFromRestartFile(str(tmp_path / "jac40.rst"))
    b40 = lb.getState()
    assert b40["total_steps"] == 40
    assert np.abs(b40["xyz"] - a40["xyz"]).max() < 1e-12
    assert np.abs(b40["dprev"] - a40["dprev"]).max() < 1e-15
    lb.propagate(20)
    b60 = lb.getState()
    dx = np.abs(b60["xyz"] - a60["xyz"]).max()
    print("[capi restart] resumed run after 20 more steps: max |dx| = %.2e A" % dx, flush=True)
    assert dx < 1e-9
    for o in (lb, ctxb, lt, rst, dcd, ctx, ctx2, fm):
        o.close()
