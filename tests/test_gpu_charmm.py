"""Systems read from CHARMM files (PSF / CRD / parameter stream) through the
HIP force path: the c47b1 waterbox from its own files against the c47b1
golden values, and the JAC DHFR benchmark system (23,558 atoms: protein with
CHARMM22 bonds, angles, Urey-Bradley, dihedral series and impropers, 7,023
TIP3P waters) against the FP64 oracle, then a short Langevin/SETTLE run."""
import os

import numpy as np
import pytest

import golden_util
from apocharmm_amd import charmm_io as C
from apocharmm_amd import systems

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

D = os.path.join(golden_util.GOLDEN, "charmm")


def f(name):
    return os.path.join(D, name + ".gz")


def _eval(spec, precision="mixed"):
    from apocharmm_amd.force_manager import ForceManager
    fm = ForceManager(spec, precision).initialize()
    x = torch.tensor(systems.xyzq_of(spec), device="cuda")
    fm.calcForce(x, reset=True, calcEnergy=True, calcVirial=True)
    torch.cuda.synchronize()
    out = dict(grad=fm.getForces().cpu().numpy().copy(), energy=fm.getEnergyComponents(),
               virial=fm.getVirial())
    fm.close()
    return out


def _rms_rel(g, ref):
    return float(np.sqrt(np.mean(np.sum((g - ref) ** 2, 1))) /
                 np.sqrt(np.mean(np.sum(ref ** 2, 1))))


def test_waterbox_from_charmm_files_matches_c47b1():
    spec = C.charmm_system(f("waterbox.psf"), [f("toppar_water_ions.str")], f("waterbox.crd"),
                           (50.0, 50.0, 50.0), 48, cutnb=10.0, ctofnb=8.0, ctonnb=7.0)
    _, _, ref = golden_util.load("waterbox")
    print("case waterbox from PSF/CRD/STR", flush=True)
    got = _eval(spec)
    for k, v in ref["energy"].items():
        tol = ref["energy_rtol"] * abs(v) + ref["energy_atol"]
        print("  %-5s gpu %.5f c47b1 %.5f" % (k, got["energy"][k], v), flush=True)
        assert abs(got["energy"][k] - v) <= tol, k
    assert np.abs(got["grad"] - ref["grad"]).max() <= ref["force_atol"]
    assert np.abs(got["virial"].reshape(-1) - ref["virial"]).max() <= ref["virial_atol"]


@pytest.fixture(scope="module")
def jac_spec():
    return C.charmm_system(f("jac_5dhfr.psf"), [f("par_all22_prot.prm"), f("toppar_water_ions.str")],
                           f("jac_5dhfr.crd"), (62.23,) * 3, 64)


def test_jac_dhfr_against_oracle(jac_spec, oracle):
    spec = jac_spec
    print("case jac_5dhfr 23558 atoms", flush=True)
    ref = oracle.evaluate(oracle.prepare(spec), systems.xyzq_of(spec), nthreads=16)
    scale = sum(abs(v) for k, v in ref["energy"].items() if k != "epot")
    for precision, ftol, etol in (("mixed", 2e-6, 2e-7), ("fp64", 5e-8, 1e-10)):
        got = _eval(spec, precision)
        err = _rms_rel(got["grad"], ref["grad"])
        print("[jac %s] force_rms_rel=%.3e" % (precision, err), flush=True)
        assert err <= ftol
        for k in ("bond", "angle", "ureyb", "dihe", "imdihe", "ewks", "ewse", "ewex", "elec",
                  "vdw"):
            r = ref["energy"][k]
            print("   %-6s gpu %.8g oracle %.8g" % (k, got["energy"][k], r), flush=True)
            assert abs(got["energy"][k] - r) <= etol * scale, k


def test_jac_dhfr_langevin_settle_run(jac_spec):
    from apocharmm_amd.force_manager import ForceManager
    from apocharmm_amd.integrator import LangevinThermostatIntegrator, maxwell_velocities
    spec = jac_spec
    fm = ForceManager(spec).initialize()
    lt = LangevinThermostatIntegrator(0.001, friction=5.0, temperature=300.0, seed=7)
    lt.setCharmmContext(fm, spec["xyz"], maxwell_velocities(spec["masses"], 300.0, 8))
    lt.propagate(200)
    st = lt.getState()
    T = lt.getTemperature()
    print("[jac md] 200 steps at 1 fs: T = %.1f K" % T[1], flush=True)
    assert np.all(np.isfinite(st["xyz"])) and 150.0 < T[1] < 450.0
    w = spec["waters"]
    x = st["xyz"]
    d = np.linalg.norm(x[w[:, 0]] - x[w[:, 1]], axis=1)
    assert np.abs(d - 0.9572).max() < 1e-9
    lt.close()
    fm.close()


def _molecule_groups(spec):
    """Molecules (connected components over bonds and waters) as CSR groups,
    for the oracle's image centering."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    n = len(spec["charges"])
    b = np.asarray(spec["bonds"], np.int64).reshape(-1, 2)
    w = np.asarray(spec["waters"], np.int64).reshape(-1, 3)
    e = np.vstack([b, w[:, [0, 1]], w[:, [0, 2]]]) if len(w) else b
    _, lab = connected_components(coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n)),
                                  directed=False)
    order = np.argsort(lab, kind="stable")
    counts = np.bincount(lab)
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return ptr, order.astype(np.int32)


def test_jac_shake_settle_steps_match_oracle(jac_spec, oracle):
    """SHAKE of the protein's bonds to hydrogen (790 groups, 1,221 bonds) with
    SETTLE of the waters: initialisation and five 2 fs steps on the GPU equal
    the oracle's integrator (iterative SHAKE, oracle/md.c) driven by the
    gradients the GPU used, to FP64 rounding; every constrained bond holds."""
    from apocharmm_amd.force_manager import ForceManager
    from apocharmm_amd.integrator import LangevinThermostatIntegrator, maxwell_velocities
    sp = jac_spec
    g, r2 = sp["shake_groups"], sp["shake_r2"]
    print("\n[jac shake] %d groups, %d bonds, %d waters" % (len(g), int((g[:, 1:] >= 0).sum()),
                                                           len(sp["waters"])), flush=True)
    ptr, atoms = _molecule_groups(sp)
    x = oracle.center(ptr, atoms, np.asarray(sp["box"], np.float64),
                      np.array(sp["xyz"], np.float64))
    v = maxwell_velocities(sp["masses"], 300.0, 17)
    fm = ForceManager(sp).initialize()
    lt = LangevinThermostatIntegrator(0.002, friction=1.0, temperature=300.0, seed=19)
    lt.setCharmmContext(fm, x, v)
    assert lt.ndegf == 3 * len(x) - 3 - 3 * len(sp["waters"]) - 1221
    lt.initialize()
    st = lt.getState()
    invm = 1.0 / np.asarray(sp["masses"], np.float64)
    ol = oracle.Langevin(invm, sp["waters"], 0.002, 300.0, 1.0, 19, shake=(g, r2))
    xo, vo = x.copy(), v.copy()
    ol.init_constrain(xo, vo)
    dprev = ol.init_backstep(st["grad"], xo, vo)
    print("[jac shake init] |dx|=%.2e |dv|=%.2e |ddprev|=%.2e" % (
        np.abs(st["xyz"] - xo).max(), np.abs(st["vel"] - vo).max(),
        np.abs(st["dprev"] - dprev).max()), flush=True)
    assert np.abs(st["xyz"] - xo).max() < 1e-10
    assert np.abs(st["vel"] - vo).max() < 1e-8
    assert np.abs(st["dprev"] - dprev).max() < 1e-11
    box = np.asarray(sp["box"], np.float64)
    for s in range(5):
        lt.propagate(1)  # each call centres molecules first (a whole-box shift)
        st = lt.getState()
        if s == 0:
            ol.remove_com(dprev)
        oracle.center(ptr, atoms, box, xo)
        ol.step(st["grad"], xo, dprev, vo, s)
        ex, ev = np.abs(st["xyz"] - xo).max(), np.abs(st["vel"] - vo).max()
        print("[jac shake step %d] |dx|=%.2e |dv|=%.2e" % (s, ex, ev), flush=True)
        assert ex < 1e-9 and ev < 1e-7
    xg = st["xyz"]
    worst = 0.0
    for k in range(1, 4):
        m = g[:, k] >= 0
        d2 = np.sum((xg[g[m, 0]] - xg[g[m, k]]) ** 2, 1)
        worst = max(worst, float(np.abs(d2 / r2[m] - 1).max()))
    print("[jac shake] worst |r^2/b0^2 - 1| = %.2e" % worst, flush=True)
    assert worst < 1e-12
    lt.close()
    fm.close()
