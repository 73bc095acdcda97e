"""The batched workload at its full size: 64 independent water boxes of
500-8,000 waters (seed 4001), each with its own ~1 A PME grid, advanced
together by the fused batch kernels (apo_lt_batch_*).

Parity: the gradient the batched kernels computed for the first step is
compared with the FP64 oracle for the smallest system, the largest and one
with an odd PME grid (force RMS relative error <= 2e-6, the same bar as a lone
system); over a whole list cycle every one of the 64 trajectories is bitwise
the one the system gets alone through apo_lt_propagate."""
import numpy as np
import pytest

from apocharmm_amd import systems

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DT_PS = 0.002
TEMP = 300.0
MIXED_FORCE_RTOL = 2e-6


def _make(i, sp):
    from apocharmm_amd.force_manager import ForceManager
    from apocharmm_amd.integrator import LangevinThermostatIntegrator, maxwell_velocities
    fm = ForceManager(sp).initialize()
    lt = LangevinThermostatIntegrator(DT_PS, friction=1.0, temperature=TEMP, seed=4500 + i)
    lt.setCharmmContext(fm, np.asarray(sp["xyz"], np.float64),
                        maxwell_velocities(sp["masses"], TEMP, 4600 + i))
    return fm, lt


def _water_groups(sp):
    w = np.sort(np.asarray(sp["waters"], np.int32), axis=1)
    return np.arange(0, 3 * len(w) + 1, 3, dtype=np.int32), w.reshape(-1).copy()


def test_c4_64_systems_batched_against_oracle_and_lone_runs(oracle):
    from apocharmm_amd.integrator import LangevinBatch
    sizes = systems.batched_sizes(64)
    specs = [systems.batched_system(i, int(w)) for i, w in enumerate(sizes)]
    grids = [int(sp["fft_grid"][0]) for sp in specs]
    odd = [i for i, k in enumerate(grids) if k % 2 == 1]
    check = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), odd[0] if odd else 0})
    print("\n[c4] 64 systems, %d..%d atoms (%d summed), grids %d..%d; oracle on %s" % (
        3 * sizes.min(), 3 * sizes.max(), 3 * sizes.sum(), min(grids), max(grids),
        [(i, 3 * int(sizes[i]), grids[i]) for i in check]), flush=True)
    assert odd, "the seeded sizes include an odd grid"

    runs = [_make(i, sp) for i, sp in enumerate(specs)]
    x0 = {}
    for i in check:
        runs[i][1].initialize()
        x0[i] = runs[i][1].getCoordinates()
    b = LangevinBatch([lt for _, lt in runs], lanes=4)
    b.propagate(1)
    torch.cuda.synchronize()
    for i in check:
        sp = specs[i]
        gptr, gatom = _water_groups(sp)
        xc = oracle.center(gptr, gatom, np.asarray(sp["box"], np.float64), x0[i].copy())
        ref = oracle.evaluate(oracle.prepare(sp), systems.xyzq_of(sp, xc))
        g = runs[i][1].getGradient()
        err = float(np.sqrt(np.mean(np.sum((g - ref["grad"]) ** 2, 1)) /
                            np.mean(np.sum(ref["grad"] ** 2, 1))))
        print("[c4] system %d (%d atoms, grid %d): batched step-0 force_rms_rel=%.3e" % (
            i, len(g), grids[i], err), flush=True)
        assert err <= MIXED_FORCE_RTOL
    b.propagate(19)
    b.close()
    states = []
    for fm, lt in runs:
        states.append(lt.getState())
        lt.close()
        fm.close()
    print("[c4] batch: 20 steps done", flush=True)
    for i, sp in enumerate(specs):
        fm, lt = _make(i, sp)
        lt.propagate(1)
        lt.propagate(19)
        st = lt.getState()
        lt.close()
        fm.close()
        assert st["step"] == states[i]["step"] == 20
        for k in ("xyz", "vel", "dprev", "grad"):
            assert np.array_equal(st[k], states[i][k]), (i, k)
    print("[c4] all 64 trajectories bitwise equal to their lone runs", flush=True)
