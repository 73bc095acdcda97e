"""GPU integrator (apo_lt_*): Langevin leapfrog + SETTLE on the HIP force path.

Parity: every step is compared with the oracle's restatement
(oracle/md.c) driven by the gradient the GPU step used, so the comparison
isolates the integrator (tolerances are FP64 rounding, ~1e-12).  The graph
replay must be bitwise identical to eager stepping; the physics checks are
SETTLE geometry, NVE energy conservation and the thermostat temperature.
"""
import numpy as np
import pytest

from apocharmm_amd import systems

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DT_PS = 0.002
FRICTION = 1.0        # 1/ps
TEMP = 300.0
SEED = 2003


def _centered_c1():
    sp = systems.config1()
    L = np.asarray(sp["box"], np.float64)
    xyz = np.asarray(sp["xyz"], np.float64) - 0.5 * L   # molecule centroids inside +-L/2
    return sp, xyz


def _make(sp, xyz, vel, dt=DT_PS, friction=FRICTION, temp=TEMP, seed=SEED, graph=True,
          precision="mixed", nbfreq=20, prune=0):
    from apocharmm_amd.force_manager import ForceManager
    from apocharmm_amd.integrator import LangevinThermostatIntegrator
    fm = ForceManager(sp, precision).initialize()
    lt = LangevinThermostatIntegrator(dt, friction=friction, temperature=temp, seed=seed)
    lt.setNonbondedListUpdateFrequency(nbfreq)
    lt.setListPruneFrequency(prune)
    lt.setUseGraph(graph)
    lt.setCharmmContext(fm, xyz, vel)
    return fm, lt


@pytest.fixture(scope="module")
def c1():
    from apocharmm_amd.integrator import maxwell_velocities
    sp, xyz = _centered_c1()
    vel = maxwell_velocities(sp["masses"], TEMP, 1002)
    return sp, xyz, vel


def _oracle_lt(oracle, sp, friction=FRICTION, temp=TEMP, seed=SEED):
    invm = 1.0 / np.asarray(sp["masses"], np.float64)
    return oracle.Langevin(invm, np.asarray(sp["waters"], np.int32), DT_PS, temp, friction, seed)


def test_md_initialize_and_steps_match_oracle(c1, oracle):
    sp, xyz, vel = c1
    fm, lt = _make(sp, xyz, vel)
    lt.initialize()
    st = lt.getState()
    ol = _oracle_lt(oracle, sp)
    x, v = xyz.copy(), vel.copy()
    ol.init_constrain(x, v)
    dprev = ol.init_backstep(st["grad"], x, v)
    print("\n[init] |dx|=%.2e |dv|=%.2e |ddprev|=%.2e" % (
        np.abs(st["xyz"] - x).max(), np.abs(st["vel"] - v).max(),
        np.abs(st["dprev"] - dprev).max()), flush=True)
    assert np.abs(st["xyz"] - x).max() < 1e-11
    assert np.abs(st["vel"] - v).max() < 1e-10
    assert np.abs(st["dprev"] - dprev).max() < 1e-12
    g0 = st["grad"]
    for s in range(10):
        lt.propagate(1)            # centering (no-op here), list rebuild, step
        st = lt.getState()
        if s == 0:
            ol.remove_com(dprev)   # comfreq 1000 removes at step 0
            # the list rebuild does not change the gradient: fixed-point sums
            # are independent of tile order
            assert np.array_equal(st["grad"], g0)
        ke = ol.step(st["grad"], x, dprev, v, s)
        th = lt.getKineticEnergy()
        ex, ev, ed = (np.abs(st["xyz"] - x).max(), np.abs(st["vel"] - v).max(),
                      np.abs(st["dprev"] - dprev).max())
        print("[step %d] |dx|=%.2e |dv|=%.2e |ddprev|=%.2e ke=%.6f/%.6f" % (
            s, ex, ev, ed, th[0], ke[0]), flush=True)
        assert ex < 1e-10 and ev < 1e-9 and ed < 1e-11
        assert abs(th[0] / ke[0] - 1) < 1e-10 and abs(th[1] / ke[1] - 1) < 1e-10
    assert lt.getCurrentStep() == 10
    lt.close()
    fm.close()


def test_md_graph_replay_bitwise_equals_eager(c1):
    sp, xyz, vel = c1
    out = []
    for graph in (True, False):
        fm, lt = _make(sp, xyz, vel, graph=graph)
        lt.propagate(45)
        out.append(lt.getState())
        lt.close()
        fm.close()
    for k in ("xyz", "vel", "dprev", "grad"):
        assert np.array_equal(out[0][k], out[1][k]), k
    assert np.array_equal(out[0]["thermo"], out[1]["thermo"])


def test_md_pruned_cycle_graphs_bitwise_equal_eager(c1):
    """With list pruning every 5 steps, whole-cycle graphs (two cycles of 20
    steps, prunes inside them) and step-by-step eager calls give the same bits."""
    sp, xyz, vel = c1
    out = []
    for graph in (True, False):
        fm, lt = _make(sp, xyz, vel, graph=graph, prune=5)
        lt.propagate(47)
        out.append(lt.getState())
        lt.close()
        fm.close()
    for k in ("xyz", "vel", "dprev", "grad"):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_md_launch_trial_cycles_bitwise_equal_eager(c1):
    """The first eight whole-cycle graphs run under four launch settings (PME
    issue priority off/on x pair kernel captured last/first; one untimed and
    one timed cycle each) and the fastest stays.  Across two propagate calls
    (ten cycles: the trial, the switch and cycles after it) the trajectory is
    bitwise the eager one, which keeps the default setting."""
    sp, xyz, vel = c1
    out = []
    for graph in (True, False):
        fm, lt = _make(sp, xyz, vel, graph=graph)
        lt.propagate(70)
        lt.propagate(145)
        out.append(lt.getState())
        lt.close()
        fm.close()
    for k in ("xyz", "vel", "dprev", "grad"):
        assert np.array_equal(out[0][k], out[1][k]), k
    assert np.array_equal(out[0]["thermo"], out[1]["thermo"])


def test_md_seed_reproducible_and_settle_geometry(c1):
    sp, xyz, vel = c1
    states = []
    for seed in (SEED, SEED, SEED + 1):
        fm, lt = _make(sp, xyz, vel, seed=seed)
        lt.propagate(30)
        states.append(lt.getState())
        lt.close()
        fm.close()
    assert np.array_equal(states[0]["xyz"], states[1]["xyz"])
    assert not np.array_equal(states[0]["xyz"], states[2]["xyz"])
    x = states[0]["xyz"]
    w = np.asarray(sp["waters"])
    d = lambda i, j: np.linalg.norm(x[w[:, i]] - x[w[:, j]], axis=1)
    assert np.abs(d(0, 1) - 0.9572).max() < 1e-9
    assert np.abs(d(0, 2) - 0.9572).max() < 1e-9
    assert np.abs(d(1, 2) - 1.5139).max() < 1e-9


def test_md_nve_energy_conservation(c1):
    """friction 0: H_n = U(x_n) + KE(v_n) is conserved by leapfrog + SETTLE;
    its fluctuation must be a small fraction of the potential's."""
    sp, xyz, vel = c1
    fm, lt = _make(sp, xyz, vel, dt=0.001, friction=0.0, nbfreq=10)
    lt.initialize()
    lt.propagate(100)                      # leave the lattice start
    # one step per call: U at the start-of-step x_n pairs with the step's
    # on-step kinetic energy
    H, U = [], []
    for _ in range(60):
        fm.calcForce(lt.getXYZQ(), calcEnergy=True)
        u = fm.getPotentialEnergy()
        lt.propagate(1)
        ke_on = lt.getKineticEnergy()[1]
        U.append(u)
        H.append(u + ke_on)
    H, U = np.array(H), np.array(U)
    print("\n[nve] std U=%.3f std H=%.4f drift=%.4f kcal/mol" % (U.std(), H.std(), H[-1] - H[0]),
          flush=True)
    assert H.std() < 0.05 * U.std()
    assert abs(H[-1] - H[0]) < 0.1 * U.std()
    lt.close()
    fm.close()


def test_md_langevin_temperature(c1):
    sp, xyz, vel = c1
    fm, lt = _make(sp, xyz, vel, friction=20.0)
    lt.propagate(500)
    T = []
    for _ in range(20):
        lt.propagate(25)
        T.append(lt.getTemperature())
    T = np.array(T)
    print("\n[nvt] <T> = %.2f / %.2f K" % (T[:, 0].mean(), T[:, 1].mean()), flush=True)
    assert abs(T[:, 0].mean() / TEMP - 1) < 0.03
    assert abs(T[:, 1].mean() / TEMP - 1) < 0.03
    lt.close()
    fm.close()


def test_md_batch_bitwise_equals_alone():
    """apo_lt_propagate_batch: each system's trajectory is bitwise the one it
    gets alone (same kernels, same order per system, integer force sums)."""
    from apocharmm_amd.integrator import maxwell_velocities, propagate_batch
    sizes = [500, 650, 830, 1100]
    specs = [systems.batched_system(i, w) for i, w in enumerate(sizes)]

    def fresh():
        out = []
        for i, sp in enumerate(specs):
            v = maxwell_velocities(sp["masses"], TEMP, 4200 + i)
            out.append(_make(sp, np.asarray(sp["xyz"], np.float64), v, seed=SEED + i, nbfreq=10))
        return out

    alone = []
    for fm, lt in fresh():
        lt.propagate(25)
        alone.append(lt.getState())
        lt.close()
        fm.close()
    runs = fresh()
    propagate_batch([lt for _, lt in runs], 25)
    for (fm, lt), ref in zip(runs, alone):
        st = lt.getState()
        assert st["step"] == ref["step"] == 25
        for k in ("xyz", "vel", "dprev", "grad"):
            assert np.array_equal(st[k], ref[k]), k
        lt.close()
        fm.close()


def test_reset_list_rejects_non_finite_coordinates():
    from apocharmm_amd import _lib
    from apocharmm_amd.force_manager import ForceManager
    sp, xyz = _centered_c1()
    fm = ForceManager(sp).initialize()
    x = systems.xyzq_of(sp, xyz)
    x[17, 1] = np.nan
    t = torch.tensor(x, device="cuda")
    with pytest.raises(_lib.ApoError, match="non-finite"):
        fm.resetNeighborList(t)
    fm.close()


def test_md_persistent_batch_lanes_bitwise_equals_alone():
    """apo_lt_batch_*: two propagate calls (graphs reused) with systems
    chained into 2 lanes give each system its lone trajectory bitwise."""
    from apocharmm_amd.integrator import LangevinBatch, maxwell_velocities
    sizes = [520, 610, 700]
    specs = [systems.batched_system(i + 7, w) for i, w in enumerate(sizes)]

    def fresh():
        out = []
        for i, sp in enumerate(specs):
            v = maxwell_velocities(sp["masses"], TEMP, 4300 + i)
            out.append(_make(sp, np.asarray(sp["xyz"], np.float64), v, seed=SEED + 11 * i,
                             nbfreq=7))
        return out

    alone = []
    for fm, lt in fresh():
        lt.propagate(9)
        lt.propagate(13)
        alone.append(lt.getState())
        lt.close()
        fm.close()
    runs = fresh()
    b = LangevinBatch([lt for _, lt in runs], lanes=2)
    b.propagate(9)
    b.propagate(13)
    b.close()
    for (fm, lt), ref in zip(runs, alone):
        st = lt.getState()
        for k in ("xyz", "vel", "dprev", "grad"):
            assert np.array_equal(st[k], ref[k]), k
        lt.close()
        fm.close()


@pytest.fixture(scope="module")
def c1_equilibrated(c1):
    """C1 after 400 Langevin steps at friction 20/ps (the lattice start is
    hot: its first steps move atoms by tenths of an A)."""
    sp, xyz, vel = c1
    fm, lt = _make(sp, xyz, vel, friction=20.0)
    lt.propagate(400)
    st = lt.getState()
    lt.close()
    fm.close()
    return sp, st["xyz"], st["vel"]


def _list_vs_fresh(start, policy, cutnb, nsteps, nbfreq):
    """Gradient at the final coordinates with the integrator's current list
    and with a list built from those coordinates: they differ by FP32
    rounding (the lists group the pairs differently) plus every pair inside
    r_off that the current list lacks."""
    from apocharmm_amd.force_manager import ForceManager
    from apocharmm_amd.integrator import LangevinThermostatIntegrator
    sp, xyz, vel = start
    sp = dict(sp, cutnb=cutnb)
    fm = ForceManager(sp).initialize()
    lt = LangevinThermostatIntegrator(DT_PS, friction=FRICTION, temperature=TEMP, seed=SEED)
    lt.setNonbondedListUpdateFrequency(nbfreq)
    lt.setListUpdatePolicy(policy)
    lt.setCharmmContext(fm, xyz, vel)
    lt.propagate(nsteps)
    info = lt.getListInfo()
    ptr = lt.getXYZQ()
    fm.calcForce(ptr)
    torch.cuda.synchronize()
    g_list = fm.getForces().cpu().numpy().copy()
    fm.calcForce(ptr, reset=True)
    torch.cuda.synchronize()
    g_fresh = fm.getForces().cpu().numpy().copy()
    lt.close()
    fm.close()
    err = float(np.abs(g_list - g_fresh).max())
    rel = float(np.sqrt(np.mean(np.sum((g_list - g_fresh) ** 2, 1))) /
                np.sqrt(np.mean(np.sum(g_fresh ** 2, 1))))
    return info, err, rel


def test_md_displacement_policy_rebuilds_before_pairs_are_missed(c1_equilibrated):
    # list buffer 0.6 A (half buffer 0.3 A), at most one rebuild per 60 steps
    info_f, err_f, rel_f = _list_vs_fresh(c1_equilibrated, "fixed", 12.6, 59, 60)
    print("\n[fixed] %s max |dg| %.2e rms rel %.2e" % (info_f, err_f, rel_f), flush=True)
    info, err, rel = _list_vs_fresh(c1_equilibrated, "displacement", 12.6, 59, 60)
    print("[displacement] %s max |dg| %.2e rms rel %.2e" % (info, err, rel), flush=True)
    assert info_f["max_displacement"] > info_f["half_buffer"]
    assert info_f["triggered_rebuilds"] == 0
    assert info["triggered_rebuilds"] > 0
    assert info["max_displacement"] <= info["half_buffer"]
    # FP32 rounding only (no missed pair); the fixed cadence missed pairs
    assert rel <= 1e-6
    assert err_f > 10 * err


def test_md_displacement_policy_with_default_buffer_keeps_all_pairs(c1_equilibrated):
    """The default 2 A buffer over 20-step intervals: the list still holds
    every pair inside r_off at the end."""
    info, err, rel = _list_vs_fresh(c1_equilibrated, "displacement", 14.0, 79, 20)
    print("\n[default buffer] %s max |dg| %.2e rms rel %.2e" % (info, err, rel), flush=True)
    assert info["max_displacement"] <= info["half_buffer"]
    assert rel <= 1e-6


def _pruned_vs_fresh(start, nbfreq, prune_every, nsteps):
    """Gradient of the last of nsteps steps, which the integrator took with its
    pruned list, against a list built from the same coordinates (the step
    began at xyz - dprev; both are exact in double): FP32 rounding plus every
    pair inside r_off that the pruned list lacks."""
    from apocharmm_amd.force_manager import ForceManager
    from apocharmm_amd.integrator import LangevinThermostatIntegrator
    sp, xyz, vel = start
    fm = ForceManager(sp).initialize()
    lt = LangevinThermostatIntegrator(DT_PS, friction=FRICTION, temperature=TEMP, seed=SEED)
    lt.setNonbondedListUpdateFrequency(nbfreq)
    lt.setListPruneFrequency(prune_every)
    lt.setCharmmContext(fm, xyz, vel)
    lt.propagate(nsteps)
    st = lt.getState()
    lt.close()
    x = (st["xyz"] - st["dprev"]).astype(np.float32)
    q = np.asarray(sp["charges"], np.float32)[:, None]
    x4 = torch.tensor(np.hstack([x, q]), device="cuda")
    fm.calcForce(x4, reset=True)
    torch.cuda.synchronize()
    g_fresh = fm.getForces().cpu().numpy().copy()
    fm.close()
    g = st["grad"].reshape(-1, 3)
    err = float(np.abs(g - g_fresh).max())
    rel = float(np.sqrt(np.mean(np.sum((g - g_fresh) ** 2, 1))) /
                np.sqrt(np.mean(np.sum(g_fresh ** 2, 1))))
    return err, rel


def test_md_list_pruning_keeps_every_pair(c1_equilibrated):
    """Pruning every 5 steps at r_off + 2 A * 5/20 (12.5 A): the last step of a
    20-step interval (list 19 steps old, prune 4 steps old) has every pair
    inside r_off; a control pruned once at 12.2 A and used for 99 steps loses
    pairs."""
    err, rel = _pruned_vs_fresh(c1_equilibrated, 20, 5, 40)
    print("\n[prune 5/20] max |dg| %.2e rms rel %.2e" % (err, rel), flush=True)
    err_c, rel_c = _pruned_vs_fresh(c1_equilibrated, 1000, 100, 100)
    print("[prune 100/1000 control] max |dg| %.2e rms rel %.2e" % (err_c, rel_c), flush=True)
    err_0, rel_0 = _pruned_vs_fresh(c1_equilibrated, 20, 0, 40)
    print("[no pruning] max |dg| %.2e rms rel %.2e" % (err_0, rel_0), flush=True)
    assert rel <= 1e-6 and rel_0 <= 1e-6
    assert err <= 2 * err_0 + 1e-4
    assert err_c > 10 * err


def test_md_batch_autotune_keeps_trajectories():
    """apo_lt_batch_autotune between propagate calls: the chosen launch shapes
    change no bit of any system's trajectory."""
    from apocharmm_amd.integrator import LangevinBatch, maxwell_velocities
    sizes = [540, 760, 980]
    specs = [systems.batched_system(i + 20, w) for i, w in enumerate(sizes)]

    def fresh():
        out = []
        for i, sp in enumerate(specs):
            v = maxwell_velocities(sp["masses"], TEMP, 4400 + i)
            out.append(_make(sp, np.asarray(sp["xyz"], np.float64), v, seed=SEED + 5 * i,
                             nbfreq=10))
        return out

    alone = []
    for fm, lt in fresh():
        lt.propagate(15)
        lt.propagate(15)
        alone.append(lt.getState())
        lt.close()
        fm.close()
    runs = fresh()
    b = LangevinBatch([lt for _, lt in runs], lanes=2)
    rep = b.autotune()
    print("\n[autotune] %s" % rep, flush=True)
    assert rep["launches_timed"] == 15 and rep["force_ms"] <= rep["default_force_ms"]
    assert rep["pme_ms"] <= rep["default_pme_ms"]
    b.propagate(15)
    b.autotune()
    b.propagate(15)
    b.close()
    for (fm, lt), ref in zip(runs, alone):
        st = lt.getState()
        for k in ("xyz", "vel", "dprev", "grad"):
            assert np.array_equal(st[k], ref[k]), k
        lt.close()
        fm.close()


def test_centering_moves_molecules_like_oracle(oracle):
    """Image centering where it acts (src/CharmmContext.cu:276-372): the C1
    lattice starts in [0, L), so most molecule centroids lie beyond +L/2 on
    some axis.  The first propagate step centres them (k_center) and steps;
    the oracle centres with orc_center and steps with the gradient the GPU
    step used.  Coordinates agree to FP64 rounding."""
    from apocharmm_amd.integrator import maxwell_velocities
    sp = systems.config1()
    xyz = np.asarray(sp["xyz"], np.float64)
    vel = maxwell_velocities(sp["masses"], TEMP, 1002)
    fm, lt = _make(sp, xyz, vel)
    lt.initialize()
    st = lt.getState()
    ol = _oracle_lt(oracle, sp)
    x, v = xyz.copy(), vel.copy()
    ol.init_constrain(x, v)
    dprev = ol.init_backstep(st["grad"], x, v)
    assert np.abs(st["xyz"] - x).max() < 1e-11
    w = np.sort(np.asarray(sp["waters"], np.int32), axis=1)
    gptr = np.arange(0, 3 * len(w) + 1, 3, dtype=np.int32)
    L = np.asarray(sp["box"], np.float64)
    xc = oracle.center(gptr, w.reshape(-1).copy(), L, x.copy())
    moved = int((np.abs(xc - x) > 0.5 * L.min()).reshape(len(w), -1).any(1).sum())
    print("\n[centering] %d of %d molecules move" % (moved, len(w)), flush=True)
    assert moved > len(w) // 2
    lt.propagate(1)
    st = lt.getState()
    ol.remove_com(dprev)
    ol.step(st["grad"], xc, dprev, v, 0)
    dx = np.abs(st["xyz"] - xc).max()
    print("[centering] max |x - x_oracle| = %.2e A" % dx, flush=True)
    assert dx < 1e-10
    cent = st["xyz"][w].mean(1)
    assert np.all(np.abs(cent) <= 0.5 * L + 0.05)
    lt.close()
    fm.close()


def test_destroy_order_is_enforced():
    """A force system outlives its integrators and an integrator its batches:
    destroying them out of order is refused with APO_INVALID_ARGUMENT and
    leaves every handle usable."""
    from apocharmm_amd import _lib
    from apocharmm_amd.integrator import LangevinBatch
    sp, xyz = _centered_c1()
    fm, lt = _make(sp, xyz, None)
    lt.propagate(2)
    with pytest.raises(_lib.ApoInvalidArgument, match="integrator"):
        fm.close()
    b = LangevinBatch([lt])
    with pytest.raises(_lib.ApoInvalidArgument, match="batch"):
        lt.close()
    b.propagate(2)
    b.close()
    lt.propagate(2)
    assert lt.getCurrentStep() == 6
    lt.close()
    fm.close()


def test_launch_settings_change_no_bit(c1):
    """The four fixed launch orders of the force branches and the timing trial
    give the same trajectory bit for bit; the trial reports its choice."""
    sp, xyz, vel = c1
    ref = None
    for setting in (None, 0, 1, 2, 3):
        fm, lt = _make(sp, xyz, vel)
        lt.setLaunchSetting(setting)
        lt.setCharmmContext(fm, xyz, vel)
        lt.propagate(400)
        st = lt.getState()
        info = lt.getLaunchInfo()
        print("\n[launch %s] %s" % (setting, info), flush=True)
        if setting is None:
            assert info["trial_cycles"] == 16 and info["setting"] in (0, 1, 2, 3)
            assert all(t > 0 for t in info["trial_ms"])
        else:
            assert info["setting"] == setting and info["trial_cycles"] == 0
        if ref is None:
            ref = st
        for k in ("xyz", "vel", "dprev", "grad"):
            assert np.array_equal(st[k], ref[k]), (setting, k)
        lt.close()
        fm.close()
