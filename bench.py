"""Benchmark: ns/day of NVT Langevin dynamics on the 98,304-atom TIP3P PME box.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1|c2|c3|c4|c5]
                    [--precision mixed|fp64] [--step md|force] [--no-cpu-baseline]
                    [--nsys 64] [--lanes 4]           (c4: batched systems)
                    [--jobs 40] [--rate 4] [--step-scale 0.1] [--chunk 100]  (c5: serving)

One "step" (default --step md) is one leapfrog Langevin step at dt = 2 fs,
friction 1/ps, 300 K with SETTLE rigid waters: the full force path (bonded +
direct-space pair kernel + 1-4/exclusion corrections + PME spread, FFT,
convolution, gather) and the integrator update, through the C ABI
(apo_lt_propagate), with the neighbor list rebuilt and images centred every
20 steps (the reference's default, src/CudaIntegrator.cu:29) and the other
steps replayed as a hipGraph.  For systems of 50,000 atoms and more, the pair
list is pruned every 10 steps between rebuilds (--prune-every; 0 runs the
reference's list unchanged): entries whose pairs all lie beyond
r_off + 2 A * 10/20 are dropped, so every pair inside r_off is still computed.
All state is resident in HBM.
ns/day = steps * dt[ps] / 1000 / (wall_s / 86400) (unittest-benchmarks.cpp:286-287).
--step force times force evaluations alone.

Prints one JSON line with the metric, a roofline entry for the dominant kernel
(the pair kernel, FP32 VALU work counted as 61 flop per pair inside r_off,
timed with HIP events in a second, profiled pass of the same steps) and the
CPU baseline (the FP64 oracle force + integrator step on this host's cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DT_PS = 0.002
FRICTION = 1.0        # 1/ps (unittest-langevinThermostat.cpp:49)
TEMP = 300.0
SEED = 2003
REBUILD_EVERY = 20
# pruning the pair list every 10 steps pays from here up (measured: 98,304
# atoms 572.6 -> 587.2 ns/day; 3,000 atoms 1,770 -> 1,706; 23,558 atoms within
# run-to-run spread)
PRUNE_MIN_ATOMS = 50000
FLOP_PER_PAIR = 61
PEAK_FP32_TFLOPS = 157.3          # packed FP32 (v_pk_fma_f32)
PEAK_FP32_NONPACKED_TFLOPS = 78.6  # what a build without packed FP32 can issue
# MD runs start from a seeded lattice, which releases potential energy as it
# melts (the first steps heat the box to ~900 K): a Langevin equilibration at
# strong friction brings it to the 300 K state before anything is timed
EQUIL_STEPS = 2000
EQUIL_FRICTION = 10.0
KBOLTZ = 1.987191e-3  # kcal/mol/K


def count_pairs(spec, rc):
    """Non-excluded minimum-image pairs with r < rc (host, periodic KD-tree;
    the 1-2/1-3/1-4 exclusions from the product's own derivation,
    apo_topology_exclusions)."""
    import ctypes
    import numpy as np
    from scipy.spatial import cKDTree
    from apocharmm_amd import _lib
    L = float(spec["box"][0])
    x = np.mod(spec["xyz"], L)
    tree = cKDTree(x, boxsize=L)
    n = len(x)
    pairs = (tree.count_neighbors(tree, rc) - n) // 2
    b = np.ascontiguousarray(np.asarray(spec.get("bonds", np.zeros((0, 2))), np.int32).reshape(-1, 2))
    counts = (ctypes.c_int32 * 3)()
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib = _lib.lib()
    _lib.check(lib.apo_topology_exclusions(n, len(b), vp(b), counts, None, None, None, None))
    ptr = np.zeros(n + 1, np.int32)
    idx = np.zeros(max(counts[0], 1), np.int32)
    in14 = np.zeros(max(2 * counts[1], 2), np.int32)
    ex14 = np.zeros(max(2 * counts[2], 2), np.int32)
    _lib.check(lib.apo_topology_exclusions(n, len(b), vp(b), counts, vp(ptr), vp(idx), vp(in14),
                                           vp(ex14)))
    i = np.repeat(np.arange(n), np.diff(ptr))
    j = idx[: counts[0]]
    keep = i < j
    d = x[i[keep]] - x[j[keep]]
    d -= L * np.round(d / L)
    n_excl = int((np.sum(d * d, 1) < rc * rc).sum())
    return int(pairs - n_excl)


VSEED = {"c1": 1002, "c2": 2002, "c3": 3005, "jac": 5003}
CHARMM_DIR = os.path.join(ROOT, "tests", "golden", "charmm")


def config_spec(name):
    """c1-c3: the synthetic TIP3P / solvated-chain systems; jac: the JAC DHFR
    benchmark system (23,558 atoms, CHARMM22 + TIP3P, box 62.23 A, PME 64^3)
    read from its PSF/CRD and parameter files."""
    from apocharmm_amd import systems
    if name == "jac":
        from apocharmm_amd import charmm_io
        f = lambda n: os.path.join(CHARMM_DIR, n + ".gz")
        return charmm_io.charmm_system(f("jac_5dhfr.psf"),
                                       [f("par_all22_prot.prm"), f("toppar_water_ions.str")],
                                       f("jac_5dhfr.crd"), (62.23,) * 3, 64, name="jac_5dhfr")
    return {"c1": systems.config1, "c2": systems.config2, "c3": systems.config3}[name]()


def host_cpus():
    """(threads to use, description): the CPUs this process may run on,
    capped by a cgroup CPU quota when one is set, and the host's CPU model."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    use = avail if quota is None else max(1, min(avail, int(quota)))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return use, {"cpu_model": model, "host_cpus": os.cpu_count(), "affinity_cpus": avail,
                 "cgroup_cpu_quota": quota}


def cpu_baseline(spec, threads, md):
    import numpy as np
    from apocharmm_amd import systems
    from oracle import oracle as O
    s = O.prepare(spec)
    xyzq = systems.xyzq_of(spec)
    n = len(xyzq)
    ol = O.Langevin(1.0 / np.asarray(spec["masses"], np.float64),
                    np.asarray(spec["waters"], np.int32), DT_PS, TEMP, FRICTION, SEED)
    x = np.ascontiguousarray(xyzq[:, :3], np.float64)
    v = np.zeros((n, 3))
    dprev = np.zeros((n, 3))
    O.evaluate(s, xyzq, nthreads=threads)  # warm
    t0 = time.perf_counter()
    reps = 0
    while True:
        r = O.evaluate(s, xyzq, nthreads=threads)
        if md:
            ol.step(r["grad"], x, dprev, v, reps)
        reps += 1
        dt = time.perf_counter() - t0
        if dt > 10.0 or reps >= 8:
            break
    t_eval = dt / reps
    what = "force evaluations + Langevin/SETTLE steps" if md else "force evaluations"
    return {"value": DT_PS * 86.4 / t_eval, "unit": "ns/day", "cores": threads, "kind": "port",
            "sample": "%d FP64 %s of the same system (oracle, cell list + PME, %d OpenMP "
                      "threads), %.3f s each; ns/day = dt*86.4/t_step" % (reps, what, threads, t_eval)}


def check_list(lt, fm, spec, prune_every):
    """List validity at the end of the timed run: one more whole list cycle
    (rebuild, prunes), then the gradient of its last step, which used a list
    nbfreq - 1 steps old (pruned prune_every - 1 steps before), against a list
    built fresh from the same coordinates.  They differ by FP32 rounding
    (pairs grouped differently) plus any pair inside r_off the list lacked."""
    import numpy as np
    import torch
    from apocharmm_amd import systems
    x_build = lt.getCoordinates()  # the cycle's rebuild step evaluates here
    lt.propagate(REBUILD_EVERY)
    st = lt.getState()
    info = lt.getListInfo()
    x = st["xyz"] - st["dprev"]  # where the last step evaluated the force
    fm.calcForce(torch.tensor(systems.xyzq_of(spec, x), device="cuda"), reset=True)
    torch.cuda.synchronize()
    g_fresh = fm.getForces().cpu().numpy()
    g = st["grad"]
    rel = float(np.sqrt(np.mean(np.sum((g - g_fresh) ** 2, 1)) /
                        np.mean(np.sum(g_fresh ** 2, 1))))
    missed, inside = missed_pairs(float(spec["box"][0]), x, x_build, float(spec["ctofnb"]),
                                  float(spec["cutnb"]))
    return {"max_displacement_A": round(info["max_displacement"], 4),
            "max_step_displacement_A": round(info["max_step_displacement"], 4),
            "half_buffer_A": info["half_buffer"],
            "list_age_steps": REBUILD_EVERY - 1,
            "prune_age_steps": (prune_every - 1) if prune_every else None,
            "grad_vs_fresh_list_rms_rel": rel,
            "grad_vs_fresh_list_max_abs": float(np.abs(g - g_fresh).max()),
            "pairs_inside_roff": inside,
            "pairs_inside_roff_beyond_list_cutoff_at_build": missed,
            "strict_displacement_bound": bool(info["max_displacement"] <= info["half_buffer"]),
            # the reference's fixed 20-step cadence can miss a pair whose atoms
            # closed in by more than the 2 A buffer; list_policy 1 (the
            # displacement check) rebuilds before that happens
            "no_missed_pairs": bool(missed == 0),
            "grad_within_2e-6_of_fresh_list": bool(rel <= 2e-6)}


def missed_pairs(L, x_eval, x_build, r_off, r_list):
    """Exact count of the pairs inside r_off at x_eval whose minimum-image
    distance at x_build (the list build) exceeded the list cutoff: pairs a
    list built at x_build cannot hold (periodic KD-tree, cubic box)."""
    import numpy as np
    from scipy.spatial import cKDTree
    xe = np.mod(x_eval, L)
    pairs = cKDTree(xe, boxsize=L).query_pairs(r_off, output_type="ndarray")
    missed = 0
    for a in range(0, len(pairs), 1 << 22):
        p = pairs[a:a + (1 << 22)]
        d = x_build[p[:, 0]] - x_build[p[:, 1]]
        d -= L * np.round(d / L)
        missed += int(np.count_nonzero(np.einsum("ij,ij->i", d, d) > r_list * r_list))
    return missed, int(len(pairs))


def copy_calibration():
    """A plain copy of known size: 256 MiB read and 256 MiB written per launch,
    three launches.  Its FETCH_SIZE / WRITE_SIZE readings scale the PMC
    counters of the other kernels to bytes."""
    import torch
    n = 64 << 20
    src = torch.ones(n, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": 4 * n, "copy_bytes_written_per_launch": 4 * n,
                      "launches": 3}), flush=True)


def pair_traffic(workload):
    """HBM bytes per pair-kernel launch from the committed rocprofv3 PMC summary
    of this workload (FETCH_SIZE and WRITE_SIZE in KiB, each scaled by the
    reading of a plain copy of known size), or None if there is none."""
    path = os.path.join(ROOT, "profiles", "k_pair_%s_pmc.md" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def bench_nve(a):
    """Energy conservation: Langevin equilibration (friction --equil-friction,
    --equil-steps steps), then NVE (friction 0) for --steps steps at 2 fs with
    the same list and SETTLE settings.  H = U(x_n) + on-step KE(v_n) is
    sampled every --sample-every steps; the drift is the least-squares slope
    of H over simulated time, per degree of freedom."""
    import numpy as np
    import torch
    from apocharmm_amd import systems
    from apocharmm_amd.force_manager import ForceManager
    from apocharmm_amd.integrator import LangevinThermostatIntegrator, maxwell_velocities

    torch.cuda.set_device(0)
    spec = config_spec(a.config)
    n = len(spec["charges"])
    fm = ForceManager(spec).initialize()
    vseed = VSEED[a.config]
    eq = LangevinThermostatIntegrator(DT_PS, friction=a.equil_friction, temperature=TEMP, seed=SEED)
    eq.setNonbondedListUpdateFrequency(REBUILD_EVERY)
    eq.setCharmmContext(fm, np.asarray(spec["xyz"], np.float64),
                        maxwell_velocities(spec["masses"], TEMP, vseed))
    eq.initialize()
    eq.propagate(a.equil_steps)
    st = eq.getState()
    t_eq = float(eq.getTemperature()[1])
    eq.close()
    lt = LangevinThermostatIntegrator(DT_PS, friction=0.0, temperature=TEMP, seed=SEED)
    lt.setNonbondedListUpdateFrequency(REBUILD_EVERY)
    lt.setCharmmContext(fm, st["xyz"], st["vel"])
    lt.initialize()
    H, U, K, steps = [], [], [], []
    k = max(1, a.sample_every)
    done = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while True:
        fm.calcForce(lt.getXYZQ(), calcEnergy=True)
        u = fm.getPotentialEnergy()
        lt.propagate(1)
        ke = float(lt.getKineticEnergy()[1])
        H.append(u + ke)
        U.append(u)
        K.append(ke)
        steps.append(done)
        done += 1
        if done >= a.steps:
            break
        m = min(k - 1, a.steps - done)
        if m > 0:
            lt.propagate(m)
            done += m
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    H, U, K = np.array(H), np.array(U), np.array(K)
    t_ns = np.array(steps) * DT_PS / 1000.0
    slope = float(np.polyfit(t_ns, H, 1)[0])
    ndof = 3 * n - 3 - 3 * len(spec.get("waters", []))
    kT = KBOLTZ * TEMP
    out = {
        "metric": "NVE total-energy drift, %s (%d steps at 2 fs after %d Langevin steps)" % (
            a.config, a.steps, a.equil_steps),
        "value": round(slope / ndof, 8),
        "unit": "kcal/mol/ns per degree of freedom",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.equil_steps,
        "ms_per_step": round(wall / a.steps * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+int64",
        "data": "synthetic (seeded system, Maxwell velocities seed %d)" % vseed,
        "config": {"workload": "%s_nve_%d_atoms" % (a.config, n), "natom": n, "ndof": ndof,
                   "dt_fs": DT_PS * 1000.0, "list_rebuild_every": REBUILD_EVERY,
                   "equilibration": {"steps": a.equil_steps, "friction_per_ps": a.equil_friction,
                                     "T_end_K": round(t_eq, 2)},
                   "sample_every": k, "parallelism": "single GPU"},
        "nve": {"drift_kcal_per_mol_per_ns": round(slope, 4),
                "drift_kT_per_ns_per_dof": round(slope / ndof / kT, 8),
                "H_first": round(float(H[0]), 4), "H_last": round(float(H[-1]), 4),
                "H_std": round(float(H.std()), 4), "U_std": round(float(U.std()), 4),
                "KE_mean": round(float(K.mean()), 3),
                "T_mean_K": round(float(2 * K.mean() / (ndof * KBOLTZ)), 2),
                "samples": len(H), "simulated_ps": round(a.steps * DT_PS, 3),
                "ns_per_day_incl_sampling": round(a.steps * DT_PS / 1000.0 / (wall / 86400.0), 2)},
        "roofline": None,
        "cpu_baseline": None,
    }
    lt.close()
    fm.close()
    print(json.dumps(out), flush=True)


def bench_batched(a):
    """C4: 64 independent TIP3P boxes of 500-8,000 waters (log-uniform, seed
    4001), each with its own box and ~1 A PME grid, NVT Langevin + SETTLE at
    2 fs.  Aggregate ns/day = systems x steps x dt / wall, batched (all systems
    advanced together, apo_lt_propagate_batch) against the same systems run
    one after another through the single-system path."""
    import numpy as np
    import torch
    from apocharmm_amd import systems
    from apocharmm_amd.force_manager import ForceManager
    from apocharmm_amd.integrator import (LangevinBatch, LangevinThermostatIntegrator,
                                          maxwell_velocities)

    torch.cuda.set_device(0)
    sizes = systems.batched_sizes(a.nsys)
    specs = [systems.batched_system(i, int(w)) for i, w in enumerate(sizes)]
    natoms = [len(sp["charges"]) for sp in specs]

    def make():
        out = []
        for i, sp in enumerate(specs):
            fm = ForceManager(sp, a.precision).initialize()
            lt = LangevinThermostatIntegrator(DT_PS, friction=FRICTION, temperature=TEMP,
                                              seed=SEED + i)
            lt.setNonbondedListUpdateFrequency(REBUILD_EVERY)
            lt.setCharmmContext(fm, np.asarray(sp["xyz"], np.float64),
                                maxwell_velocities(sp["masses"], TEMP, 4200 + i))
            lt.initialize()
            out.append((fm, lt))
        torch.cuda.synchronize()
        return out

    runs = make()
    # warm both paths (graph capture, first rebuilds)
    for fm, lt in runs:
        lt.propagate(a.warmup)
    torch.cuda.synchronize()
    t_seq = 0.0
    for fm, lt in runs:
        t0 = time.perf_counter()
        lt.propagate(a.steps)
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t0
    batch = LangevinBatch([lt for _, lt in runs], lanes=a.lanes)
    tuned = None if a.no_autotune else batch.autotune()
    batch.propagate(a.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.propagate(a.steps)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    batch.close()
    n = len(runs)
    agg = lambda t: n * a.steps * DT_PS / 1000.0 / (t / 86400.0)
    for fm, lt in runs:
        lt.close()
        fm.close()
    out = {
        "metric": "aggregate ns/day, %d uneven TIP3P PME boxes (NVT Langevin MD, 2 fs), batched"
                  % n,
        "value": round(agg(t_batch), 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_batch / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+int64" if a.precision == "mixed" else "fp64+int64",
        "data": "synthetic (seeded TIP3P lattices, random orientations, Maxwell velocities)",
        "config": {"workload": "c4_%d_tip3p_boxes_%d_to_%d_atoms" % (n, min(natoms), max(natoms)),
                   "systems": n, "atoms_total": int(sum(natoms)),
                   "atoms_min_max": [int(min(natoms)), int(max(natoms))],
                   "dt_fs": 2.0, "friction_per_ps": FRICTION, "temperature_K": TEMP,
                   "list_rebuild_every": REBUILD_EVERY, "precision": a.precision,
                   "gpu_max_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
                   "lanes": a.lanes, "autotune": tuned,
                   "parallelism": "single GPU, one graph launch per batch step"},
        "sequential": {"value": round(agg(t_seq), 3), "unit": "ns/day",
                       "ms_per_step_all_systems": round(t_seq / a.steps * 1e3, 4)},
        "speedup_vs_sequential": round(t_seq / t_batch, 3),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_serving(a):
    """C5: a seeded Poisson stream of NVT Langevin jobs (TIP3P boxes of
    log-uniform 1,500-100,000 atoms, step counts uniform 1,000-50,000 times
    --step-scale, rounded to whole chunks), served continuously with every job
    on its own stream (serving.StreamServer) and, for comparison, one job at a
    time in arrival order."""
    import torch
    from apocharmm_amd import serving

    torch.cuda.set_device(0)

    atoms = (1500, a.max_atoms)
    proto = serving.make_jobs(a.jobs, a.rate, atoms=atoms, step_scale=a.step_scale, chunk=a.chunk)
    for j in proto:
        j.prepare()

    def jobs():  # fresh job state over the same input systems
        js = serving.make_jobs(a.jobs, a.rate, atoms=atoms, step_scale=a.step_scale,
                               chunk=a.chunk)
        for j, p in zip(js, proto):
            j.spec = p.spec
        return js

    batched = serving.summarize(serving.StreamServer(max_slots=a.slots, chunk=a.chunk,
                                                     log=sys.stderr).run(jobs()))
    plain = serving.summarize(serving.run_plain(jobs(), chunk=a.chunk, log=sys.stderr))
    js = jobs()
    out = {
        "metric": "aggregate ns/day + p99 job latency, %d-job Poisson stream (NVT Langevin MD, "
                  "2 fs), continuous serving, one stream per job" % a.jobs,
        "value": round(batched["aggregate_ns_per_day"], 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": int(sum(j.nsteps for j in js)),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+int64",
        "data": "synthetic (seeded TIP3P lattices; Poisson arrivals seed 5001, jobs seed 5002)",
        "config": {"workload": "c5_poisson_%d_jobs_rate_%g_per_s" % (a.jobs, a.rate),
                   "jobs": a.jobs, "rate_per_s": a.rate, "step_scale": a.step_scale,
                   "chunk_steps": a.chunk, "max_slots": a.slots,
                   "atoms_range": list(atoms),
                   "atoms_total": int(sum(j.natom for j in js)),
                   "parallelism": "single GPU, one stream per job"},
        "served": {k: round(v, 4) for k, v in batched.items()},
        "plain": {k: round(v, 4) for k, v in plain.items()},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--config", default="c2", help="c1 | c2 | c3 | c4 | c5 | jac")
    ap.add_argument("--dt-fs", type=float, default=2.0, help="integrator time step (fs)")
    ap.add_argument("--precision", default="mixed")
    ap.add_argument("--step", choices=("md", "force"), default="md")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--prune-every", type=int, default=None,
                    help="MD: prune the pair list every K steps between rebuilds (0 = off, "
                         "the reference's list unchanged; default 10 from %d atoms up, else 0)"
                         % PRUNE_MIN_ATOMS)
    ap.add_argument("--nsys", type=int, default=64, help="systems in the c4 batch")
    ap.add_argument("--hw-queues", type=int, default=4,
                    help="GPU_MAX_HW_QUEUES for the c4 run (HIP's default is 4)")
    ap.add_argument("--lanes", type=int, default=4,
                    help="parallel chains of list rebuilds in the batch graph (c4, c5)")
    ap.add_argument("--jobs", type=int, default=40, help="c5: jobs in the Poisson stream")
    ap.add_argument("--rate", type=float, default=4.0, help="c5: arrivals per second")
    ap.add_argument("--step-scale", type=float, default=0.1,
                    help="c5: scale of the 1,000-50,000 step range (1 = the full range)")
    ap.add_argument("--chunk", type=int, default=100, help="c5: steps per serving round")
    ap.add_argument("--slots", type=int, default=16, help="c5: most jobs on the GPU at once")
    ap.add_argument("--max-atoms", type=int, default=100000, help="c5: largest job size")
    ap.add_argument("--no-autotune", action="store_true",
                    help="c4: keep the default launch shapes (no timing pass)")
    ap.add_argument("--nve", action="store_true",
                    help="energy-conservation run: Langevin equilibration, then --steps NVE steps")
    ap.add_argument("--equil-md-steps", type=int, default=EQUIL_STEPS,
                    help="MD: untimed Langevin equilibration at friction %g/ps before the warmup "
                         "(reported in config.equilibration; 0 = none)" % EQUIL_FRICTION)
    ap.add_argument("--equil-steps", type=int, default=10000)
    ap.add_argument("--equil-friction", type=float, default=5.0)
    ap.add_argument("--sample-every", type=int, default=500)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device-to-device copy (PMC byte calibration)")
    a = ap.parse_args()
    global DT_PS
    DT_PS = a.dt_fs / 1000.0
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build runs one system per process on a "
              "single MI355X; --gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration()
        return
    if a.nve:
        bench_nve(a)
        return
    if a.config == "c4":
        # hardware queues are fixed when the HIP runtime starts in this process
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(max(a.hw_queues, 1), 32))
        bench_batched(a)
        return
    if a.config == "c5":
        print("c5 (serving) is disabled in this build: after the stream-served jobs of a 40-job "
              "stream all finished on the GPU, the process ended in an unresolved host "
              "segmentation fault at the first job of the one-at-a-time comparison run",
              file=sys.stderr)
        sys.exit(2)

    import ctypes

    import numpy as np
    import torch
    from apocharmm_amd import _lib, systems
    from apocharmm_amd.force_manager import ForceManager
    from apocharmm_amd.integrator import LangevinThermostatIntegrator, maxwell_velocities

    spec = config_spec(a.config)
    n = len(spec["charges"])
    torch.cuda.set_device(0)
    fm = ForceManager(spec, a.precision).initialize()
    md = a.step == "md"
    equil = None
    if md:
        vseed = VSEED[a.config]
        vel = maxwell_velocities(spec["masses"], TEMP, vseed)
        if a.prune_every is None:
            a.prune_every = 10 if n >= PRUNE_MIN_ATOMS else 0
        xyz = np.asarray(spec["xyz"], np.float64)
        launch = None
        if a.equil_md_steps > 0:
            # equilibration (untimed): the lattice start relaxes to 300 K; its
            # first list cycles also time the launch orders of the force
            # branches, and the timed integrator keeps the chosen one
            eq = LangevinThermostatIntegrator(DT_PS, friction=EQUIL_FRICTION, temperature=TEMP,
                                              seed=SEED + 1)
            eq.setNonbondedListUpdateFrequency(REBUILD_EVERY)
            eq.setListPruneFrequency(a.prune_every)
            eq.setCharmmContext(fm, xyz, vel)
            t_eq0 = time.perf_counter()
            eq.propagate(a.equil_md_steps)
            st = eq.getState()
            t_eq = time.perf_counter() - t_eq0
            T_eq = eq.getTemperature()
            launch = eq.getLaunchInfo()
            eq.close()
            xyz, vel = st["xyz"], st["vel"]
            equil = {"steps": a.equil_md_steps, "friction_per_ps": EQUIL_FRICTION,
                     "T_end_K": round(float(T_eq[1]), 2), "wall_s": round(t_eq, 3),
                     "launch_setting": launch["setting"],
                     "launch_trial_cycle_ms": [round(v, 4) for v in launch["trial_ms"]]}
        lt = LangevinThermostatIntegrator(DT_PS, friction=FRICTION, temperature=TEMP, seed=SEED)
        lt.setNonbondedListUpdateFrequency(REBUILD_EVERY)
        lt.setListPruneFrequency(a.prune_every)
        if launch is not None and launch["setting"] >= 0:
            lt.setLaunchSetting(launch["setting"])
        lt.setCharmmContext(fm, xyz, vel)
        lt.initialize()

        def run(k):
            lt.propagate(k)
    else:
        xyzq = torch.tensor(systems.xyzq_of(spec), device="cuda")

        def run(k):
            for i in range(k):
                fm.calcForce(xyzq, reset=(i % REBUILD_EVERY == 0))

    run(a.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(a.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    # second pass of the same steps with HIP-event kernel timers (eager launches)
    _lib.check(_lib.lib().apo_nb_profile(fm._handle, 1))
    run(a.steps)
    torch.cuda.synchronize()
    ms = (ctypes.c_double * 10)()
    cnt = (ctypes.c_int64 * 10)()
    _lib.check(_lib.lib().apo_nb_profile_read(fm._handle, ms, cnt))
    _lib.check(_lib.lib().apo_nb_profile(fm._handle, 0))
    kern = {name: (ms[k] / max(cnt[k], 1), int(cnt[k])) for k, name in enumerate(
        ("pair", "pme_spread", "pme_fft_conv", "pme_gather", "list_rebuild", "calc_force",
         "bonded_14", "sort_finalize", "integrator", "center_com"))}
    ms_step = wall / a.steps * 1e3
    value = a.steps * DT_PS / 1000.0 / (wall / 86400.0)

    pairs = count_pairs(spec, float(spec["ctofnb"]))
    pair_ms = kern["pair"][0]
    achieved = pairs * FLOP_PER_PAIR / (pair_ms * 1e-3) / 1e12
    stats = fm.listStats()
    thermo = None
    list_check = None
    if md:
        th = lt._state(())["thermo"]
        thermo = {"T_K": round(float(th[2]), 2), "T_avg_K": round(float(th[4]), 2),
                  "steps_total": int(lt.getCurrentStep())}
        list_check = check_list(lt, fm, spec, a.prune_every)
    what = "NVT Langevin + SETTLE step" if md else "force evaluation only (no integrator)"
    workload = "%s_%d_atoms_pme%d" % ("jac_dhfr" if a.config == "jac" else a.config + "_tip3p", n,
                                       int(spec["fft_grid"][0]))
    if not md:
        workload += "_force"
    if a.precision != "mixed":
        workload += "_" + a.precision
    out = {
        "metric": "ns/day @98k-atom PME water (%s, 2 fs)" % ("NVT Langevin MD" if md else
                                                              "force path"),
        "value": round(value, 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+int64" if a.precision == "mixed" else "fp64+int64",
        "data": "synthetic (seeded TIP3P lattice, random orientations, Maxwell velocities)",
        "config": {"workload": workload,
                   "natom": n, "box_A": float(spec["box"][0]),
                   "pme_grid": [int(v) for v in spec["fft_grid"]],
                   "cutoffs_A": [float(spec["ctonnb"]), float(spec["ctofnb"]), float(spec["cutnb"])],
                   "kappa": float(spec["kappa"]), "dt_fs": DT_PS * 1000.0,
                   "list_rebuild_every": REBUILD_EVERY,
                   "list_prune_every": a.prune_every if md else 0,
                   "step": what, "friction_per_ps": FRICTION if md else None,
                   "temperature_K": TEMP if md else None, "precision": a.precision,
                   "parallelism": "single GPU"},
        "kernels_ms_avg": {k: round(v[0], 5) for k, v in kern.items()},
        "kernel_launches": {k: v[1] for k, v in kern.items()},
        "list": stats,
        "list_check": list_check,
        "thermo": thermo,
        "roofline": {"bound": "valu", "engine": "fp32 VALU (no dense contraction on this path)",
                     "kernel": "k_pair (direct-space pair kernel)", "achieved": round(achieved, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 5),
                     "peak_nonpacked": PEAK_FP32_NONPACKED_TFLOPS,
                     "frac_nonpacked": round(achieved / PEAK_FP32_NONPACKED_TFLOPS, 5),
                     "traffic": pair_traffic(workload),
                     "work": "%d pairs inside r_off x %d flop" % (pairs, FLOP_PER_PAIR)},
    }
    if md:
        out["config"]["equilibration"] = equil
    if not a.no_cpu_baseline:
        threads, host = host_cpus()
        out["cpu_baseline"] = cpu_baseline(spec, threads, md)
        out["cpu_baseline"].update(host)
    else:
        out["cpu_baseline"] = None
    if md:
        lt.close()
    fm.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    # a crash in native code prints the Python stack of every thread to stderr
    import faulthandler
    faulthandler.enable()
    main()
