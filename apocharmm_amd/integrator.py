"""LangevinThermostatIntegrator: Python mirror of the reference's
CudaLangevinThermostatIntegrator (include/CudaLangevinThermostatIntegrator.h,
src/CudaLangevinThermostatIntegrator.cu) over the C ABI (apo_lt_*).

    fm = ForceManager(spec).initialize()
    lt = LangevinThermostatIntegrator(0.002)          # ps
    lt.setThermostatFriction(1.0)                     # 1/ps
    lt.setReferenceTemperature(300.0)
    lt.setThermostatRngSeed(2003)
    lt.setCharmmContext(fm, xyz, vel)
    lt.propagate(1000)

The whole step (force path, leapfrog Langevin update, SETTLE, kinetic energy)
runs in libapocharmm_amd.so; this class only carries parameters and copies
state out.  Velocities are in A per AKMA time unit, as in the reference.
"""
import ctypes

import numpy as np

from . import _lib

TIMFAC = 0.0488882129       # AKMA time unit in ps (src/CudaIntegrator.cu:22)
KBOLTZ = 1.987191e-03       # kcal/mol/K (include/Constants.h:17)


def _torch():
    import torch
    return torch


class LangevinThermostatIntegrator:
    def __init__(self, timeStep, friction=0.0, temperature=300.0, seed=0):
        self._dt = float(timeStep)
        self._friction = float(friction)
        self._temperature = float(temperature)
        self._seed = int(seed)
        self._nbfreq = 20
        self._comfreq = 1000
        self._use_graph = True
        self._list_policy = 0
        self._prune_every = 0
        self._launch_setting = 0
        self._shake = True
        self._handle = ctypes.c_void_p()
        self._fm = None
        self._keep = []

    # ---- setters (src/CudaLangevinThermostatIntegrator.cu:41-74) -------------------
    def setThermostatFriction(self, friction):
        self._friction = float(friction)

    def setReferenceTemperature(self, temperature):
        self._temperature = float(temperature)

    def setThermostatRngSeed(self, seed):
        self._seed = int(seed)

    def setNonbondedListUpdateFrequency(self, nfreq):
        self._nbfreq = int(nfreq)

    def setRemoveCenterOfMassFrequency(self, freq):
        self._comfreq = int(freq)

    def setUseGraph(self, use):
        self._use_graph = bool(use)

    def setListUpdatePolicy(self, policy):
        """"fixed": rebuild every nbfreq steps (the reference's cadence);
        "displacement": also when an atom may have moved more than half the
        list buffer since the last build."""
        self._list_policy = {"fixed": 0, "displacement": 1}[policy]

    def setListPruneFrequency(self, every):
        """Prunes the pair kernel's list every `every` steps between builds
        at r_off + (cutnb - r_off) * every / nbfreq (no counterpart in the
        reference; needs the fixed cadence); 0 turns it off."""
        self._prune_every = int(every)

    def setLaunchSetting(self, setting):
        """None: the first list cycles time the four launch orders of the
        force branches and keep the fastest (waits for the GPU on the host
        during those cycles); 0..3: use that order without a trial (bit 0 =
        PME kernels at raised wave issue priority, bit 1 = pair kernel
        captured first).  No result bit depends on it."""
        self._launch_setting = 0 if setting is None else int(setting) + 1

    def useHolonomicConstraints(self, use):
        """SHAKE for the system's bonds to hydrogen (its shake_groups) on or
        off; the SETTLE waters stay rigid either way."""
        self._shake = bool(use)

    def getLaunchInfo(self):
        """dict: setting (0..3, -1 while the trial runs), trial_cycles, and
        the fastest timed cycle (ms) of each setting."""
        out = (ctypes.c_double * 6)()
        _lib.check(_lib.lib().apo_lt_launch_info(self._handle, out))
        return dict(setting=int(out[0]), trial_cycles=int(out[1]), trial_ms=list(out[2:6]))

    def getTimeStep(self):
        return self._dt

    def getThermostatFriction(self):
        return self._friction

    def getReferenceTemperature(self):
        return self._temperature

    def getThermostatRngSeed(self):
        return self._seed

    # ---- binding (CudaIntegrator::setCharmmContext) ------------------------------
    def setCharmmContext(self, force_manager, xyz, vel=None, water_geom=None):
        """Binds to an initialised ForceManager.  xyz, vel: (natom, 3) float64
        host arrays (A, A/AKMA).  Masses, charges and the water list come from
        the ForceManager's system spec."""
        self.close()
        spec = force_manager.spec
        n = force_manager.natom
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(n, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float64).reshape(n, 3)
        mass = np.ascontiguousarray(spec["masses"], np.float64)
        charge = np.ascontiguousarray(spec["charges"], np.float32)
        waters = np.ascontiguousarray(spec.get("waters", np.zeros((0, 3))), np.int32).reshape(-1)
        p = _lib.LtParams()
        p.timestep = self._dt
        p.temperature = self._temperature
        p.friction = self._friction
        p.seed = self._seed
        p.nbfreq = self._nbfreq
        p.comfreq = self._comfreq
        p.nwater = len(waters) // 3
        p.use_graph = 1 if self._use_graph else 0
        p.list_policy = self._list_policy
        p.prune_every = self._prune_every
        p.launch_setting = self._launch_setting
        p.waters = waters.ctypes.data_as(_lib._i32p)
        if water_geom is not None:
            for k in range(4):
                p.water_geom[k] = float(water_geom[k])
        p.mass = mass.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        p.charge = charge.ctypes.data_as(_lib._f32p)
        # bonds to hydrogen held by SHAKE when the system lists them (the
        # reference's context turns holonomic constraints on by default)
        sg = np.ascontiguousarray(spec.get("shake_groups", np.zeros((0, 4))), np.int32).reshape(-1)
        sr2 = np.ascontiguousarray(spec.get("shake_r2", np.zeros(0)), np.float64).reshape(-1)
        if not self._shake:
            sg, sr2 = sg[:0], sr2[:0]
        p.nshake = len(sr2)
        p.shake_groups = sg.ctypes.data_as(_lib._i32p)
        p.shake_r2 = sr2.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self._keep = [waters, mass, charge, sg, sr2]
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().apo_lt_create(
            force_manager._handle, ctypes.byref(p), xyz.ctypes.data_as(ctypes.c_void_p),
            None if vel is None else vel.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h)))
        self._handle = h
        self._fm = force_manager
        self.natom = n
        self.ndegf = 3 * n - 3 - 3 * p.nwater - int((sg.reshape(-1, 4)[:, 1:] >= 0).sum())
        return self

    def _stream(self):
        torch = _torch()
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    # ---- dynamics ----------------------------------------------------------------
    def initialize(self):
        _lib.check(_lib.lib().apo_lt_initialize(self._handle, self._stream()))

    def propagate(self, numSteps):
        _lib.check(_lib.lib().apo_lt_propagate(self._handle, int(numSteps), self._stream()))

    # ---- state ---------------------------------------------------------------------
    def _state(self, which):
        n = self.natom
        bufs = {k: None for k in ("xyz", "vel", "dprev", "grad")}
        out = {}
        for k in which:
            if k in bufs:
                out[k] = np.zeros((n, 3))
                bufs[k] = out[k].ctypes.data_as(ctypes.c_void_p)
        thermo = np.zeros(6)
        step = ctypes.c_int64()
        _torch().cuda.current_stream().synchronize()
        _lib.check(_lib.lib().apo_lt_get_state(
            self._handle, bufs["xyz"], bufs["vel"], bufs["dprev"], bufs["grad"],
            thermo.ctypes.data_as(ctypes.c_void_p), ctypes.byref(step)))
        out["thermo"] = thermo
        out["step"] = step.value
        return out

    def getCoordinates(self):
        return self._state(("xyz",))["xyz"]

    def getVelocities(self):
        return self._state(("vel",))["vel"]

    def getCoordsDelta(self):
        """Last half-step displacement (the reference's m_CoordsDeltaPrevious
        after the step's final copy)."""
        return self._state(("dprev",))["dprev"]

    def getGradient(self):
        """Gradient dU/dx (natom, 3) the last step used."""
        return self._state(("grad",))["grad"]

    def getState(self):
        return self._state(("xyz", "vel", "dprev", "grad"))

    def getKineticEnergy(self):
        """[three-estimate mean, on-step] (getKineticEnergy)."""
        return self._state(())["thermo"][:2].copy()

    def getTemperature(self):
        return self._state(())["thermo"][2:4].copy()

    def getAverageTemperature(self):
        return self._state(())["thermo"][4:6].copy()

    def getCurrentStep(self):
        return self._state(())["step"]

    def getListInfo(self):
        """dict: max_displacement (A, since the last list build),
        max_step_displacement (A), triggered_rebuilds, half_buffer (A)."""
        out = (ctypes.c_double * 4)()
        _lib.check(_lib.lib().apo_lt_list_info(self._handle, out))
        return dict(max_displacement=out[0], max_step_displacement=out[1],
                    triggered_rebuilds=int(out[2]), half_buffer=out[3])

    def getXYZQ(self):
        """Device pointer (int) of the float xyzq the force path reads."""
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().apo_lt_xyzq(self._handle, ctypes.byref(p)))
        return p.value

    def close(self):
        if self._handle:
            _lib.check(_lib.lib().apo_lt_destroy(self._handle))
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def propagate_batch(integrators, numSteps):
    """Advances several LangevinThermostatIntegrator objects (each with its own
    ForceManager) by numSteps together, their kernels overlapping on the GPU
    (apo_lt_propagate_batch).  Each trajectory is bitwise the one propagate()
    gives the system alone."""
    torch = _torch()
    arr = (ctypes.c_void_p * len(integrators))(*[lt._handle.value for lt in integrators])
    _lib.check(_lib.lib().apo_lt_propagate_batch(
        arr, len(integrators), int(numSteps),
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))


class LangevinBatch:
    """Persistent batch over several integrators (apo_lt_batch_*): one graph
    launch per batch step, systems chained into `lanes` parallel branches."""

    def __init__(self, integrators, lanes=0):
        self.members = list(integrators)
        arr = (ctypes.c_void_p * len(self.members))(*[lt._handle.value for lt in self.members])
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().apo_lt_batch_create(arr, len(self.members), int(lanes),
                                                  ctypes.byref(h)))
        self._handle = h

    def propagate(self, numSteps):
        torch = _torch()
        _lib.check(_lib.lib().apo_lt_batch_propagate(
            self._handle, int(numSteps), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def autotune(self):
        """Picks the pair-kernel block budget and the FFT block shape by timing
        (apo_lt_batch_autotune); returns the choice and the timings."""
        torch = _torch()
        out = (ctypes.c_double * 9)()
        _lib.check(_lib.lib().apo_lt_batch_autotune(
            self._handle, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), out))
        return dict(pair_blocks_per_cu=out[0], fft_threads=int(out[1]), fft_rows=int(out[2]),
                    fft_cols=int(out[3]), force_ms=out[4], pme_ms=out[5],
                    default_force_ms=out[6], default_pme_ms=out[7], launches_timed=int(out[8]))

    def close(self):
        if self._handle:
            _torch().cuda.synchronize()
            _lib.check(_lib.lib().apo_lt_batch_destroy(self._handle))
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def maxwell_velocities(masses, temperature, seed):
    """Gaussian velocities with variance kB T / m per component (A/AKMA) and
    zero mass-weighted mean, as CharmmContext::assignVelocitiesAtTemperature
    (src/CharmmContext.cu:478-526) produces them (numpy generator here)."""
    m = np.asarray(masses, np.float64)
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(len(m), 3)) * np.sqrt(KBOLTZ * temperature / m)[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    return v
