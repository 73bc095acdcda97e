"""ctypes mirror of the reference's Python binding (python/apocharmm/*.py) over
the reference-compatible C ABI of libapocharmm_amd.so
(include/apocharmm_amd_capi.h).

Class and method names follow python/apocharmm: CharmmPsf, CharmmCrd,
CharmmParameters, ForceManager(psf, parameters), CharmmContext(force_manager),
CudaIntegrator (setTimeStep, setCharmmContext, subscribe, propagate,
initializeFromRestartFile), RestartSubscriber, DcdSubscriber.  The Langevin
thermostat integrator, which the reference's C ABI does not expose, is
CudaLangevinThermostatIntegrator here (the Nose-Hoover class is not built).
Errors raise apocharmm_amd._lib.ApoError (ValueError for invalid arguments).
"""
import ctypes
import os

import numpy as np

from . import _lib

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_d = ctypes.c_double
_i = ctypes.c_int
_u64 = ctypes.c_uint64
_done = False


def _lib_c():
    global _done
    L = _lib.lib()
    if _done:
        return L
    pp = ctypes.POINTER(_vp)
    sig = {
        "apo_charmm_psf_create": [pp, ctypes.c_char_p],
        "apo_charmm_psf_get_num_atoms": [ctypes.POINTER(_sz), _vp],
        "apo_charmm_psf_get_num_bonds": [ctypes.POINTER(_sz), _vp],
        "apo_charmm_psf_get_num_angles": [ctypes.POINTER(_sz), _vp],
        "apo_charmm_psf_get_num_dihedrals": [ctypes.POINTER(_sz), _vp],
        "apo_charmm_psf_get_num_impropers": [ctypes.POINTER(_sz), _vp],
        "apo_charmm_psf_get_num_cross_terms": [ctypes.POINTER(_sz), _vp],
        "apo_charmm_psf_get_segment_identifiers": [ctypes.c_char_p, _sz, _vp],
        "apo_charmm_psf_get_residue_identifiers": [ctypes.POINTER(_i), _sz, _vp],
        "apo_charmm_psf_get_residue_names": [ctypes.c_char_p, _sz, _vp],
        "apo_charmm_psf_get_atom_names": [ctypes.c_char_p, _sz, _vp],
        "apo_charmm_psf_get_atom_types": [ctypes.c_char_p, _sz, _vp],
        "apo_charmm_psf_get_charges": [ctypes.POINTER(_d), _sz, _vp],
        "apo_charmm_psf_get_masses": [ctypes.POINTER(_d), _sz, _vp],
        "apo_charmm_psf_get_net_charge": [ctypes.POINTER(_d), _vp],
        "apo_charmm_psf_get_total_mass": [ctypes.POINTER(_d), _vp],
        "apo_charmm_psf_get_file_name": [ctypes.c_char_p, _sz, _vp],
        "apo_charmm_crd_create": [pp, ctypes.c_char_p],
        "apo_charmm_crd_get_num_atoms": [ctypes.POINTER(_sz), _vp],
        "apo_charmm_crd_get_coordinates": [ctypes.POINTER(_d), _sz, _vp],
        "apo_charmm_parameters_create_from_files": [pp, ctypes.POINTER(ctypes.c_char_p), _sz],
        "apo_force_manager_create": [pp, _vp, _vp],
        "apo_force_manager_set_box_dimensions": [_vp, _d, _d, _d],
        "apo_force_manager_set_fft_grid": [_vp, _i, _i, _i],
        "apo_force_manager_set_kappa": [_vp, _d],
        "apo_force_manager_set_cutoff": [_vp, _d],
        "apo_force_manager_set_ctonnb": [_vp, _d],
        "apo_force_manager_set_ctofnb": [_vp, _d],
        "apo_force_manager_get_nb_system": [_vp, pp],
        "apo_charmm_context_create": [pp, _vp],
        "apo_charmm_context_set_coordinates": [_vp, _vp],
        "apo_charmm_context_set_random_seed_for_velocities": [_vp, _u64],
        "apo_charmm_context_use_holonomic_constraints": [_vp, ctypes.c_bool],
        "apo_charmm_context_assign_velocities_at_temperature": [_vp, _d],
        "apo_charmm_context_compute_temperature": [ctypes.POINTER(_d), _vp],
        "apo_subscriber_set_report_frequency": [_vp, _i],
        "apo_subscriber_get_report_frequency": [ctypes.POINTER(_i), _vp],
        "apo_restart_subscriber_create_with_report_frequency": [pp, ctypes.c_char_p, _i],
        "apo_restart_subscriber_as_subscriber": [pp, _vp],
        "apo_dcd_subscriber_create_with_report_frequency": [pp, ctypes.c_char_p, _i],
        "apo_dcd_subscriber_as_subscriber": [pp, _vp],
        "apo_cuda_integrator_set_time_step": [_vp, _d],
        "apo_cuda_integrator_set_charmm_context": [_vp, _vp],
        "apo_cuda_integrator_subscribe": [_vp, _vp],
        "apo_cuda_integrator_propagate": [_vp, _i],
        "apo_cuda_integrator_initialize_from_restart_file": [_vp, ctypes.c_char_p],
        "apo_cuda_integrator_get_state": [_vp, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_ulonglong)],
        "apo_cuda_langevin_thermostat_integrator_create": [pp, _d],
        "apo_cuda_langevin_thermostat_integrator_set_friction": [_vp, _d],
        "apo_cuda_langevin_thermostat_integrator_set_reference_temperature": [_vp, _d],
        "apo_cuda_langevin_thermostat_integrator_set_rng_seed": [_vp, _u64],
        "apo_cuda_langevin_thermostat_integrator_as_cuda_integrator": [pp, _vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    for name in ("apo_charmm_psf_destroy", "apo_charmm_crd_destroy",
                 "apo_charmm_parameters_destroy", "apo_force_manager_destroy",
                 "apo_charmm_context_destroy", "apo_restart_subscriber_destroy",
                 "apo_dcd_subscriber_destroy", "apo_cuda_langevin_thermostat_integrator_destroy"):
        f = getattr(L, name)
        f.argtypes = [_vp]
        f.restype = None
    _done = True
    return L


def _path(p):
    return os.fspath(p).encode()


class _ApoObject:
    """Owns one handle; close() (or garbage collection) destroys it."""
    _destroy = None

    def __init__(self):
        self._handle = _vp()

    @property
    def handle(self):
        if not self._handle:
            raise RuntimeError("apoCHARMM object has been destroyed")
        return self._handle

    def close(self):
        if self._handle and self._destroy:
            getattr(_lib_c(), self._destroy)(self._handle)
        self._handle = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _strings(fn, handle, n):
    buf = ctypes.create_string_buffer(8 * n)
    _lib.check(fn(buf, 8 * n, handle))
    raw = buf.raw
    return [raw[8 * i:8 * i + 8].decode().replace(" ", "") for i in range(n)]


class CharmmPsf(_ApoObject):
    _destroy = "apo_charmm_psf_destroy"

    def __init__(self, path):
        super().__init__()
        _lib.check(_lib_c().apo_charmm_psf_create(ctypes.byref(self._handle), _path(path)))

    def _count(self, what):
        n = _sz()
        _lib.check(getattr(_lib_c(), "apo_charmm_psf_get_num_" + what)(ctypes.byref(n),
                                                                        self.handle))
        return n.value

    def getNumAtoms(self):
        return self._count("atoms")

    def getNumBonds(self):
        return self._count("bonds")

    def getNumAngles(self):
        return self._count("angles")

    def getNumDihedrals(self):
        return self._count("dihedrals")

    def getNumImpropers(self):
        return self._count("impropers")

    def getNumCrossTerms(self):
        return self._count("cross_terms")

    def getSegmentIdentifiers(self):
        return _strings(_lib_c().apo_charmm_psf_get_segment_identifiers, self.handle,
                        self.getNumAtoms())

    def getResidueNames(self):
        return _strings(_lib_c().apo_charmm_psf_get_residue_names, self.handle, self.getNumAtoms())

    def getAtomNames(self):
        return _strings(_lib_c().apo_charmm_psf_get_atom_names, self.handle, self.getNumAtoms())

    def getAtomTypes(self):
        return _strings(_lib_c().apo_charmm_psf_get_atom_types, self.handle, self.getNumAtoms())

    def getResidueIdentifiers(self):
        n = self.getNumAtoms()
        out = (_i * n)()
        _lib.check(_lib_c().apo_charmm_psf_get_residue_identifiers(out, n, self.handle))
        return list(out)

    def _doubles(self, name):
        n = self.getNumAtoms()
        out = np.zeros(n)
        _lib.check(getattr(_lib_c(), name)(out.ctypes.data_as(ctypes.POINTER(_d)), n,
                                           self.handle))
        return out

    def getCharges(self):
        return self._doubles("apo_charmm_psf_get_charges")

    def getMasses(self):
        return self._doubles("apo_charmm_psf_get_masses")

    def getNetCharge(self):
        v = _d()
        _lib.check(_lib_c().apo_charmm_psf_get_net_charge(ctypes.byref(v), self.handle))
        return v.value

    def getTotalMass(self):
        v = _d()
        _lib.check(_lib_c().apo_charmm_psf_get_total_mass(ctypes.byref(v), self.handle))
        return v.value

    def getFileName(self):
        buf = ctypes.create_string_buffer(4096)
        _lib.check(_lib_c().apo_charmm_psf_get_file_name(buf, 4096, self.handle))
        return buf.value.decode()


class CharmmCrd(_ApoObject):
    _destroy = "apo_charmm_crd_destroy"

    def __init__(self, path):
        super().__init__()
        _lib.check(_lib_c().apo_charmm_crd_create(ctypes.byref(self._handle), _path(path)))

    def getNumAtoms(self):
        n = _sz()
        _lib.check(_lib_c().apo_charmm_crd_get_num_atoms(ctypes.byref(n), self.handle))
        return n.value

    def getCoordinates(self):
        n = self.getNumAtoms()
        out = np.zeros((n, 3))
        _lib.check(_lib_c().apo_charmm_crd_get_coordinates(out.ctypes.data_as(ctypes.POINTER(_d)),
                                                           3 * n, self.handle))
        return out


class CharmmParameters(_ApoObject):
    _destroy = "apo_charmm_parameters_destroy"

    def __init__(self, paths):
        super().__init__()
        if isinstance(paths, (str, bytes, os.PathLike)):
            paths = [paths]
        arr = (ctypes.c_char_p * len(paths))(*[_path(p) for p in paths])
        _lib.check(_lib_c().apo_charmm_parameters_create_from_files(ctypes.byref(self._handle),
                                                                    arr, len(paths)))


class ForceManager(_ApoObject):
    _destroy = "apo_force_manager_destroy"

    def __init__(self, psf, parameters):
        super().__init__()
        _lib.check(_lib_c().apo_force_manager_create(ctypes.byref(self._handle), psf.handle,
                                                     parameters.handle))
        self._keep = (psf, parameters)

    def setBoxDimensions(self, box):
        x, y, z = (float(v) for v in box)
        _lib.check(_lib_c().apo_force_manager_set_box_dimensions(self.handle, x, y, z))

    def setFFTGrid(self, nx, ny, nz):
        _lib.check(_lib_c().apo_force_manager_set_fft_grid(self.handle, nx, ny, nz))

    def setKappa(self, kappa):
        _lib.check(_lib_c().apo_force_manager_set_kappa(self.handle, kappa))

    def setCutoff(self, cutoff):
        _lib.check(_lib_c().apo_force_manager_set_cutoff(self.handle, cutoff))

    def setCtonnb(self, ctonnb):
        _lib.check(_lib_c().apo_force_manager_set_ctonnb(self.handle, ctonnb))

    def setCtofnb(self, ctofnb):
        _lib.check(_lib_c().apo_force_manager_set_ctofnb(self.handle, ctofnb))

    def nbSystem(self):
        """The force path as an apo_nb_system handle (borrowed) for the
        apo_nb_* entries, e.g. apo_nb_calc_force."""
        h = _vp()
        _lib.check(_lib_c().apo_force_manager_get_nb_system(self.handle, ctypes.byref(h)))
        return h


class CharmmContext(_ApoObject):
    _destroy = "apo_charmm_context_destroy"

    def __init__(self, force_manager):
        super().__init__()
        _lib.check(_lib_c().apo_charmm_context_create(ctypes.byref(self._handle),
                                                      force_manager.handle))
        self._fm = force_manager

    def setCoordinates(self, crd):
        _lib.check(_lib_c().apo_charmm_context_set_coordinates(self.handle, crd.handle))

    def setRandomSeedForVelocities(self, seed):
        _lib.check(_lib_c().apo_charmm_context_set_random_seed_for_velocities(self.handle, seed))

    def useHolonomicConstraints(self, use):
        _lib.check(_lib_c().apo_charmm_context_use_holonomic_constraints(self.handle, bool(use)))

    def assignVelocitiesAtTemperature(self, temperature):
        _lib.check(_lib_c().apo_charmm_context_assign_velocities_at_temperature(self.handle,
                                                                                temperature))

    def computeTemperature(self):
        v = _d()
        _lib.check(_lib_c().apo_charmm_context_compute_temperature(ctypes.byref(v), self.handle))
        return v.value


class Subscriber(_ApoObject):
    """Base of the subscribers (python/apocharmm/subscriber.py)."""
    _as = None

    def subscriber_handle(self):
        h = _vp()
        _lib.check(getattr(_lib_c(), self._as)(ctypes.byref(h), self.handle))
        return h

    def setReportFrequency(self, freq):
        _lib.check(_lib_c().apo_subscriber_set_report_frequency(self.subscriber_handle(), freq))

    def getReportFrequency(self):
        v = _i()
        _lib.check(_lib_c().apo_subscriber_get_report_frequency(ctypes.byref(v),
                                                                self.subscriber_handle()))
        return v.value


class RestartSubscriber(Subscriber):
    _destroy = "apo_restart_subscriber_destroy"
    _as = "apo_restart_subscriber_as_subscriber"

    def __init__(self, path, report_frequency=1000):
        super().__init__()
        _lib.check(_lib_c().apo_restart_subscriber_create_with_report_frequency(
            ctypes.byref(self._handle), _path(path), report_frequency))


class DcdSubscriber(Subscriber):
    _destroy = "apo_dcd_subscriber_destroy"
    _as = "apo_dcd_subscriber_as_subscriber"

    def __init__(self, path, report_frequency=1000):
        super().__init__()
        _lib.check(_lib_c().apo_dcd_subscriber_create_with_report_frequency(
            ctypes.byref(self._handle), _path(path), report_frequency))


class CudaIntegrator(_ApoObject):
    """Methods of the reference's CudaIntegrator binding on the integrator's
    base handle."""

    def integrator_handle(self):
        raise NotImplementedError

    def setTimeStep(self, time_step):
        _lib.check(_lib_c().apo_cuda_integrator_set_time_step(self.integrator_handle(), time_step))

    def setCharmmContext(self, context):
        _lib.check(_lib_c().apo_cuda_integrator_set_charmm_context(self.integrator_handle(),
                                                                   context.handle))
        self._ctx = context

    def subscribe(self, subscriber):
        _lib.check(_lib_c().apo_cuda_integrator_subscribe(self.integrator_handle(),
                                                          subscriber.subscriber_handle()))
        self._subs = getattr(self, "_subs", []) + [subscriber]

    def propagate(self, num_steps):
        _lib.check(_lib_c().apo_cuda_integrator_propagate(self.integrator_handle(), num_steps))

    def initializeFromRestartFile(self, path):
        _lib.check(_lib_c().apo_cuda_integrator_initialize_from_restart_file(
            self.integrator_handle(), _path(path)))

    def getState(self):
        """dict of xyz, vel, dprev ((n, 3) arrays) and total_steps."""
        n = self._ctx._fm._keep[0].getNumAtoms()
        out = {k: np.zeros((n, 3)) for k in ("xyz", "vel", "dprev")}
        steps = ctypes.c_ulonglong()
        _lib.check(_lib_c().apo_cuda_integrator_get_state(
            self.integrator_handle(), out["xyz"].ctypes.data, out["vel"].ctypes.data,
            out["dprev"].ctypes.data, ctypes.byref(steps)))
        out["total_steps"] = steps.value
        return out


class CudaLangevinThermostatIntegrator(CudaIntegrator):
    _destroy = "apo_cuda_langevin_thermostat_integrator_destroy"

    def __init__(self, time_step, friction=0.0, temperature=300.0, seed=0):
        super().__init__()
        _lib.check(_lib_c().apo_cuda_langevin_thermostat_integrator_create(
            ctypes.byref(self._handle), time_step))
        self.setFriction(friction)
        self.setReferenceTemperature(temperature)
        self.setRngSeed(seed)

    def integrator_handle(self):
        h = _vp()
        _lib.check(_lib_c().apo_cuda_langevin_thermostat_integrator_as_cuda_integrator(
            ctypes.byref(h), self.handle))
        return h

    def setFriction(self, friction):
        _lib.check(_lib_c().apo_cuda_langevin_thermostat_integrator_set_friction(self.handle,
                                                                                 friction))

    def setReferenceTemperature(self, temperature):
        _lib.check(_lib_c().apo_cuda_langevin_thermostat_integrator_set_reference_temperature(
            self.handle, temperature))

    def setRngSeed(self, seed):
        _lib.check(_lib_c().apo_cuda_langevin_thermostat_integrator_set_rng_seed(self.handle,
                                                                                 seed))
