"""ForceManager: Python mirror of the reference's ForceManager surface
(include/ForceManager.h:200-476, src/ForceManager.cu) over the C ABI.

Method names and argument meaning follow the reference:
setBoxDimensions / setFFTGrid / setKappa / setCutoff / setCtonnb / setCtofnb /
setPmeSplineOrder, initialize(), resetNeighborList(xyzq),
calcForce(xyzq, reset, calcEnergy, calcVirial), getForces(), getForceStride(),
getEnergyComponents(), getPotentialEnergy(), getVirial().

Device memory comes from torch (plumbing only); every force evaluation runs in
libapocharmm_amd.so.  Forces are the gradient dU/dx, as in the reference.
"""
import ctypes

import numpy as np

from . import _lib
from .systems import lj_table


def _torch():
    import torch
    return torch


class ForceManager:
    def __init__(self, spec, precision="mixed"):
        self.spec = spec
        self.natom = len(spec["charges"])
        self._box = [float(v) for v in spec["box"]]
        self._grid = [int(v) for v in spec["fft_grid"]]
        self._kappa = float(spec["kappa"])
        self._cutnb = float(spec["cutnb"])
        self._ctofnb = float(spec["ctofnb"])
        self._ctonnb = float(spec["ctonnb"])
        self._e14fac = float(spec.get("e14fac", 1.0))
        self._order = int(spec.get("order", 4))
        self._precision = _lib.PRECISION_FP64 if precision in ("fp64", 1) else _lib.PRECISION_MIXED
        self._handle = ctypes.c_void_p()
        self._keep = []
        self._views = []  # subscribed plugin forces
        self.stream = None
        self.force = None
        self.energy = None
        self.virial = None

    # ---- setters (include/ForceManager.h:200-231) --------------------------
    def setBoxDimensions(self, box):
        self._box = [float(v) for v in box]
        if self._handle:
            arr = (ctypes.c_double * 3)(*self._box)
            _lib.check(_lib.lib().apo_nb_set_box(self._handle, arr))
        for force, _ in self._views:
            force.setBoxDimensions(list(self._box))

    def setFFTGrid(self, nx, ny, nz):
        self._grid = [int(nx), int(ny), int(nz)]

    def setKappa(self, kappa):
        self._kappa = float(kappa)

    def setCutoff(self, cutnb):
        self._cutnb = float(cutnb)

    def setCtonnb(self, ctonnb):
        self._ctonnb = float(ctonnb)

    def setCtofnb(self, ctofnb):
        self._ctofnb = float(ctofnb)

    def setPmeSplineOrder(self, order):
        self._order = int(order)

    def getBoxDimensions(self):
        return list(self._box)

    # ---- plugin forces (ForceManager::subscribe, include/ForceManager.h:427-444) ----
    def subscribe(self, force, tag=""):
        """Adds a plugin force (ForceView, include/ForceManager.h:37-105).  It
        provides initialize(natom, box, stride), clear(),
        calcForce(xyzq_tensor, calcEnergy, calcVirial), setBoxDimensions(box),
        getForce() -> int64 device tensor [3 * stride] (2^-40 gradient) and
        getEnergy() -> float64 device tensor [1], and works on the current
        stream.  After every calcForce its gradient and energy are added to
        the totals (src/ForceManager.cu:678-682, 774-778); like the
        reference, the total virial leaves it out."""
        self._views.append((force, tag))
        if self._handle:
            force.initialize(self.natom, list(self._box), self.stride)

    # ---- set-up --------------------------------------------------------------
    def initialize(self, stream=None):
        torch = _torch()
        L = _lib.lib()
        s = self.spec
        p = _lib.NbParams()
        for d in range(3):
            p.box[d] = self._box[d]
            p.fft_grid[d] = self._grid[d]
        p.kappa, p.cutnb, p.ctofnb, p.ctonnb = self._kappa, self._cutnb, self._ctofnb, self._ctonnb
        p.e14fac, p.pme_order, p.precision = self._e14fac, self._order, self._precision

        def i32(a, w):
            a = np.ascontiguousarray(np.asarray(a, np.int32).reshape(-1, w))
            self._keep.append(a)
            return a.shape[0], a.ctypes.data_as(_lib._i32p)

        def f32(a, w):
            a = np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, w))
            self._keep.append(a)
            return a.ctypes.data_as(_lib._f32p)

        t = _lib.NbTopology()
        t.natom = self.natom
        eps, rh = s["eps"], s["rmin_half"]
        t.ntypes = len(eps)
        _, t.vdwtype = i32(s["vdwtype"], 1)
        # explicit (c6, c12) tables (NBFIX applied) take precedence over the
        # combination rule
        t.vdwtab = f32(s["vdwtab"] if "vdwtab" in s else lj_table(eps, rh), 2)
        t.vdwtab14 = f32(s["vdwtab14"] if "vdwtab14" in s else
                         lj_table(s.get("eps14", eps), s.get("rmin_half14", rh)), 2)
        t.nbond, t.bonds = i32(s["bonds"], 2)
        t.bond_par = f32(s["bond_par"], 2)
        t.nureyb, t.ureyb = i32(s["ureyb"], 2)
        t.ub_par = f32(s["ub_par"], 2)
        t.nangle, t.angles = i32(s["angles"], 3)
        t.angle_par = f32(s["angle_par"], 2)
        t.ndihe, t.dihe = i32(s["dihe"], 4)
        t.dihe_par = f32(s["dihe_par"], 3)
        t.nimdihe, t.imdihe = i32(s["imp"], 4)
        t.imdihe_par = f32(s["imp_par"], 3)
        _lib.check(L.apo_nb_create(ctypes.byref(p), ctypes.byref(t), ctypes.byref(self._handle)))
        st = ctypes.c_int32()
        _lib.check(L.apo_nb_force_stride(self._handle, ctypes.byref(st)))
        self.stride = st.value
        dev = torch.device("cuda", torch.cuda.current_device())
        self.force = torch.zeros(3 * self.stride, dtype=torch.int64, device=dev)
        self.energy = torch.zeros(_lib.NUM_ENERGY, dtype=torch.float64, device=dev)
        self.virial = torch.zeros(9, dtype=torch.float64, device=dev)
        self._fdouble = torch.zeros(3 * self.stride, dtype=torch.float64, device=dev)
        self.stream = stream
        for force, _ in self._views:
            force.initialize(self.natom, list(self._box), self.stride)
        return self

    def _stream_ptr(self):
        torch = _torch()
        s = self.stream if self.stream is not None else torch.cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    def getForceStride(self):
        return self.stride

    # ---- force evaluation ------------------------------------------------------
    def resetNeighborList(self, xyzq):
        _lib.check(_lib.lib().apo_nb_reset_list(self._handle, ctypes.c_void_p(xyzq.data_ptr()),
                                                self._stream_ptr()))

    def calcForce(self, xyzq, reset=False, calcEnergy=False, calcVirial=False, terms=0):
        """xyzq: device float32 tensor (natom, 4), or the integer address of
        such a device array (e.g. LangevinThermostatIntegrator.getXYZQ()).
        Overwrites self.force with the int64 2^-40 gradient; energies/virial
        when requested."""
        if isinstance(xyzq, int):
            if self._views:
                raise ValueError("plugin forces need the xyzq tensor, not its address")
            ptr = xyzq
        else:
            assert xyzq.dtype == _torch().float32 and xyzq.is_cuda and xyzq.is_contiguous()
            ptr = xyzq.data_ptr()
        flags = terms
        if reset:
            flags |= _lib.APO_RESET_LIST
        if calcEnergy:
            flags |= _lib.APO_CALC_ENERGY
        if calcVirial:
            flags |= _lib.APO_CALC_VIRIAL
        for force, _ in self._views:
            force.clear()
            force.calcForce(xyzq, calcEnergy, calcVirial)
        _lib.check(_lib.lib().apo_nb_calc_force(
            self._handle, ctypes.c_void_p(ptr), flags,
            ctypes.c_void_p(self.force.data_ptr()), self.stride,
            ctypes.c_void_p(self.energy.data_ptr()), ctypes.c_void_p(self.virial.data_ptr()),
            self._stream_ptr()))
        for force, _ in self._views:
            _lib.check(_lib.lib().apo_nb_add_force(
                ctypes.c_void_p(force.getForce().data_ptr()), self.natom, self.stride,
                ctypes.c_void_p(self.force.data_ptr()), self._stream_ptr()))
            if calcEnergy:
                _lib.check(_lib.lib().apo_nb_add_energy(
                    ctypes.c_void_p(force.getEnergy().data_ptr()),
                    ctypes.c_void_p(self.energy.data_ptr()), self._stream_ptr()))

    def getForces(self):
        """Gradient as a (natom, 3) float64 device tensor (Force<double>)."""
        _lib.check(_lib.lib().apo_nb_force_to_double(
            ctypes.c_void_p(self.force.data_ptr()), self.natom, self.stride,
            ctypes.c_void_p(self._fdouble.data_ptr()), self._stream_ptr()))
        return self._fdouble.view(3, self.stride)[:, : self.natom].t()

    def getForcesFixed(self):
        """Raw int64 2^-40 gradient, (natom, 3)."""
        return self.force.view(3, self.stride)[:, : self.natom].t()

    def getEnergyComponents(self):
        e = self.energy.cpu().numpy()
        return {k: float(v) for k, v in zip(_lib.ENERGY_NAMES, e)}

    def getPotentialEnergy(self):
        return float(self.energy[_lib.NUM_ENERGY - 1].item())

    def getVirial(self):
        return self.virial.cpu().numpy().reshape(3, 3)

    def listStats(self):
        out = (ctypes.c_int64 * 6)()
        _lib.check(_lib.lib().apo_nb_list_stats(self._handle, out, self._stream_ptr()))
        keys = ("cells", "entries", "items", "pairs_in_list", "columns", "max_column")
        return dict(zip(keys, list(out)))

    def close(self):
        if self._handle:
            _lib.check(_lib.lib().apo_nb_destroy(self._handle))
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
