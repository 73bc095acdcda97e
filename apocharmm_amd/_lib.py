"""ctypes binding of libapocharmm_amd.so (the C ABI in include/apocharmm_amd.h).

The shared library is built in-tree by apocharmm_amd/build.py.  There is no
fallback: if it is missing, loading raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libapocharmm_amd.so")

APO_OK, APO_INVALID_ARGUMENT, APO_RUNTIME_ERROR, APO_GPU_ERROR, APO_NOT_INITIALIZED, \
    APO_NOT_IMPLEMENTED = range(6)

APO_CALC_ENERGY = 1 << 0
APO_CALC_VIRIAL = 1 << 1
APO_RESET_LIST = 1 << 2
APO_ACCUMULATE = 1 << 3
APO_TERM_BONDED = 1 << 8
APO_TERM_DIRECT = 1 << 9
APO_TERM_PAIRS14 = 1 << 10
APO_TERM_RECIP = 1 << 11
APO_TERM_SELF = 1 << 12
APO_TERM_ALL = 0x1F00

ENERGY_NAMES = ("bond", "angle", "ureyb", "dihe", "imdihe", "ewks", "ewse", "ewex", "elec", "vdw",
                "epot")
NUM_ENERGY = len(ENERGY_NAMES)

PRECISION_MIXED = 0
PRECISION_FP64 = 1


class ApoError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("apocharmm_amd status %d: %s" % (status, msg))
        self.status = status


class ApoInvalidArgument(ApoError, ValueError):
    pass


class NbParams(ctypes.Structure):
    _fields_ = [("box", ctypes.c_double * 3), ("kappa", ctypes.c_double),
                ("cutnb", ctypes.c_double), ("ctofnb", ctypes.c_double),
                ("ctonnb", ctypes.c_double), ("e14fac", ctypes.c_double),
                ("fft_grid", ctypes.c_int32 * 3), ("pme_order", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("reserved", ctypes.c_int32)]


_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)


class NbTopology(ctypes.Structure):
    _fields_ = [("natom", ctypes.c_int32), ("ntypes", ctypes.c_int32),
                ("vdwtype", _i32p), ("vdwtab", _f32p), ("vdwtab14", _f32p),
                ("nbond", ctypes.c_int32), ("bonds", _i32p), ("bond_par", _f32p),
                ("nureyb", ctypes.c_int32), ("ureyb", _i32p), ("ub_par", _f32p),
                ("nangle", ctypes.c_int32), ("angles", _i32p), ("angle_par", _f32p),
                ("ndihe", ctypes.c_int32), ("dihe", _i32p), ("dihe_par", _f32p),
                ("nimdihe", ctypes.c_int32), ("imdihe", _i32p), ("imdihe_par", _f32p)]


class LtParams(ctypes.Structure):
    _fields_ = [("timestep", ctypes.c_double), ("temperature", ctypes.c_double),
                ("friction", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("nbfreq", ctypes.c_int32), ("comfreq", ctypes.c_int32),
                ("nwater", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("waters", _i32p), ("water_geom", ctypes.c_double * 4),
                ("mass", ctypes.POINTER(ctypes.c_double)), ("charge", _f32p),
                ("list_policy", ctypes.c_int32), ("prune_every", ctypes.c_int32),
                ("launch_setting", ctypes.c_int32), ("nshake", ctypes.c_int32),
                ("shake_groups", _i32p), ("shake_r2", ctypes.POINTER(ctypes.c_double))]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s is not built; run `python -m apocharmm_amd.build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64p, dp = ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), \
            ctypes.POINTER(ctypes.c_double)
        L.apo_last_error.restype = ctypes.c_char_p
        L.apo_version.restype = ctypes.c_char_p
        L.apo_nb_create.argtypes = [ctypes.POINTER(NbParams), ctypes.POINTER(NbTopology),
                                    ctypes.POINTER(vp)]
        L.apo_nb_destroy.argtypes = [vp]
        L.apo_nb_set_box.argtypes = [vp, dp]
        L.apo_nb_reset_list.argtypes = [vp, vp, vp]
        L.apo_nb_calc_force.argtypes = [vp, vp, i32, vp, i32, vp, vp, vp]
        L.apo_nb_force_stride.argtypes = [vp, _i32p]
        L.apo_nb_num_atoms.argtypes = [vp, _i32p]
        L.apo_nb_force_to_double.argtypes = [vp, i32, i32, vp, vp]
        L.apo_nb_add_force.argtypes = [vp, i32, i32, vp, vp]
        L.apo_nb_add_energy.argtypes = [vp, vp, vp]
        L.apo_nb_list_stats.argtypes = [vp, i64p, vp]
        L.apo_topology_exclusions.argtypes = [i32, i32, vp, _i32p, vp, vp, vp, vp]
        L.apo_lt_create.argtypes = [vp, ctypes.POINTER(LtParams), vp, vp, ctypes.POINTER(vp)]
        L.apo_lt_destroy.argtypes = [vp]
        L.apo_lt_initialize.argtypes = [vp, vp]
        L.apo_lt_propagate.argtypes = [vp, i32, vp]
        L.apo_lt_propagate_batch.argtypes = [ctypes.POINTER(vp), i32, i32, vp]
        L.apo_lt_batch_create.argtypes = [ctypes.POINTER(vp), i32, i32, ctypes.POINTER(vp)]
        L.apo_lt_batch_propagate.argtypes = [vp, i32, vp]
        L.apo_lt_batch_destroy.argtypes = [vp]
        L.apo_lt_batch_autotune.argtypes = [vp, vp, dp]
        L.apo_lt_get_state.argtypes = [vp, vp, vp, vp, vp, vp, i64p]
        L.apo_lt_xyzq.argtypes = [vp, ctypes.POINTER(vp)]
        L.apo_lt_list_info.argtypes = [vp, dp]
        L.apo_lt_launch_info.argtypes = [vp, dp]
        for name in ("apo_nb_create", "apo_nb_destroy", "apo_nb_set_box", "apo_nb_reset_list",
                     "apo_nb_calc_force", "apo_nb_force_stride", "apo_nb_num_atoms",
                     "apo_nb_force_to_double", "apo_nb_add_force", "apo_nb_add_energy",
                     "apo_nb_list_stats", "apo_topology_exclusions",
                     "apo_lt_create", "apo_lt_destroy", "apo_lt_initialize", "apo_lt_propagate",
                     "apo_lt_propagate_batch", "apo_lt_batch_create", "apo_lt_batch_propagate",
                     "apo_lt_batch_destroy", "apo_lt_batch_autotune",
                     "apo_lt_get_state", "apo_lt_xyzq", "apo_lt_list_info", "apo_lt_launch_info"):
            getattr(L, name).restype = ctypes.c_int
        _LIB = L
    return _LIB


def check(status):
    if status != APO_OK:
        msg = lib().apo_last_error().decode(errors="replace")
        if status == APO_INVALID_ARGUMENT:
            raise ApoInvalidArgument(status, msg)
        raise ApoError(status, msg)


def version():
    return lib().apo_version().decode()
