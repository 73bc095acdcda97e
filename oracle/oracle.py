"""FP64 CPU oracle: Python driver over liboracle.so.

TEST INFRASTRUCTURE ONLY -- the checker for the HIP product path and the timed
CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module; the product library never does.

A "system" here is a plain dict of numpy arrays (see `evaluate`).  Topology
derivation (1-2/1-3/1-4 sets) restates src/CharmmPSF.cu:331-385 and the
inclusion/exclusion split of src/CharmmPSF.cu:235-279; the vdW tables restate
the CHARMM combination rule of src/CharmmParameters.cu:718-844.
"""
import ctypes
import os

import numpy as np

CCELEC = 332.0716
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        i, d = ctypes.c_int, ctypes.c_double
        L.orc_direct.argtypes = [i, _dp, _ip, i, _dp, _dp, d, d, d, ctypes.c_void_p,
                                 ctypes.c_void_p, _dp, _dp, _dp, i]
        L.orc_pairs14.argtypes = [i, ctypes.c_void_p, i, ctypes.c_void_p, _dp, _ip, i, _dp,
                                  d, d, d, d, _dp, _dp, _dp]
        L.orc_recip.argtypes = [i, _dp, _dp, _ip, i, d, _dp, _dp, _dp, i]
        L.orc_self_energy.argtypes = [i, _dp, d]
        L.orc_self_energy.restype = d
        L.orc_bonded.argtypes = [_dp] + [i, ctypes.c_void_p, ctypes.c_void_p] * 5 + [_dp, _dp, _dp]
        L.orc_fft3d.argtypes = [_ip, _dp, i, i]
        L.orc_bspline.argtypes = [i, d, _dp, _dp]
        L.orc_bsp_mod.argtypes = [i, i, _dp]
        u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
        u64 = ctypes.c_uint64
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_langevin_normals.argtypes = [u64, u64, i, _dp]
        L.orc_settle_geometry.argtypes = [_dp, _dp]
        L.orc_settle.argtypes = [i, _ip, _dp, _dp, _dp]
        L.orc_shake_water.argtypes = [i, _ip, _dp, _dp, _dp, d, i]
        L.orc_shake_groups.argtypes = [i, _ip, _dp, _dp, _dp, _dp, d, i]
        L.orc_lt_init_constrain.argtypes = [i, i, _ip, _dp, i, _ip, _dp, _dp, d, _dp, _dp, _dp]
        L.orc_lt_init_backstep.argtypes = [i, _dp, i, _ip, _dp, i, _ip, _dp, d, d, _dp, _dp, _dp,
                                           _dp, _dp]
        L.orc_lt_remove_com.argtypes = [i, _dp, _dp]
        L.orc_lt_step.argtypes = [i, _dp, i, _ip, _dp, i, _ip, _dp, d, d, d, u64, u64, _dp, _dp,
                                  _dp, _dp, _dp, _dp]
        L.orc_center.argtypes = [i, _ip, _ip, _dp, _dp]
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None or len(a) == 0 else a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------- #
# topology and parameters
# --------------------------------------------------------------------------- #
def lj_table(eps, rmin_half, dtype=np.float64):
    """(c6, c12) per lower-triangular type pair, index a(a+1)/2 + b (a >= b).

    eps_ij = sqrt(eps_i eps_j), rmin_ij = rmin/2_i + rmin/2_j,
    c6 = 2 eps rmin^6, c12 = eps rmin^12 (src/CharmmParameters.cu:739-768).
    The reference rounds the table to float; pass dtype=np.float32 to do so.
    """
    eps = np.asarray(eps, np.float64)
    rh = np.asarray(rmin_half, np.float64)
    nt = len(eps)
    out = []
    for a in range(nt):
        for b in range(a + 1):
            e = np.sqrt(eps[a] * eps[b])
            rm = rh[a] + rh[b]
            out += [2.0 * e * rm ** 6, e * rm ** 12]
    return np.array(out, dtype=dtype).astype(np.float64)


def connectivity(n, bonds):
    """1-2, 1-3 and 1-4 neighbour sets (src/CharmmPSF.cu:331-361)."""
    c12 = [set() for _ in range(n)]
    c13 = [set() for _ in range(n)]
    c14 = [set() for _ in range(n)]
    for i, j in bonds:
        c12[i].add(int(j))
        c12[j].add(int(i))
    for i, j in bonds:
        for a in c12[j]:
            if a != i:
                c13[i].add(a)
        for b in c12[i]:
            if b != j:
                c13[j].add(b)
    for i in range(n):
        for j in c13[i]:
            for k in c12[j]:
                if k not in c12[i]:
                    c14[i].add(k)
    return c12, c13, c14


def exclusion_lists(n, bonds):
    """Returns (excl_ptr, excl_idx, in14 pairs, ex14 pairs).

    The direct-space exclusion set is 1-2 u 1-3 u 1-4 (src/CharmmPSF.cu:363-380);
    in14 = 1-4 pairs that are neither 1-2 nor 1-3, ex14 = 1-2 u 1-3
    (src/CharmmPSF.cu:235-279), each pair listed once with i < j.
    """
    c12, c13, c14 = connectivity(n, bonds)
    ptr = [0]
    idx = []
    in14, ex14 = [], []
    for i in range(n):
        s = (c12[i] | c13[i] | c14[i]) - {i}
        idx += sorted(s)
        ptr.append(len(idx))
        for j in sorted(c14[i]):
            if j > i and j not in c12[i] and j not in c13[i]:
                in14.append((i, j))
        for j in sorted(c12[i] | c13[i]):
            if j > i:
                ex14.append((i, j))
    return (np.array(ptr, np.int32), np.array(idx, np.int32),
            np.array(in14, np.int32).reshape(-1, 2), np.array(ex14, np.int32).reshape(-1, 2))


# --------------------------------------------------------------------------- #
# evaluation
# --------------------------------------------------------------------------- #
TERM_NAMES = ("bond", "angle", "ureyb", "dihe", "imdihe", "ewks", "ewse", "ewex", "elec", "vdw")


def evaluate(sysd, xyzq, terms=("bonded", "direct", "pairs14", "recip", "self"), nthreads=None):
    """FP64 energies/gradient/virial of one system.

    sysd keys: box (3,), kappa, roff (ctofnb), ron (ctonnb), e14fac, fft_grid (3,),
    order, vdwtype (n,), vdwtab, vdwtab14, excl_ptr, excl_idx, in14, ex14,
    bonds/bond_par, ureyb/ub_par, angles/angle_par (theta0 rad), dihe/dihe_par,
    imp/imp_par (any may be absent).  xyzq: (n, 4) float array (promoted to
    float64).  Returns dict(energy={name: value}, grad=(n,3), virial=(3,3)).
    """
    L = lib()
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    xyzq = np.ascontiguousarray(np.asarray(xyzq, np.float64).reshape(-1, 4))
    n = xyzq.shape[0]
    flat = xyzq.reshape(-1)
    grad = np.zeros(3 * n)
    vir = np.zeros(9)
    en = {k: 0.0 for k in TERM_NAMES}
    box = np.ascontiguousarray(sysd["box"], np.float64)
    vdwtype = np.ascontiguousarray(sysd["vdwtype"], np.int32)
    ntypes = int(vdwtype.max()) + 1 if n else 1
    kappa, roff, ron = float(sysd["kappa"]), float(sysd["roff"]), float(sysd["ron"])
    if "direct" in terms:
        e = np.zeros(2)
        rc = L.orc_direct(n, flat, vdwtype, ntypes, np.ascontiguousarray(sysd["vdwtab"], np.float64),
                          box, kappa, roff, ron, _ptr(sysd["excl_ptr"]), _ptr(sysd["excl_idx"]),
                          grad, e, vir, nthreads)
        if rc != 0:
            raise RuntimeError("orc_direct failed (%d)" % rc)
        en["elec"] += e[0]
        en["vdw"] += e[1]
    if "pairs14" in terms:
        e = np.zeros(3)
        in14 = np.ascontiguousarray(sysd["in14"], np.int32).reshape(-1)
        ex14 = np.ascontiguousarray(sysd["ex14"], np.int32).reshape(-1)
        L.orc_pairs14(len(in14) // 2, _ptr(in14), len(ex14) // 2, _ptr(ex14), flat, vdwtype, ntypes,
                      np.ascontiguousarray(sysd["vdwtab14"], np.float64), kappa, roff, ron,
                      float(sysd.get("e14fac", 1.0)), grad, e, vir)
        en["elec"] += e[0]
        en["vdw"] += e[1]
        en["ewex"] += e[2]
    if "recip" in terms:
        e = np.zeros(1)
        rc = L.orc_recip(n, flat, box, np.ascontiguousarray(sysd["fft_grid"], np.int32),
                         int(sysd.get("order", 4)), kappa, grad, e, vir, nthreads)
        if rc != 0:
            raise RuntimeError("orc_recip failed (%d)" % rc)
        en["ewks"] += e[0]
    if "self" in terms:
        en["ewse"] += L.orc_self_energy(n, flat, kappa)
    if "bonded" in terms:
        e = np.zeros(5)

        def arr(key, width, dt):
            a = sysd.get(key)
            if a is None:
                return 0, None
            a = np.ascontiguousarray(np.asarray(a, dt).reshape(-1, width))
            return a.shape[0], a

        nb, b = arr("bonds", 2, np.int32)
        _, bp = arr("bond_par", 2, np.float64)
        nub, u = arr("ureyb", 2, np.int32)
        _, up = arr("ub_par", 2, np.float64)
        na, a = arr("angles", 3, np.int32)
        _, ap = arr("angle_par", 2, np.float64)
        nd, dd = arr("dihe", 4, np.int32)
        _, dp = arr("dihe_par", 3, np.float64)
        ni, im = arr("imp", 4, np.int32)
        _, ip = arr("imp_par", 3, np.float64)
        keep = [b, bp, u, up, a, ap, dd, dp, im, ip]  # noqa: F841 (keep buffers alive)
        P = lambda x: None if x is None or x.size == 0 else x.ctypes.data_as(ctypes.c_void_p)
        L.orc_bonded(flat, nb, P(b), P(bp), nub, P(u), P(up), na, P(a), P(ap), nd, P(dd), P(dp),
                     ni, P(im), P(ip), grad, e, vir)
        for k, name in enumerate(("bond", "angle", "ureyb", "dihe", "imdihe")):
            en[name] += e[k]
    en["epot"] = sum(en[k] for k in TERM_NAMES)
    return {"energy": en, "grad": grad.reshape(n, 3), "virial": vir.reshape(3, 3)}


def fft3d(grid, sign):
    """Unnormalised complex 3-D DFT of a [z][y][x] array (oracle FFT)."""
    g = np.ascontiguousarray(grid, np.complex128)
    K = np.array([g.shape[2], g.shape[1], g.shape[0]], np.int32)
    buf = g.view(np.float64).reshape(-1).copy()
    rc = lib().orc_fft3d(K, buf, int(sign), 1)
    if rc != 0:
        raise RuntimeError("orc_fft3d failed")
    return buf.view(np.complex128).reshape(g.shape)


def bspline(order, w):
    th = np.zeros(order)
    dth = np.zeros(order)
    lib().orc_bspline(order, float(w), th, dth)
    return th, dth


def bsp_mod(order, nfft):
    out = np.zeros(nfft)
    lib().orc_bsp_mod(order, nfft, out)
    return out


def prepare(spec, table_dtype=np.float32):
    """Resolve a raw system spec into the arrays `evaluate` needs.

    spec keys: charges, vdwtype, eps, rmin_half (optional eps14, rmin_half14),
    bonds (the connectivity used for exclusions), bonded term lists, box,
    fft_grid, kappa, ctofnb, ctonnb, e14fac, order.  vdW tables are rounded
    like the reference's float tables (table_dtype) so that the FP64 oracle
    and the device path consume identical parameters.
    """
    n = len(spec["charges"])
    bonds = np.asarray(spec.get("bonds", np.zeros((0, 2))), np.int64).reshape(-1, 2)
    ptr, idx, in14, ex14 = exclusion_lists(n, bonds)
    eps14 = spec.get("eps14", spec["eps"])
    rh14 = spec.get("rmin_half14", spec["rmin_half"])
    d = dict(spec)
    # bonded coefficients are float in the reference (bondcoef float2,
    # anglecoef float2, dihecoef float4, src/CudaBondedForce.cu) and in the C
    # ABI: evaluate the same rounded values
    for key in ("bond_par", "ub_par", "angle_par", "dihe_par", "imp_par"):
        if key in d and d[key] is not None:
            d[key] = np.asarray(d[key], np.float64).astype(np.float32).astype(np.float64)
    # explicit (c6, c12) tables (NBFIX applied, float) take precedence over
    # the combination rule
    tab = (np.asarray(spec["vdwtab"], np.float64) if "vdwtab" in spec
           else lj_table(spec["eps"], spec["rmin_half"], table_dtype))
    tab14 = (np.asarray(spec["vdwtab14"], np.float64) if "vdwtab14" in spec
             else lj_table(eps14, rh14, table_dtype))
    d.update(
        vdwtab=tab,
        vdwtab14=tab14,
        excl_ptr=ptr, excl_idx=idx, in14=in14, ex14=ex14,
        roff=float(spec["ctofnb"]), ron=float(spec["ctonnb"]),
        order=int(spec.get("order", 4)), e14fac=float(spec.get("e14fac", 1.0)),
    )
    return d


# --------------------------------------------------------------------------- #
# Langevin thermostat leapfrog + SETTLE (oracle/md.c)
# --------------------------------------------------------------------------- #
TIMFAC = 0.0488882129        # AKMA time unit in ps (src/CudaIntegrator.cu:22)
KBOLTZ = 1.987191e-03        # kcal/mol/K (include/Constants.h:17)
TIP3P_GEOM = (15.99940, 1.00800, 0.9572, 1.5139)   # mO, mH, rOH, rHH


def philox4x32_10(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.ascontiguousarray(ctr, np.uint32),
                            np.ascontiguousarray(key, np.uint32), out)
    return out


def langevin_normals(seed, step, atom):
    out = np.zeros(3)
    lib().orc_langevin_normals(seed, step, atom, out)
    return out


def settle_geometry(geom=TIP3P_GEOM):
    out = np.zeros(3)
    lib().orc_settle_geometry(np.asarray(geom, np.float64), out)
    return out


def settle(waters, xref, x, geom=TIP3P_GEOM):
    """SETTLE x (n,3) against xref in place; waters (nw,3) int32."""
    w = np.ascontiguousarray(waters, np.int32).reshape(-1)
    lib().orc_settle(len(w) // 3, w, np.asarray(geom, np.float64),
                     np.ascontiguousarray(xref, np.float64).reshape(-1), x.reshape(-1))
    return x


def shake_water(waters, xref, x, geom=TIP3P_GEOM, tol=1e-15, maxit=10000):
    w = np.ascontiguousarray(waters, np.int32).reshape(-1)
    it = lib().orc_shake_water(len(w) // 3, w, np.asarray(geom, np.float64),
                               np.ascontiguousarray(xref, np.float64).reshape(-1),
                               x.reshape(-1), tol, maxit)
    if it < 0:
        raise RuntimeError("SHAKE did not converge")
    return x


class Langevin:
    """CPU restatement of CudaLangevinThermostatIntegrator on numpy state.

    x, v, dprev are (n, 3) float64 arrays updated in place; grad is dU/dx
    (n, 3).  dt_ps, friction_ps (1/ps), temperature (K) as in the reference
    setters; internally dt is in AKMA units and gamma = dt_ps * friction.
    shake: optional (groups (k, 4) int, r2 (k,)) SHAKE groups of bonds to
    hydrogen, constrained with the waters after every update.
    """

    def __init__(self, invmass, waters, dt_ps, temperature, friction_ps, seed, geom=TIP3P_GEOM,
                 shake=None):
        self.invmass = np.ascontiguousarray(invmass, np.float64)
        self.n = len(self.invmass)
        self.w = np.ascontiguousarray(waters, np.int32).reshape(-1)
        self.nw = len(self.w) // 3
        self.geom = np.asarray(geom, np.float64)
        self.dt = dt_ps / TIMFAC
        self.gamma = dt_ps * friction_ps
        self.kbt = KBOLTZ * temperature
        self.seed = int(seed)
        self.scratch = np.zeros(3 * self.n)
        if shake is None:
            shake = (np.zeros((0, 4), np.int32), np.zeros(0))
        self.sg = np.ascontiguousarray(shake[0], np.int32).reshape(-1)
        self.sr2 = np.ascontiguousarray(shake[1], np.float64).reshape(-1)
        if len(self.sg) == 0:  # ndpointer arguments need a real array
            self.sg, self.sr2 = np.full(4, -1, np.int32), np.ones(1)
            self.ns = 0
        else:
            self.ns = len(self.sr2)

    def init_constrain(self, x, v):
        lib().orc_lt_init_constrain(self.n, self.nw, self.w, self.geom, self.ns, self.sg,
                                    self.sr2, self.invmass, self.dt, x.reshape(-1), v.reshape(-1),
                                    self.scratch)

    def init_backstep(self, grad, x, v):
        dprev = np.zeros((self.n, 3))
        lib().orc_lt_init_backstep(self.n, self.invmass, self.nw, self.w, self.geom, self.ns,
                                   self.sg, self.sr2, self.dt, self.gamma,
                                   np.ascontiguousarray(grad, np.float64).reshape(-1),
                                   np.ascontiguousarray(x).reshape(-1),
                                   np.ascontiguousarray(v).reshape(-1), dprev.reshape(-1),
                                   self.scratch)
        return dprev

    def remove_com(self, dprev):
        lib().orc_lt_remove_com(self.n, self.invmass, dprev.reshape(-1))

    def step(self, grad, x, dprev, v, step):
        ke = np.zeros(2)
        lib().orc_lt_step(self.n, self.invmass, self.nw, self.w, self.geom, self.ns, self.sg,
                          self.sr2, self.dt, self.gamma, self.kbt, self.seed, int(step),
                          np.ascontiguousarray(grad, np.float64).reshape(-1), x.reshape(-1),
                          dprev.reshape(-1), v.reshape(-1), ke, self.scratch)
        return ke


def shake_groups(groups, r2, invmass, xref, x, tol=1e-13, maxit=1000):
    """Iterative SHAKE of hydrogen groups (orc_shake_groups) on (n, 3) x in
    place; returns the sweeps used (-1: not converged)."""
    g = np.ascontiguousarray(groups, np.int32).reshape(-1)
    return lib().orc_shake_groups(len(g) // 4, g, np.ascontiguousarray(r2, np.float64),
                                  np.ascontiguousarray(invmass, np.float64),
                                  np.ascontiguousarray(xref, np.float64).reshape(-1),
                                  x.reshape(-1), tol, maxit)


def center(groups_ptr, groups_atom, box, x):
    lib().orc_center(len(groups_ptr) - 1, np.ascontiguousarray(groups_ptr, np.int32),
                     np.ascontiguousarray(groups_atom, np.int32), np.asarray(box, np.float64),
                     x.reshape(-1))
    return x
