"""Seeded synthetic workloads: TIP3P water boxes and a solvated random-walk chain.

A system is a plain dict ("spec") of numpy arrays describing atoms, resolved
CHARMM parameters and PME settings -- the information ForceManager::initialize
draws from CharmmPSF + CharmmParameters (src/ForceManager.cu:436-550).  Keys:

  xyz (n,3) A, charges (n,), masses (n,), vdwtype (n,) int32,
  eps (ntypes,), rmin_half (ntypes,)   [CHARMM epsilon (<=0 allowed) and Rmin/2]
  bonds (nb,2), bond_par (nb,2) = (b0, kb)
  angles (na,3), angle_par (na,2) = (theta0 [rad], ktheta)
  ureyb (nu,2), ub_par (nu,2); dihe (nd,4), dihe_par (nd,3) = (n, k, delta[rad])
  imp (ni,4), imp_par (ni,3)
  box (3,), fft_grid (3,), kappa, cutnb, ctofnb, ctonnb, e14fac, order
  waters (nw,3) atom indices (O, H1, H2) of rigid waters

TIP3P geometry/parameters follow test/data/toppar_water_ions.str (OH 0.9572 A,
HOH 104.52 deg, HT-HT 1.5139 A; OT eps -0.1521 Rmin/2 1.7682, HT eps -0.046
Rmin/2 0.2245; q(O) -0.834, q(H) +0.417).  vdW types are numbered in sorted
type-name order as the reference does (HT = 0, OT = 1).
"""
import math

import numpy as np

WATER_DENSITY = 0.0334  # molecules / A^3
TIP3P = dict(roh=0.9572, theta=math.radians(104.52), rhh=1.5139, qo=-0.834, qh=0.417,
             mo=15.9994, mh=1.008, kb_oh=450.0, kb_hh=0.0, ka=55.0)
HT, OT = 0, 1
WATER_EPS = np.array([-0.046, -0.1521])
WATER_RMIN_HALF = np.array([0.2245, 1.7682])


def smooth_grid(L, spacing=1.0):
    """Smallest {2,3,5,7}-smooth integer >= L/spacing."""
    k = max(4, int(math.ceil(L / spacing - 1e-9)))
    while True:
        m = k
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        if m == 1:
            return k
        k += 1


def random_rotations(rng, n):
    """Uniform random rotation matrices (Shoemake quaternions)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], 1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_frame():
    h = TIP3P["theta"] / 2
    b = TIP3P["roh"]
    return np.array([[0.0, 0.0, 0.0], [b * math.sin(h), b * math.cos(h), 0.0],
                     [-b * math.sin(h), b * math.cos(h), 0.0]])


def _empty_lists(spec):
    for k, w in (("ureyb", 2), ("dihe", 4), ("imp", 4)):
        spec.setdefault(k, np.zeros((0, w), np.int32))
    spec.setdefault("ub_par", np.zeros((0, 2)))
    spec.setdefault("dihe_par", np.zeros((0, 3)))
    spec.setdefault("imp_par", np.zeros((0, 3)))
    return spec


def water_topology(n_waters, first=0):
    """Bonds O-H1, O-H2, H1-H2 and angle H1-O-H2 per water (waterbox.psf layout)."""
    o = first + 3 * np.arange(n_waters)
    bonds = np.stack([np.stack([o, o + 1], 1), np.stack([o, o + 2], 1), np.stack([o + 1, o + 2], 1)],
                     1).reshape(-1, 2)
    bond_par = np.tile(np.array([[TIP3P["roh"], TIP3P["kb_oh"]], [TIP3P["roh"], TIP3P["kb_oh"]],
                                 [TIP3P["rhh"], TIP3P["kb_hh"]]]), (n_waters, 1))
    angles = np.stack([o + 1, o, o + 2], 1)
    angle_par = np.tile(np.array([[TIP3P["theta"], TIP3P["ka"]]]), (n_waters, 1))
    return bonds.astype(np.int32), bond_par, angles.astype(np.int32), angle_par


def lattice_sites(m, L):
    a = L / m
    g = (np.arange(m) + 0.5) * a
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


def water_positions(sites, rng):
    R = random_rotations(rng, len(sites))
    fr = water_frame()
    return (sites[:, None, :] + np.einsum("nij,aj->nai", R, fr)).reshape(-1, 3)


def tip3p_box(n_waters, L, fft_grid=None, seed=0, kappa=0.34, cutnb=14.0, ctofnb=12.0,
              ctonnb=10.0):
    """n_waters rigid TIP3P molecules on a seeded cubic lattice in a cubic box
    of edge L with uniformly random orientations (O at the lattice sites)."""
    rng = np.random.default_rng(seed)
    m = int(math.ceil(round(n_waters ** (1.0 / 3.0), 9)))
    sites = lattice_sites(m, L)[:n_waters]
    xyz = water_positions(sites, rng)
    n = 3 * n_waters
    charges = np.tile([TIP3P["qo"], TIP3P["qh"], TIP3P["qh"]], n_waters)
    masses = np.tile([TIP3P["mo"], TIP3P["mh"], TIP3P["mh"]], n_waters)
    vdwtype = np.tile(np.array([OT, HT, HT], np.int32), n_waters)
    bonds, bond_par, angles, angle_par = water_topology(n_waters)
    K = fft_grid if fft_grid is not None else smooth_grid(L)
    if np.isscalar(K):
        K = (K, K, K)
    spec = dict(xyz=xyz, charges=charges, masses=masses, vdwtype=vdwtype, eps=WATER_EPS.copy(),
                rmin_half=WATER_RMIN_HALF.copy(), bonds=bonds, bond_par=bond_par, angles=angles,
                angle_par=angle_par, box=np.array([L, L, L], float),
                fft_grid=np.array(K, np.int32), kappa=kappa, cutnb=cutnb, ctofnb=ctofnb,
                ctonnb=ctonnb, e14fac=1.0, order=4,
                waters=(3 * np.arange(n_waters)[:, None] + np.arange(3)).astype(np.int32),
                name="tip3p_%d" % n_waters)
    assert len(xyz) == n
    return _empty_lists(spec)


def config1():
    """1,000 waters (3,000 atoms), box 31.04 A, PME 32^3 (seed 1001)."""
    return tip3p_box(1000, 31.04, 32, seed=1001)


def config2():
    """32,768 waters (98,304 atoms), box 99.36 A, PME 100^3 (seed 2001)."""
    return tip3p_box(32768, 99.36, 100, seed=2001)


def batched_sizes(n=64, seed=4001, lo=500, hi=8000):
    """Water counts log-uniform in [lo, hi] (seed 4001)."""
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), n))).astype(int)


def batched_system(i, n_waters):
    L = (n_waters / WATER_DENSITY) ** (1.0 / 3.0)
    return tip3p_box(int(n_waters), L, smooth_grid(L), seed=4100 + i)


def _chain(n, rng, b=1.53, theta=math.radians(111.0), clash=3.0, L=None):
    """Self-avoiding random walk: fixed bond length and angle, uniform
    torsions, rejection when a non-neighbour atom (more than three bonds away)
    comes within `clash` A.  With a box edge L the check uses minimum-image
    distances, so the chain does not overlap its own periodic images."""
    pos = np.zeros((n, 3))
    pos[1] = [b, 0, 0]
    pos[2] = pos[1] + b * np.array([-math.cos(theta), math.sin(theta), 0.0])
    if L is None:
        ncell, cell = None, clash
    else:
        ncell = max(3, int(L // clash))
        cell = L / ncell
    grid = {}

    def key(p):
        if ncell is None:
            return tuple(np.floor(p / cell).astype(int))
        return tuple(np.floor(np.mod(p, L) / cell).astype(int) % ncell)

    def near(kk):
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    c = (kk[0] + dx, kk[1] + dy, kk[2] + dz)
                    if ncell is not None:
                        c = (c[0] % ncell, c[1] % ncell, c[2] % ncell)
                    yield c

    for k in range(3):
        grid.setdefault(key(pos[k]), []).append(k)
    k = 3
    tries = 0
    while k < n:
        a, bb, c = pos[k - 3], pos[k - 2], pos[k - 1]
        phi = rng.uniform(-math.pi, math.pi)
        bc = c - bb
        bc /= np.linalg.norm(bc)
        nrm = np.cross(bb - a, bc)
        nrm /= np.linalg.norm(nrm)
        m = np.cross(nrm, bc)
        d = np.array([-math.cos(theta), math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi)])
        p = c + b * (d[0] * bc + d[1] * m + d[2] * nrm)
        kk = key(p)
        ok = True
        for cc in set(near(kk)):
            for j in grid.get(cc, ()):
                if j < k - 3:
                    dv = pos[j] - p
                    if L is not None:
                        dv -= L * np.round(dv / L)
                    if np.sum(dv * dv) < clash * clash:
                        ok = False
        tries += 1
        if not ok:
            if tries > 100:
                # dead end: back off a few atoms and regrow
                back = min(k - 3, 6)
                for j in range(k - back, k):
                    grid[key(pos[j])].remove(j)
                k -= back
                tries = 0
            continue
        tries = 0
        pos[k] = p
        grid.setdefault(kk, []).append(k)
        k += 1
    return pos


def solvated_chain(n_chain=2489, n_waters=7023, L=62.23, fft_grid=64, seeds=(3001, 3002, 3003, 3004)):
    """Random-walk solute (bonds 1.53 A k=222, angles 111 deg k=58, dihedrals
    n=3 k=0.2, q ~ N(0,0.3) shifted to zero, eps ~ U(0.02,0.2), Rmin/2 ~
    U(1,2)) in TIP3P water on lattice sites > 2.6 A from the chain."""
    rng = np.random.default_rng(seeds[0])
    chain = _chain(n_chain, rng, L=L)
    chain += 0.5 * L - chain.mean(0)
    rq = np.random.default_rng(seeds[1])
    q = rq.normal(0.0, 0.3, n_chain)
    q -= q.mean()
    rl = np.random.default_rng(seeds[2])
    ceps = -rl.uniform(0.02, 0.2, n_chain)
    crh = rl.uniform(1.0, 2.0, n_chain)
    # lattice sites whose oxygen is > 2.6 A from every chain atom (periodic);
    # the lattice is refined until n_waters such sites exist (a random walk of
    # this length excludes more volume than the bulk-density lattice leaves)
    from scipy.spatial import cKDTree
    tree = cKDTree(np.mod(chain, L), boxsize=L)
    m = int(round((L ** 3 * WATER_DENSITY) ** (1.0 / 3.0)))
    while True:
        sites = lattice_sites(m, L)
        dmin, _ = tree.query(sites)
        free = np.nonzero(dmin > 2.6)[0]
        if len(free) >= n_waters or m > 64:
            break
        m += 1
    rw = np.random.default_rng(seeds[3])
    waters = water_positions(sites[free[:n_waters]], rw).reshape(-1, 3, 3)
    nw = len(waters)
    n = n_chain + 3 * nw
    xyz = np.concatenate([chain, waters.reshape(-1, 3)])
    charges = np.concatenate([q, np.tile([TIP3P["qo"], TIP3P["qh"], TIP3P["qh"]], nw)])
    masses = np.concatenate([np.full(n_chain, 12.011), np.tile([TIP3P["mo"], TIP3P["mh"], TIP3P["mh"]], nw)])
    # types: water HT=0, OT=1, chain atom i -> 2+i
    eps = np.concatenate([WATER_EPS, ceps])
    rh = np.concatenate([WATER_RMIN_HALF, crh])
    vdwtype = np.concatenate([2 + np.arange(n_chain), np.tile([OT, HT, HT], nw)]).astype(np.int32)
    cb = np.stack([np.arange(n_chain - 1), np.arange(1, n_chain)], 1)
    ca = np.stack([np.arange(n_chain - 2), np.arange(1, n_chain - 1), np.arange(2, n_chain)], 1)
    cd = np.stack([np.arange(n_chain - 3), np.arange(1, n_chain - 2), np.arange(2, n_chain - 1),
                   np.arange(3, n_chain)], 1)
    wb, wbp, wa, wap = water_topology(nw, first=n_chain)
    bonds = np.concatenate([cb, wb]).astype(np.int32)
    bond_par = np.concatenate([np.tile([[1.53, 222.0]], (len(cb), 1)), wbp])
    angles = np.concatenate([ca, wa]).astype(np.int32)
    angle_par = np.concatenate([np.tile([[math.radians(111.0), 58.0]], (len(ca), 1)), wap])
    spec = dict(xyz=xyz, charges=charges, masses=masses, vdwtype=vdwtype, eps=eps, rmin_half=rh,
                bonds=bonds, bond_par=bond_par, angles=angles, angle_par=angle_par,
                dihe=cd.astype(np.int32), dihe_par=np.tile([[3.0, 0.2, 0.0]], (len(cd), 1)),
                box=np.array([L, L, L]), fft_grid=np.array([fft_grid] * 3, np.int32), kappa=0.34,
                cutnb=14.0, ctofnb=12.0, ctonnb=10.0, e14fac=1.0, order=4,
                waters=(n_chain + 3 * np.arange(nw)[:, None] + np.arange(3)).astype(np.int32),
                name="chain%d_w%d" % (n_chain, nw))
    assert len(xyz) == n
    return _empty_lists(spec)


def config3():
    """2,489-atom chain in 7,023 TIP3P waters, box 62.23 A, PME 64^3."""
    return solvated_chain()


def xyzq_of(spec, xyz=None):
    x = spec["xyz"] if xyz is None else xyz
    return np.concatenate([x, spec["charges"][:, None]], 1).astype(np.float32)


def lj_table(eps, rmin_half):
    """(c6, c12) per lower-triangular type pair with the CHARMM combination
    rule eps_ij = sqrt(eps_i eps_j), Rmin_ij = Rmin/2_i + Rmin/2_j,
    c6 = 2 eps Rmin^6, c12 = eps Rmin^12 (src/CharmmParameters.cu:739-768),
    rounded to float like the reference's table."""
    eps = np.asarray(eps, np.float64)
    rh = np.asarray(rmin_half, np.float64)
    a, b = np.tril_indices(len(eps))
    e = np.sqrt(eps[a] * eps[b])
    rm = rh[a] + rh[b]
    out = np.empty(2 * len(a))
    out[0::2] = 2.0 * e * rm ** 6
    out[1::2] = e * rm ** 12
    return out.astype(np.float32)
