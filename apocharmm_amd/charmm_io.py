"""CHARMM input files: PSF topology, CRD coordinates, parameter / stream files.

Readers with the reference's class and method names:
  CharmmPSF         src/CharmmPSF.cu:387+ (atom section, bond/angle/dihedral/
                    improper/cross-term lists), getWaterMolecules :281-291
  CharmmCrd         src/CharmmCrd.cu:19+ (title, count, one atom per line)
  CharmmParameters  src/CharmmParameters.cu:57-335 (BONDS, ANGLES with
                    Urey-Bradley, DIHEDRALS, IMPROPER, NONBONDED, NBFIX),
                    getBondedParamsAndLists :357-714, getVdwParamsAndTypes
                    :718-844
and `charmm_system`, which turns them into the system spec that ForceManager,
the integrator and the oracle consume.  Files may be gzip-compressed.

Key matching follows the reference: type names are upper-cased; the first
entry of a bond, angle, improper or NBFIX key wins, a repeated NONBONDED type
takes the last entry, and every line of a dihedral key adds a term.  Bond keys are
sorted, angle keys put the smaller outer type first, dihedral and improper keys
put the smaller outer type first (inner pair sorted when the outer ones are
equal); a dihedral without an exact entry falls back to X-b-c-X, an improper to
a-X-X-d.  A dihedral key with several lines is a multi-term (Fourier) series;
each term becomes its own dihedral entry here.  Angles whose Urey-Bradley force
constant is at most 0.01 get no Urey-Bradley term (:410-412).  vdW types are
the sorted set of the PSF's atom types (:718-735); NBFIX pairs replace the
combination rule in both the normal and the 1-4 tables (:739-800).
"""
import gzip
import math

import numpy as np

from .systems import lj_table

DEG = math.pi / 180.0


def _lines(path):
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt") as f:
        return f.read().splitlines()


class CharmmPSF:
    """Protein structure file (XPLOR-style, string atom types)."""

    def __init__(self, fileName):
        self.fileName = str(fileName)
        lines = _lines(fileName)
        sections = {}
        i = 0
        while i < len(lines):
            tok = lines[i].split()
            if len(tok) >= 2 and tok[1].startswith("!N"):
                key = tok[1].rstrip(":").split(":")[0]
                count = int(tok[0])
                sections[key] = (i + 1, count)
            i += 1
        if "!NATOM" not in sections:
            raise ValueError("no !NATOM section in PSF %s" % fileName)
        start, n = sections["!NATOM"]
        self.segids, self.resids, self.resnames, self.names, self.types = [], [], [], [], []
        charges, masses = [], []
        for k in range(n):
            f = lines[start + k].split()
            if len(f) < 8:
                raise ValueError("short atom line %d in PSF %s" % (k + 1, fileName))
            self.segids.append(f[1])
            self.resids.append(int(f[2]))
            self.resnames.append(f[3])
            self.names.append(f[4])
            self.types.append(f[5].upper())
            charges.append(float(f[6]))
            masses.append(float(f[7]))
        self.charges = np.array(charges)
        self.masses = np.array(masses)

        def ints(key, per):
            if key not in sections:
                return np.zeros((0, per), np.int32)
            s, cnt = sections[key]
            vals = []
            j = s
            while len(vals) < cnt * per:
                vals += [int(v) for v in lines[j].split()]
                j += 1
            return (np.array(vals[: cnt * per], np.int64).reshape(cnt, per) - 1).astype(np.int32)

        self.bonds = ints("!NBOND", 2)
        self.angles = ints("!NTHETA", 3)
        self.dihedrals = ints("!NPHI", 4)
        self.impropers = ints("!NIMPHI", 4)
        self.crossterms = ints("!NCRTERM", 8)

    # the reference's getters
    def getNumAtoms(self):
        return len(self.types)

    def getNumBonds(self):
        return len(self.bonds)

    def getNumAngles(self):
        return len(self.angles)

    def getNumDihedrals(self):
        return len(self.dihedrals)

    def getNumImpropers(self):
        return len(self.impropers)

    def getNumCrossTerms(self):
        return len(self.crossterms)

    def getSegmentIdentifiers(self):
        return self.segids

    def getResidueIdentifiers(self):
        return self.resids

    def getResidueNames(self):
        return self.resnames

    def getAtomNames(self):
        return self.names

    def getAtomTypes(self):
        return self.types

    def getCharges(self):
        return self.charges

    def getMasses(self):
        return self.masses

    def getBonds(self):
        return self.bonds

    def getAngles(self):
        return self.angles

    def getDihedrals(self):
        return self.dihedrals

    def getImpropers(self):
        return self.impropers

    def getCrossTerms(self):
        return self.crossterms

    def getWaterMolecules(self):
        """(n, 3) atom indices of consecutive (OT, HT, HT) triples."""
        t = self.types
        out = []
        i = 0
        while i < len(t) - 2:
            if t[i] == "OT" and t[i + 1] == "HT" and t[i + 2] == "HT":
                out.append((i, i + 1, i + 2))
                i += 3
            else:
                i += 1
        return np.array(out, np.int32).reshape(-1, 3)


class CharmmCrd:
    """CHARMM coordinate file (standard or EXT layout)."""

    def __init__(self, fileName):
        lines = [l for l in _lines(fileName) if not l.startswith("*") and l.strip()]
        n = int(lines[0].split()[0])
        xyz = np.empty((n, 3))
        for k in range(n):
            f = lines[1 + k].split()
            xyz[k] = (float(f[4]), float(f[5]), float(f[6]))
        self.coords = xyz

    def getCoordinates(self):
        return self.coords


def _dihedral_key(a, b, c, d):
    if a > d:
        a, b, c, d = d, c, b, a
    elif a == d and b > c:
        b, c = c, b
    return (a, b, c, d)


class CharmmParameters:
    """Force-field parameters from one or more parameter / stream files."""

    def __init__(self, fileNames):
        if isinstance(fileNames, (str, bytes)) or not hasattr(fileNames, "__iter__"):
            fileNames = [fileNames]
        self.bonds, self.angles, self.ureyb = {}, {}, {}
        self.dihedrals, self.impropers = {}, {}
        self.vdw, self.vdw14, self.nbfix = {}, {}, {}
        for name in fileNames:
            self.readCharmmParameterFile(name)

    def readCharmmParameterFile(self, fileName):
        lines = _lines(fileName)
        # a stream file carries topology first: parameters follow "READ PARA"
        for k, raw in enumerate(lines):
            ln = raw.split("!")[0].strip().upper()
            if ln.startswith("READ") and "PARA" in ln:
                lines = lines[k + 1:]
                break
        state = None
        k = 0
        while k < len(lines):
            ln = lines[k].split("!")[0].strip().upper()
            k += 1
            if not ln or ln.startswith("*"):
                continue
            head = ln.split()[0]
            new_state = None
            for prefix, st in (("ATOM", "atoms"), ("BOND", "bonds"), ("ANGL", "angles"),
                               ("THET", "angles"), ("DIHE", "dihedrals"), ("PHI", "dihedrals"),
                               ("IMPR", "impropers"), ("IMPH", "impropers"), ("CMAP", "cmap"),
                               ("NONB", "nonbonded"), ("NBON", "nonbonded"), ("NBFI", "nbfix"),
                               ("HBON", "hbond")):
                if head.startswith(prefix) and (len(ln.split()) == 1 or st == "nonbonded"
                                                or st == "hbond"):
                    new_state = st
                    break
            if head == "END" or head == "RETURN":
                state = None
                continue
            if new_state is not None:
                state = new_state
                if state == "nonbonded":
                    # the header's option list may continue over lines ending in "-"
                    while ln.endswith("-") and k < len(lines):
                        ln = lines[k].split("!")[0].strip().upper()
                        k += 1
                continue
            tok = ln.split()
            if state == "bonds" and len(tok) >= 4:
                a, b = sorted(tok[:2])
                self.bonds.setdefault((a, b), (float(tok[3]), float(tok[2])))  # (b0, kb)
            elif state == "angles" and len(tok) >= 5:
                a, b, c = tok[:3]
                if a > c:
                    a, c = c, a
                # (theta0, k) and (s0, kub)
                self.angles.setdefault((a, b, c), (float(tok[4]) * DEG, float(tok[3])))
                self.ureyb.setdefault((a, b, c), (float(tok[6]), float(tok[5])) if len(tok) >= 7
                                      else (0.0, 0.0))
            elif state == "dihedrals" and len(tok) >= 7:
                key = _dihedral_key(*tok[:4])
                # (n, k, delta)
                self.dihedrals.setdefault(key, []).append(
                    (float(int(tok[5])), float(tok[4]), float(tok[6]) * DEG))
            elif state == "impropers" and len(tok) >= 7:
                key = _dihedral_key(*tok[:4])
                self.impropers.setdefault(key, (0.0, float(tok[4]), float(tok[6]) * DEG))
            elif state == "nonbonded":
                if len(tok) == 4:
                    self.vdw[tok[0]] = (float(tok[2]), float(tok[3]))
                elif len(tok) == 7:
                    self.vdw[tok[0]] = (float(tok[2]), float(tok[3]))
                    self.vdw14[tok[0]] = (float(tok[5]), float(tok[6]))
                else:
                    raise ValueError("unexpected NONBONDED line in %s: %s" % (fileName, ln))
            elif state == "nbfix" and len(tok) >= 4:
                a, b = sorted(tok[:2])
                emin, rmin = abs(float(tok[2])), float(tok[3])
                self.nbfix.setdefault((a, b), (emin, rmin))
            elif state == "atoms" or state in ("cmap", "hbond"):
                continue

    # ---- lookups (src/CharmmParameters.cu:357-714) --------------------------------
    def _dihedral_terms(self, ta, tb, tc, td):
        key = _dihedral_key(ta, tb, tc, td)
        if key in self.dihedrals:
            return self.dihedrals[key]
        b, c = sorted((tb, tc))
        key = ("X", b, c, "X")
        if key in self.dihedrals:
            return self.dihedrals[key]
        raise KeyError("dihedral not found: %s %s %s %s" % (ta, tb, tc, td))

    def _improper(self, ta, tb, tc, td):
        key = _dihedral_key(ta, tb, tc, td)
        if key in self.impropers:
            return self.impropers[key]
        key = (key[0], "X", "X", key[3])
        if key in self.impropers:
            return self.impropers[key]
        raise KeyError("improper not found: %s %s %s %s" % (ta, tb, tc, td))

    def getBondedParamsAndLists(self, psf):
        """Bonded term lists with their parameters, in the spec layout:
        bonds/bond_par (b0, kb), ureyb/ub_par (s0, kub), angles/angle_par
        (theta0 [rad], k), dihe/dihe_par (n, k, delta [rad]) one row per
        Fourier term, imp/imp_par (0, k, psi0 [rad])."""
        t = psf.getAtomTypes()
        out = {}
        bp = []
        for i, j in psf.getBonds():
            key = tuple(sorted((t[i], t[j])))
            if key not in self.bonds:
                raise KeyError("bond not found: %s %s" % key)
            bp.append(self.bonds[key])
        out["bonds"] = psf.getBonds().astype(np.int32)
        out["bond_par"] = np.array(bp, np.float64).reshape(-1, 2)
        ap, ub, ubp = [], [], []
        for i, j, k in psf.getAngles():
            a, c = (t[i], t[k]) if t[i] <= t[k] else (t[k], t[i])
            key = (a, t[j], c)
            if key not in self.angles:
                raise KeyError("angle not found: %s %s %s" % key)
            ap.append(self.angles[key])
            s0, kub = self.ureyb.get(key, (0.0, 0.0))
            if abs(kub) > 0.01:
                ub.append((i, k))
                ubp.append((s0, kub))
        out["angles"] = psf.getAngles().astype(np.int32)
        out["angle_par"] = np.array(ap, np.float64).reshape(-1, 2)
        out["ureyb"] = np.array(ub, np.int32).reshape(-1, 2)
        out["ub_par"] = np.array(ubp, np.float64).reshape(-1, 2)
        dl, dp = [], []
        for q in psf.getDihedrals():
            for term in self._dihedral_terms(*(t[a] for a in q)):
                dl.append(q)
                dp.append(term)
        out["dihe"] = np.array(dl, np.int32).reshape(-1, 4)
        out["dihe_par"] = np.array(dp, np.float64).reshape(-1, 3)
        il, ip = [], []
        for q in psf.getImpropers():
            il.append(q)
            ip.append(self._improper(*(t[a] for a in q)))
        out["imp"] = np.array(il, np.int32).reshape(-1, 4)
        out["imp_par"] = np.array(ip, np.float64).reshape(-1, 3)
        return out

    def getVdwParamsAndTypes(self, psf):
        """Sorted vdW type names, per-atom type index, per-type (eps, rmin/2)
        for the normal and 1-4 tables, and the two (c6, c12) tables with the
        NBFIX pairs applied."""
        names = sorted(set(psf.getAtomTypes()))
        index = {n: i for i, n in enumerate(names)}
        missing = [n for n in names if n not in self.vdw]
        if missing:
            raise KeyError("no NONBONDED entry for types: %s" % " ".join(missing))
        eps = np.array([self.vdw[n][0] for n in names])
        rh = np.array([self.vdw[n][1] for n in names])
        eps14 = np.array([self.vdw14.get(n, self.vdw[n])[0] for n in names])
        rh14 = np.array([self.vdw14.get(n, self.vdw[n])[1] for n in names])
        tab = lj_table(eps, rh).reshape(-1, 2).astype(np.float64)
        tab14 = lj_table(eps14, rh14).reshape(-1, 2).astype(np.float64)
        for (a, b), (emin, rmin) in self.nbfix.items():
            if a in index and b in index:
                ia, ib = index[a], index[b]
                hi, lo = max(ia, ib), min(ia, ib)
                row = hi * (hi + 1) // 2 + lo
                for tb in (tab, tab14):
                    tb[row] = (2.0 * emin * rmin ** 6, emin * rmin ** 12)
        vdwtype = np.array([index[n] for n in psf.getAtomTypes()], np.int32)
        return dict(type_names=names, vdwtype=vdwtype, eps=eps, rmin_half=rh, eps14=eps14,
                    rmin_half14=rh14, vdwtab=tab.astype(np.float32).reshape(-1),
                    vdwtab14=tab14.astype(np.float32).reshape(-1))


def shake_groups(psf, prm):
    """SHAKE groups of bonds to hydrogen, as
    ForceManager::initializeHolonomicConstraintsVariables builds them
    (src/ForceManager.cu:844-918): every bond with an atom whose type starts
    with H, except water O-H (OT-HT) and H-H bonds, grouped by its heavy atom
    in bond order; a group holds the heavy atom and up to three hydrogens and
    one squared length, the b0 of its last hydrogen bond.  Returns
    ((k, 4) int32 with -1 padding, (k,) float64)."""
    t = psf.getAtomTypes()
    hyd = {}
    for i, j in psf.getBonds():
        i, j = int(i), int(j)
        hi, hj = t[i][0] == "H", t[j][0] == "H"
        if not (hi or hj):
            continue
        if {t[i], t[j]} == {"OT", "HT"} or (hi and hj):
            continue
        heavy, h = (j, i) if hi else (i, j)
        hyd.setdefault(heavy, []).append(h)
    groups, r2 = [], []
    for heavy in sorted(hyd):
        hs = hyd[heavy]
        if len(hs) > 3:
            raise ValueError("atom %d carries more than three hydrogens" % (heavy + 1))
        groups.append([heavy] + hs + [-1] * (3 - len(hs)))
        key = tuple(sorted((t[heavy], t[hs[-1]])))
        r2.append(prm.bonds[key][0] ** 2)
    return (np.array(groups, np.int32).reshape(-1, 4), np.array(r2, np.float64))


def charmm_system(psf, prm, xyz, box, fft_grid, kappa=0.34, cutnb=14.0, ctofnb=12.0,
                  ctonnb=10.0, e14fac=1.0, order=4, name=None):
    """System spec (the input of ForceManager, the integrator and the oracle)
    from a PSF, its parameters and coordinates.  xyz: (n, 3) array or a
    CharmmCrd; box: three edge lengths in A; fft_grid: int or three ints.
    Rigid waters are the PSF's (OT, HT, HT) triples and the other bonds to
    hydrogen form SHAKE groups (shake_groups, shake_r2); cross terms (CMAP)
    are not evaluated, as in the reference (src/CharmmParameters.cu:710)."""
    if isinstance(psf, str):
        psf = CharmmPSF(psf)
    if not isinstance(prm, CharmmParameters):
        prm = CharmmParameters(prm)
    if isinstance(xyz, CharmmCrd):
        xyz = xyz.getCoordinates()
    elif isinstance(xyz, str):
        xyz = CharmmCrd(xyz).getCoordinates()
    xyz = np.asarray(xyz, np.float64)
    if xyz.shape != (psf.getNumAtoms(), 3):
        raise ValueError("coordinates: %s for %d atoms" % (xyz.shape, psf.getNumAtoms()))
    K = np.array([fft_grid] * 3 if np.isscalar(fft_grid) else fft_grid, np.int32)
    spec = dict(xyz=xyz, charges=np.asarray(psf.getCharges(), np.float64),
                masses=np.asarray(psf.getMasses(), np.float64),
                box=np.array(box, np.float64), fft_grid=K, kappa=float(kappa),
                cutnb=float(cutnb), ctofnb=float(ctofnb), ctonnb=float(ctonnb),
                e14fac=float(e14fac), order=int(order), waters=psf.getWaterMolecules(),
                name=name or psf.fileName)
    spec.update(prm.getVdwParamsAndTypes(psf))
    spec.update(prm.getBondedParamsAndLists(psf))
    spec["shake_groups"], spec["shake_r2"] = shake_groups(psf, prm)
    return spec
