"""Writes the golden fixtures under tests/golden/ from the reference's own test data.

Run once in a container where the apoCHARMM source tree is mounted (it is read as
text only; nothing from it is imported or executed):

    python tests/golden/make_fixtures.py /root/reference

Inputs (all under <ref>/test/data): waterbox.psf/.crd, waterDimer.psf/.crd,
argon_10.psf/.crd, toppar_water_ions.str, argon.prm and the CHARMM c47b1 force
files waterbox.forces.c47b1.dat, waterDimer.forces.c47b1.dat.  The expected
energies, virials and argon forces are the literal values asserted in
test/unittests/unittest-eneForces.cpp (waterDimer :110-206, argon10 :208-286,
waterbox :288-365); they are restated below as data.

Each fixture is an .npz (no pickles) with the atoms, resolved force-field
parameters and expected outputs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_psf(path):
    """Minimal PSF reader: atom types/charges, bonds, angles, dihedrals, impropers."""
    lines = open(path).read().splitlines()
    out = {}
    i = 0
    while i < len(lines):
        ln = lines[i]
        if "!NATOM" in ln:
            n = int(ln.split()[0])
            types, charges, names = [], [], []
            for k in range(n):
                f = lines[i + 1 + k].split()
                names.append(f[4])
                types.append(f[5])
                charges.append(float(f[6]))
            out["types"] = types
            out["names"] = names
            out["charges"] = np.array(charges)
            i += n + 1
            continue
        for key, per in (("!NBOND", 2), ("!NTHETA", 3), ("!NPHI", 4), ("!NIMPHI", 4)):
            if key in ln:
                n = int(ln.split()[0])
                vals = []
                j = i + 1
                while len(vals) < n * per:
                    vals += [int(v) for v in lines[j].split()]
                    j += 1
                out[key] = np.array(vals, dtype=np.int64).reshape(n, per) - 1
                i = j
                break
        else:
            i += 1
    return out


def read_crd(path):
    lines = [l for l in open(path).read().splitlines() if not l.startswith("*")]
    n = int(lines[0].split()[0])
    xyz = np.array([[float(v) for v in l.split()[4:7]] for l in lines[1 : n + 1]])
    return xyz


def read_params(path):
    """BONDS / ANGLES / NONBONDED sections of a CHARMM parameter (or stream) file."""
    bonds, angles, nonb = {}, {}, {}
    section = None
    lines = open(path).read().splitlines()
    # a stream file carries topology first: parameters start after "read para"
    for k, raw in enumerate(lines):
        if raw.strip().lower().startswith("read para"):
            lines = lines[k + 1:]
            break
    for raw in lines:
        ln = raw.split("!")[0].strip()
        if not ln:
            continue
        head = ln.split()[0].upper()
        if head.startswith("BOND"):
            section = "bond"; continue
        if head.startswith("ANGL") or head.startswith("THET"):
            section = "angle"; continue
        if head.startswith("DIHE") or head.startswith("IMPR") or head.startswith("CMAP"):
            section = None; continue
        if head.startswith("NONB"):
            section = "nonb_hdr"; continue
        if head.startswith("NBFIX") or head.startswith("HBOND") or head == "END":
            section = None; continue
        if section == "nonb_hdr":
            if ln.endswith("-") or head.startswith("CUTNB"):
                continue
            section = "nonb"
        f = ln.split()
        if section == "bond" and len(f) >= 4:
            bonds[tuple(sorted(f[:2]))] = (float(f[2]), float(f[3]))
        elif section == "angle" and len(f) >= 5:
            angles[(f[0], f[1], f[2])] = (float(f[3]), float(f[4]))
            angles[(f[2], f[1], f[0])] = (float(f[3]), float(f[4]))
        elif section == "nonb" and len(f) >= 4:
            try:
                nonb[f[0]] = (float(f[2]), float(f[3]))
            except ValueError:
                pass
    return bonds, angles, nonb


def read_forces(path):
    rows = []
    for l in open(path).read().splitlines():
        l = l.strip()
        if not l or l.startswith("!"):
            continue
        rows.append([float(v) for v in l.split()[:3]])
    return np.array(rows)


def build(ref, psf_name, crd_name, prm_name, box, K, kappa, cutnb, ctofnb, ctonnb,
          energies, virial, forces, energy_rtol, energy_atol, force_atol, virial_atol):
    data = os.path.join(ref, "test", "data")
    psf = read_psf(os.path.join(data, psf_name))
    xyz = read_crd(os.path.join(data, crd_name))
    pb, pa, pn = read_params(os.path.join(data, prm_name))
    types = psf["types"]
    # the reference orders vdW types by sorted type name (std::set,
    # src/CharmmParameters.cu:723-735)
    uniq = sorted(set(types))
    tidx = np.array([uniq.index(t) for t in types], dtype=np.int32)
    eps = np.array([pn[t][0] for t in uniq])
    rmin2 = np.array([pn[t][1] for t in uniq])
    bonds = psf.get("!NBOND", np.zeros((0, 2), np.int64))
    angles = psf.get("!NTHETA", np.zeros((0, 3), np.int64))
    bond_par = np.array([pb[tuple(sorted((types[i], types[j])))][::-1] for i, j in bonds]).reshape(-1, 2)
    # bond_par rows: (b0, kb); angle rows: (theta0 [deg], ktheta)
    ang_par = np.array([pa[(types[i], types[j], types[k])][::-1] for i, j, k in angles]).reshape(-1, 2)
    names = sorted(energies)
    np.savez_compressed(
        os.path.join(HERE, psf_name.split(".")[0] + ".npz"),
        xyz=xyz, charges=psf["charges"], vdwtype=tidx, type_names=np.array(uniq),
        eps=eps, rmin_half=rmin2, bonds=bonds.astype(np.int32), bond_par=bond_par,
        angles=angles.astype(np.int32), angle_par_deg=ang_par,
        box=np.array(box, float), fft_grid=np.array(K, np.int32),
        kappa=kappa, cutnb=cutnb, ctofnb=ctofnb, ctonnb=ctonnb,
        ref_energy_names=np.array(names), ref_energy=np.array([energies[k] for k in names]),
        ref_virial=np.array(virial, float), ref_grad=forces,
        energy_rtol=energy_rtol, energy_atol=energy_atol, force_atol=force_atol,
        virial_atol=virial_atol,
    )


def main(ref):
    data = os.path.join(ref, "test", "data")
    # waterDimer: unittest-eneForces.cpp:110-206 (margins 1e-5 energies, 1e-4 virial/forces)
    build(ref, "waterDimer.psf", "waterDimer.crd", "toppar_water_ions.str",
          box=(520.0, 520.0, 520.0), K=(48, 48, 48), kappa=1e-8,
          cutnb=255.0, ctofnb=254.0, ctonnb=250.0,
          energies={"bond": 0.00003, "angle": 0.00002, "vdw": -0.00132, "elec": 0.12185,
                    "epot": 0.12057},
          virial=[0.4435916880, 0.0, 0.1160047166, -0.0, -0.007127478, 0.0,
                  0.1160047166, 0.0, 0.0512243359],
          forces=read_forces(os.path.join(data, "waterDimer.forces.c47b1.dat")),
          energy_rtol=0.0, energy_atol=1e-5, force_atol=1e-4, virial_atol=1e-4)
    # argon10: unittest-eneForces.cpp:208-286 (Approx energy, 1e-5 forces, 1e-6 virial)
    argon_forces = np.array([
        [0.00567, -0.00087, -0.01986], [-0.00791, 0.00989, 0.00778],
        [0.00513, -0.00134, 0.00932], [-0.00427, -0.00227, -0.00540],
        [-0.00334, 0.00684, 0.01087], [-0.00301, -0.01385, -0.02313],
        [0.01136, 0.00056, -0.00589], [-0.00778, 0.00468, 0.00800],
        [0.00337, 0.00260, 0.00934], [0.00078, -0.00625, 0.00898]])
    build(ref, "argon_10.psf", "argon_10.crd", "argon.prm",
          box=(50.0, 50.0, 50.0), K=(48, 48, 48), kappa=0.34,
          cutnb=9.0, ctofnb=8.0, ctonnb=7.0,
          energies={"epot": -0.17984},
          virial=[-0.0095837290, 0.0066020280, 0.0224834946, 0.0066020280, -0.0307467278,
                  -0.0684782285, 0.0224834946, -0.0684782285, -0.0510912119],
          forces=argon_forces, energy_rtol=1.2e-5, energy_atol=0.0, force_atol=1e-5,
          virial_atol=1e-6)
    # waterbox: unittest-eneForces.cpp:288-365 (Approx energies, 0.01 forces,
    # virial to 1e-4*|W00|)
    build(ref, "waterbox.psf", "waterbox.crd", "toppar_water_ions.str",
          box=(50.0, 50.0, 50.0), K=(48, 48, 48), kappa=0.34,
          cutnb=10.0, ctofnb=8.0, ctonnb=7.0,
          energies={"epot": -49755.02125, "bond": 1915.85339, "angle": 1071.29551,
                    "vdw": 8697.43632, "elec": -56211.95319, "ewks": 156.45211,
                    "ewse": -260256.60420, "ewex": 254872.49980},
          virial=[-7071.0775896888, -315.6040942237, -605.0510971046,
                  -315.6040942237, -6636.9115981748, -255.8901308852,
                  -605.0510971047, -255.8901308853, -6935.7815227056],
          forces=read_forces(os.path.join(data, "waterbox.forces.c47b1.dat")),
          energy_rtol=1.2e-5, energy_atol=0.0, force_atol=0.01,
          virial_atol=0.0001 * 7071.0775896888)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
