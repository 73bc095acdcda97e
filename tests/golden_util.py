"""Loads the committed c47b1 golden fixtures into raw system specs."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("waterDimer", "argon_10", "waterbox")


def load(name):
    f = np.load(os.path.join(GOLDEN, name + ".npz"))
    ang = f["angle_par_deg"]
    spec = dict(
        charges=f["charges"], vdwtype=f["vdwtype"].astype(np.int32), eps=f["eps"],
        rmin_half=f["rmin_half"], bonds=f["bonds"], bond_par=f["bond_par"],
        angles=f["angles"],
        angle_par=np.stack([np.deg2rad(ang[:, 0]), ang[:, 1]], 1) if len(ang) else np.zeros((0, 2)),
        box=f["box"], fft_grid=f["fft_grid"], kappa=float(f["kappa"]),
        cutnb=float(f["cutnb"]), ctofnb=float(f["ctofnb"]), ctonnb=float(f["ctonnb"]),
        e14fac=1.0, order=4,
    )
    xyzq = np.concatenate([f["xyz"], f["charges"][:, None]], 1)
    ref = dict(
        energy=dict(zip([str(s) for s in f["ref_energy_names"]], f["ref_energy"])),
        virial=f["ref_virial"], grad=f["ref_grad"],
        energy_rtol=float(f["energy_rtol"]), energy_atol=float(f["energy_atol"]),
        force_atol=float(f["force_atol"]), virial_atol=float(f["virial_atol"]),
    )
    return spec, xyzq, ref
