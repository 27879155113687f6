"""Shared helpers for the test-suite: workloads, oracle construction from a
prepared target, and the parity criteria."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
if REPO not in sys.path:
    sys.path.insert(0, REPO)
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import massccs_amd as M  # noqa: E402
from massccs_amd import synth  # noqa: E402
import oracle_py as O  # noqa: E402

GOLDEN = os.path.join(HERE, "golden")
C60_XYZ = os.path.join(GOLDEN, "c60.xyz")
UBQ_PQR = os.path.join(GOLDEN, "04_1ubq_charge_4e_min.pqr")

LONG = {"Short-range cutoff": "yes", "Long-range forces": "yes", "Long-range cutoff": "yes"}

# Configurations of the reference's published runs (log headers of ccs.out,
# ccs_He.log, ccs_N2.log, ccs_Ar.log): seed 2104, 10 x 10^4 trajectories.
PUBLISHED = {
    "ubq_He": dict(targetFileName=UBQ_PQR, GasBuffer="He", polarizability="yes", **LONG),
    "c60_He": dict(targetFileName=C60_XYZ, GasBuffer="He", polarizability="yes", **LONG),
    "c60_N2": dict(targetFileName=C60_XYZ, GasBuffer="N2", polarizability="yes", **LONG),
    "c60_Ar": dict(targetFileName=C60_XYZ, GasBuffer="Ar", polarizability="yes", **LONG),
}
for _c in PUBLISHED.values():
    _c.update(seed=2104, dt=10.0, Temp=298.0, skin=0.01, numberProbe=10000, nIter=10)


def globule_file(tmpdir):
    path = os.path.join(str(tmpdir), "globule_1231.xyz")
    if not os.path.exists(path):
        s, xyz, q = synth.globule()
        synth.write_xyz(path, s, xyz, q, comment="synthetic globule 1231 atoms, net +4e")
    return path


def c60_synth_file(tmpdir):
    path = os.path.join(str(tmpdir), "c60_analytic.xyz")
    if not os.path.exists(path):
        synth.write_xyz(path, ["C"] * 60, synth.c60(), comment="C60 truncated icosahedron")
    return path


def c60_ff_file(tmpdir):
    """LB-mixed C/He pair parameters (eps = sqrt(0.0860*0.0200), sigma = (3.40+2.64)/2)."""
    path = os.path.join(str(tmpdir), "c60_he.ff")
    if not os.path.exists(path):
        with open(path, "w") as fh:
            fh.write("1\nSymbol mass eps sigma\nC 12.011 0.041473 3.02\n")
    return path


def globule_cfg(tmpdir, n_probe=10000, n_iter=10):
    """The 1,231-atom globule with the N2 probe, LJ + Coulomb + induced dipole
    with cutoffs (force type 6)."""
    return dict(targetFileName=globule_file(tmpdir), GasBuffer="N2", polarizability="yes",
                alpha=1.7403, seed=2104, dt=10.0, Temp=298.0, numberProbe=n_probe, nIter=n_iter,
                **LONG)


def c60_lj_cfg(tmpdir, n_probe=10000, n_iter=3, cutoff=True):
    """C60 / He, Lennard-Jones only (force type 2, or 1 without cutoff)."""
    return {"targetFileName": c60_synth_file(tmpdir), "force-field": c60_ff_file(tmpdir),
            "GasBuffer": "He", "Short-range cutoff": "yes" if cutoff else "no",
            "Long-range forces": "no", "polarizability": "no", "seed": 2104, "dt": 10.0,
            "Temp": 298.0, "numberProbe": n_probe, "nIter": n_iter}


def prepared(cfg):
    return M.Prepared(M.config_from_dict(cfg))


def oracle_for(P, cfg):
    """Oracle context over the product's prepared target (cell-sorted arrays,
    grid, ellipsoid); the preparation itself is pinned separately."""
    i = P.info
    a = P.atoms()
    prep = dict(natoms=i.natoms, force_type=i.force_type, grid=[i.nx, i.ny, i.nz, i.ncells],
                alpha=i.alpha, mu=i.mu, abc=[i.a, i.b, i.c], bmax=i.bmax,
                halfbox=list(i.half_box), **a)
    return O.Oracle(prep, cfg)


def pair_scales(P, cfg, r, qi, lj, coul, dip, cut=True):
    """Scale of the parity criterion: sums of |pair term| (force per component
    and energy), plus alpha |E| |grad E| for the induced dipole."""
    i = P.info
    a = P.atoms()
    d = O.input_defaults(cfg)
    dx = r[0] - a["x"]
    dy = r[1] - a["y"]
    dz = r[2] - a["z"]
    r2 = dx * dx + dy * dy + dz * dz
    rr = np.sqrt(r2)
    rlj, rc = d["lj_cutoff"], d["coul_cutoff"]
    mlj = (rr < rlj) if cut else np.ones_like(rr, bool)
    mc = (rr < rc) if cut else np.ones_like(rr, bool)
    fs = np.zeros(3)
    us = 0.0
    if lj:
        eps, sig = (a["eps_c"], a["sig_c"]) if lj == "central" else (a["eps"], a["sig"])
        l1 = 4 * eps * sig ** 6
        l2 = l1 * sig ** 6
        r6 = 1 / r2 ** 3
        flj = np.abs(r6 * (12 * l2 * r6 - 6 * l1) / r2)[mlj]
        fs += np.array([np.sum(flj * np.abs(v[mlj])) for v in (dx, dy, dz)])
        us += np.sum(np.abs(r6 * (l2 * r6 - l1))[mlj]) + np.sum(np.abs(l1[mlj]) / rlj ** 6)
    if coul:
        uc = np.abs(qi * a["q"] * 332.06348078020324 / rr)[mc]
        fs += np.array([np.sum(uc / rr[mc] ** 2 * np.abs(v[mc])) for v in (dx, dy, dz)])
        us += 2.5 * np.sum(uc) + np.sum(np.abs(qi * a["q"][mc]) * 332.1 * 2 / rc)
    if dip:
        q = np.abs(a["q"])[mc]
        E = np.sum(q / rr[mc] ** 2)
        G = np.sum(3 * q / rr[mc] ** 3) + np.sum(q) / rc ** 3
        fs += i.alpha * E * G * 3
        us += 0.5 * i.alpha * E * E * 3
    return fs + 1e-300, us + 1e-300


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
