"""Writes the golden fixtures of tests/golden/ from the reference program.

Runs only where oracle/_ref/ref_harness exists (built from the reference
sources by oracle/Makefile).  Every vector here comes from the reference's
own code paths, driven through the harness:

  prep_*.json        prepared-target summaries (force type, ellipsoid, grid,
                     mass, charge, checksums of the cell-sorted arrays)
  forces_*.json      per-point site forces and energies for each kernel
  traj_*.json        per-trajectory (hit, success, chi) of the first
                     trajectories of each configuration
  published.json     the reference's published per-iteration results
                     (ccs.out, ccs_He.log, ccs_N2.log, ccs_Ar.log), parsed
                     from those files
  equi_*.json        equipotential-ellipsoid axes (src/Equipotential.cpp),
                     with the reference's OpenMP reductions on one thread

Usage: python tests/golden/make_golden.py [equi]
"""
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_py as O  # noqa: E402
from common import LONG, C60_XYZ, UBQ_PQR  # noqa: E402

REF_LOGS = "/root/reference"

CONFIGS = {
    "c60_He_t4": dict(targetFileName=C60_XYZ, GasBuffer="He", polarizability="yes", **LONG),
    "c60_He_t2": dict(targetFileName=C60_XYZ, GasBuffer="He", polarizability="no",
                      **{"Long-range forces": "no"}),
    "c60_N2_t6": dict(targetFileName=C60_XYZ, GasBuffer="N2", polarizability="yes", **LONG),
    "ubq_He_t4": dict(targetFileName=UBQ_PQR, GasBuffer="He", polarizability="yes", **LONG),
    "ubq_N2_t6": dict(targetFileName=UBQ_PQR, GasBuffer="N2", polarizability="yes", **LONG),
    "ubq_N2_t4": dict(targetFileName=UBQ_PQR, GasBuffer="N2", polarizability="no", **LONG),
    "ubq_CO2_t6": dict(targetFileName=UBQ_PQR, GasBuffer="CO2", polarizability="yes", **LONG),
    "ubq_Ar_t4": dict(targetFileName=UBQ_PQR, GasBuffer="Ar", polarizability="yes", **LONG),
}
for c in CONFIGS.values():
    c.update(seed=2104)

KERNELS = {
    "He": [("lennardjones_induced_dipole_LC", 0), ("lennardjones_LC", 0), ("lennardjones", 0)],
    "Ar": [("lennardjones_induced_dipole_LC", 0), ("lennardjones_LC", 0)],
    "N2": [("lennardjones_coulomb_LC", 0), ("coulomb_induced_dipole_iso_LC", 2),
           ("lennardjones_coulomb", 0), ("coulomb", 2), ("lennardjones_LC", 0),
           ("coulomb_LC", 2)],
    "CO2": [("lennardjones_coulomb_LC", 0), ("lennardjones_coulomb_induced_dipole_iso_LC_CO2", 1),
            ("lennardjones_LC_CO2", 1), ("lennardjones_coulomb_LC_CO2", 1),
            ("lennardjones_CO2", 1), ("lennardjones_coulomb_CO2", 1)],
}

N_TRAJ = {"c60_He_t4": 2000, "c60_He_t2": 2000, "c60_N2_t6": 300, "ubq_He_t4": 1000,
          "ubq_N2_t6": 100, "ubq_N2_t4": 100, "ubq_CO2_t6": 40, "ubq_Ar_t4": 300}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def rel(cfg):
    c = dict(cfg)
    c["targetFileName"] = os.path.basename(c["targetFileName"])
    return c


def points(prep, n, rng):
    a = np.array(prep["abc"])
    pts = rng.uniform(-0.6, 0.6, size=(n, 3)) * a
    k = rng.integers(0, prep["natoms"], size=n // 3)
    pts[: n // 3] = np.stack([prep["x"][k], prep["y"][k], prep["z"][k]], 1) + \
        rng.normal(0, 2.0, (n // 3, 3))
    hb = np.array(prep["halfbox"])
    pts[-3:] = hb * np.array([[1.01, 0.0, 0.1], [0.0, 0.2, -1.02], [-1.3, -1.3, -1.3]])
    return pts


def parse_log(path):
    txt = open(path).read()
    om = [float(x) for x in re.findall(r"omega: ([-0-9.eE+]+)", txt)]
    fr = [int(x) for x in re.findall(r"Nfree: (\d+)", txt)]
    sc = [int(x) for x in re.findall(r"Nscatter: (\d+)", txt)]
    lo = [int(x) for x in re.findall(r"Nlost: (\d+)", txt)]
    mean = float(re.search(r"average value of CCS = ([-0-9.eE+]+)", txt).group(1))
    err = float(re.search(r"error value of CCS = ([-0-9.eE+]+)", txt).group(1))
    return dict(omega=om, nfree=fr, nscatter=sc, nlost=lo, mean=mean, err=err)


EQUI = {
    "c60_He_t4": dict(CONFIGS["c60_He_t4"], Equipotential="yes"),
    "ubq_He_t4": dict(CONFIGS["ubq_He_t4"], Equipotential="yes"),
    "ubq_N2_t6": dict(CONFIGS["ubq_N2_t6"], Equipotential="yes"),
    "c60_Ar_t4": dict(targetFileName=C60_XYZ, GasBuffer="Ar", polarizability="yes", seed=2104,
                      Equipotential="yes", **LONG),
}


def equi():
    os.environ["OMP_NUM_THREADS"] = "1"  # the reference sums its potentials with OpenMP reductions
    for name, cfg in EQUI.items():
        prep = O.harness_prep(cfg)
        with open(os.path.join(HERE, "equi_%s.json" % name), "w") as fh:
            json.dump({"config": rel(cfg), "abc": prep["abc"], "bmax": prep["bmax"],
                       "grid": prep["grid"], "halfbox": prep["halfbox"]}, fh, indent=1)
        print(name, prep["abc"], flush=True)


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/ref_harness not built (make -C oracle ref)")
    rng = np.random.default_rng(20162104)
    for name, cfg in CONFIGS.items():
        prep = O.harness_prep(cfg)
        summary = {k: prep[k] for k in ("force_type", "natoms", "mass", "Q", "abc", "bmax", "mu",
                                        "alpha", "halfbox", "grid", "maxXYZ", "stencil_sums")}
        summary["sha256"] = {k: digest(prep[k]) for k in ("x", "y", "z", "q", "m", "eps", "sig")}
        summary["cells"] = prep["cells"]
        with open(os.path.join(HERE, "prep_%s.json" % name), "w") as fh:
            json.dump({"config": rel(cfg), "prep": summary}, fh, indent=1)
        gas = cfg["GasBuffer"]
        lines = []
        pts = points(prep, 30, rng)
        for kname, site in KERNELS[gas]:
            # kernels that read the second-shell stencil need "Long-range forces";
            # dipole kernels need a polarizability (alpha is unset otherwise)
            if cfg.get("Long-range forces") != "yes" and kname not in ("lennardjones", "lennardjones_LC"):
                continue
            if cfg.get("polarizability") != "yes" and "dipole" in kname:
                continue
            # coulomb_LC never takes its out-of-grid exit (src/Force.cpp:753)
            # and reads out of bounds there: only points inside the grid
            for r in (pts[:-3] if kname == "coulomb_LC" else pts):
                lines.append((kname, site, r))
        res = O.harness_force(cfg, lines)
        with open(os.path.join(HERE, "forces_%s.json" % name), "w") as fh:
            json.dump({"config": rel(cfg),
                       "rows": [[k, s, list(map(float, r)), list(map(float, f)), float(U)]
                                for (k, s, r), (_, _, f, U) in zip(lines, res)]}, fh)
        hit, ok, chi = O.harness_traj(cfg, 0, N_TRAJ[name])
        with open(os.path.join(HERE, "traj_%s.json" % name), "w") as fh:
            json.dump({"config": rel(cfg), "hit": hit.tolist(), "success": ok.tolist(),
                       "chi": [float(c) for c in chi]}, fh)
        print(name, "done", flush=True)
    pub = {}
    for key, fn in [("ubq_He", "ccs.out"), ("c60_He", "ccs_He.log"), ("c60_N2", "ccs_N2.log"),
                    ("c60_Ar", "ccs_Ar.log"), ("c60_CO2", "ccs_CO2.log")]:
        pub[key] = parse_log(os.path.join(REF_LOGS, fn))
    with open(os.path.join(HERE, "published.json"), "w") as fh:
        json.dump(pub, fh, indent=1)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "equi":
        equi()
    else:
        main()
