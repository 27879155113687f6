"""Shared helpers for the test-suite: workloads, oracle construction from a
prepared target, and the parity criteria."""
import os
import sys
import time

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


def oracle_for(P, cfg, max_total_steps=0):
    """Oracle context over the product's prepared target (cell-sorted arrays,
    grid, ellipsoid); the preparation itself is pinned separately."""
    i = P.info
    a = P.atoms()
    prep = dict(natoms=i.natoms, force_type=i.force_type, grid=[i.nx, i.ny, i.nz, i.ncells],
                alpha=i.alpha, mu=i.mu, abc=[i.a, i.b, i.c], bmax=i.bmax,
                halfbox=list(i.half_box), **a)
    return O.Oracle(prep, cfg, max_total_steps=max_total_steps)


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


# ---------------------------------------------------------------- trajectory parity
# A trajectory is chaotic when its fate or its scattering angle depends on
# the integration itself.  The oracle runs it twice more: at dt/2 (every
# try's step halved), and at dt from an impact parameter moved by four units
# in the last place (4 x 2^-52 relative).  It counts as non-chaotic when the
# tries and the fate are the same in all three runs, the 4-ulp run takes
# exactly as many steps as the dt run (an orbit around a charged target can
# keep its tries and fate by chance while its length changes by thousands of
# steps: every try of it is chaotic), and, if it scattered,
# |chi(dt) - chi(dt/2)| <= CHAOS_DCHI (the step's truncation error stays
# small) and |chi - chi(b + 4 ulp)| <= CHAOS_ULP_DCHI (a perturbation of the
# size of a reordered FP64 sum is not amplified).  On a non-chaotic
# trajectory the rounding of the engine's other summation order cannot move
# chi by 1e-8 rad; on a chaotic one it moves it by up to O(1).  The ulp run
# catches the trajectories whose dt/2 change happens to be small although a
# dt/4 run changes their fate (seen on the 3-site CO2 probe, whose energy
# test allows 20 %).
CHAOS_DCHI = 0.1
CHAOS_ULP_DCHI = 1e-10
CHI_TOL = 1e-8        # rad, non-chaotic trajectories with chi > CHI_SMALL
CHI_SMALL = 1e-6      # below: compared as 1 - cos chi
COS_TOL = 1e-15

_REF_CACHE = {}


def oracle_refs(P, cfg, u, key=None):
    """Oracle records of samples u at the run's dt, at dt/2, and at dt with
    the impact parameter moved by 4 ulp (cached by key within the test
    session: several tests compare launch variants against the same
    trajectories).  The two classification runs stop a trajectory once its
    steps exceed 2.5 (dt/2) or 1.5 (4 ulp) times the longest trajectory of
    the dt run, plus 10^4: it is then reported lost (tries = -1), differs
    from the dt run, and counts as chaotic, which it is, having changed its
    length that much; the budget keeps an orbit that a chaotic trajectory
    enters at dt/2 from running for minutes."""
    if key is not None and key in _REF_CACHE:
        return _REF_CACHE[key]
    t0 = time.perf_counter()
    ref = oracle_for(P, cfg).run(u)
    print("oracle: %d trajectories at dt in %.1f s" % (len(u), time.perf_counter() - t0), flush=True)
    longest = int(ref["steps"].max()) if len(ref) else 0
    half = oracle_for(P, dict(cfg, dt=float(cfg.get("dt", 10.0)) / 2),
                      max_total_steps=int(2.5 * longest) + 10000).run(u)
    print("oracle: at dt/2, %.1f s" % (time.perf_counter() - t0), flush=True)
    up = np.array(u, dtype=np.float64, copy=True)
    up[:, 0] *= 1.0 + 4.0 * 2.0 ** -52
    pert = oracle_for(P, cfg, max_total_steps=int(1.5 * longest) + 10000).run(up)
    print("oracle: from b + 4 ulp, %.1f s" % (time.perf_counter() - t0), flush=True)
    if key is not None:
        _REF_CACHE[key] = (ref, half, pert)
    return ref, half, pert


def non_chaotic(ref, half, pert):
    same = ((ref["tries"] == half["tries"]) & (ref["success"] == half["success"]) &
            (ref["tries"] == pert["tries"]) & (ref["success"] == pert["success"]) &
            (ref["steps"] == pert["steps"]))
    # near-zero angles (free flights through the field, acos of 1 - ulp) are
    # compared as 1 - cos chi, as in the criterion itself
    big = np.maximum(ref["chi"], pert["chi"]) > CHI_SMALL
    ulp_ok = np.where(big, np.abs(ref["chi"] - pert["chi"]) <= CHAOS_ULP_DCHI,
                      np.abs(np.cos(ref["chi"]) - np.cos(pert["chi"])) <= COS_TOL)
    stable_chi = (ref["success"] == 0) | ((np.abs(ref["chi"] - half["chi"]) <= CHAOS_DCHI) & ulp_ok)
    return (ref["hit"] == 1) & same & stable_chi


def assert_traj_parity(rec, refs, name, min_big=20):
    """The trajectory criteria against the oracle on the same samples: hit
    flags identical everywhere; on every non-chaotic hit the same tries and
    fate; on every non-chaotic scattered trajectory with chi > 1e-6 rad,
    |dchi| <= 1e-8 rad, and below that |d(1 - cos chi)| <= 1e-15.  No
    fraction may fail.  Returns the largest |dchi| of the checked set.
    refs = oracle_refs(...)."""
    ref, half, pert = refs
    assert np.array_equal(rec["hit"], ref["hit"]), name
    nc = non_chaotic(ref, half, pert)
    fate = (rec["success"] == ref["success"]) & (rec["tries"] == ref["tries"])
    bad_fate = np.where(nc & ~fate)[0]
    assert bad_fate.size == 0, (name, "fate differs on non-chaotic trajectories", bad_fate[:10],
                                rec["tries"][bad_fate[:10]], ref["tries"][bad_fate[:10]])
    sc = nc & (ref["success"] == 1)
    big = sc & (ref["chi"] > CHI_SMALL)
    small = sc & (ref["chi"] <= CHI_SMALL)
    dchi = np.abs(rec["chi"] - ref["chi"])
    worst = float(dchi[big].max()) if big.any() else 0.0
    bad = np.where(big & (dchi > CHI_TOL))[0]
    assert bad.size == 0, (name, "chi differs", bad[:10], rec["chi"][bad[:10]], ref["chi"][bad[:10]])
    dcos = np.abs(np.cos(rec["chi"]) - np.cos(ref["chi"]))
    bad = np.where(small & (dcos > COS_TOL))[0]
    assert bad.size == 0, (name, "small chi differs", bad[:10])
    print("%s: hits %d, non-chaotic %d, chi>1e-6 checked %d (max |dchi| %.3g), small chi %d, "
          "chaotic %d" % (name, int(ref["hit"].sum()), int(nc.sum()), int(big.sum()), worst,
                          int(small.sum()), int(((ref["hit"] == 1) & ~nc).sum())), flush=True)
    assert big.sum() >= min_big, (name, "too few non-chaotic scattered trajectories", big.sum())
    return worst


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
