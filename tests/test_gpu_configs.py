"""The large benchmark configurations on the GPU against the oracle: the
50,000-atom ellipsoid (N2, force type 6), the 1,000,000-atom hollow shell
(N2, force type 6) and the 64-assembly batch (He type 4 / N2 type 6):
per-point site forces at 1e-12 of the pair-term scale, trajectories under
the strict criterion (common.assert_traj_parity), and for the shell a
statistical check of a larger run (32 against 64 lanes per trajectory,
which changes only the summation order)."""
import numpy as np
import pytest

from common import (M, assert_traj_parity, gpu_available, oracle_for, oracle_refs, pair_scales,
                    synth)

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

N2_CFG = {"targetFileName": "memory.xyz", "GasBuffer": "N2", "Short-range cutoff": "yes",
          "LJ-cutoff": 12.0, "Long-range forces": "yes", "Long-range cutoff": "yes",
          "Coul-cutoff": 25.0, "polarizability": "yes", "alpha": 1.7403, "seed": 2104,
          "dt": 10.0, "Temp": 298.0}

_TARGETS = {}


def _target(which):
    if which not in _TARGETS:
        if which == "ellipsoid50k":
            s, xyz, q = synth.globule(n=50000, axes=(60.0, 45.0, 35.0), net_charge=20.0, seed=2104)
        else:
            s, xyz, q = synth.shell()
        P = M.Prepared(M.config_from_dict(N2_CFG), symbols=list(s), xyz=xyz, q=q)
        _TARGETS[which] = P
    return _TARGETS[which]


def _surface_points(P, n, rng):
    """Points a probe reaches: around random atoms, 2 A or more from all."""
    a = P.atoms()
    k = rng.integers(0, P.info.natoms, 4 * n)
    pts = np.stack([a["x"][k], a["y"][k], a["z"][k]], 1) + rng.normal(0, 5.0, (4 * n, 3))
    return pts[:n]


@pytest.mark.parametrize("which", ["ellipsoid50k", "shell1m"])
def test_large_target_site_forces_match_oracle(which):
    print("case site forces", which, flush=True)
    P = _target(which)
    orc = oracle_for(P, N2_CFG)
    E = M.Engine(P)
    rng = np.random.default_rng(3)
    pts = _surface_points(P, 60, rng)
    for site, kernel, qi, terms in [(M.SITE_N, "lennardjones_coulomb_LC", -0.4825, (True, True, False)),
                                    (M.SITE_DUMMY, "coulomb_induced_dipole_iso_LC", 0.965,
                                     (False, True, True))]:
        f, U = E.eval_sites(site, pts)
        worst = 0.0
        for k in range(len(pts)):
            fr, ur = orc.site_force(kernel, pts[k], qi)
            sf, su = pair_scales(P, N2_CFG, pts[k], qi, *terms, cut=True)
            err = max(np.max(np.abs(f[k] - fr) / sf), abs(U[k] - ur) / su)
            worst = max(worst, err)
            assert err <= 1e-12, (which, kernel, k, pts[k], f[k], fr, U[k], ur)
        print(which, kernel, "worst", worst, flush=True)


@pytest.mark.parametrize("which,n,min_big", [("ellipsoid50k", 240, 5), ("shell1m", 40, 2)])
def test_large_target_trajectories_match_oracle(which, n, min_big):
    """Hit flags everywhere, and every non-chaotic trajectory's tries, fate
    and angle (the oracle at dt, dt/2 and from a 4-ulp impact parameter).
    The shell's +60 e makes most of its trajectories chaotic orbits, and each
    costs the oracle seconds: 40 of them, of which 11 are non-chaotic and 2
    scatter by more than 1e-6 rad."""
    print("case trajectories", which, n, flush=True)
    P = _target(which)
    u = P.samples(n)
    refs = oracle_refs(P, N2_CFG, u)
    rec = M.Engine(P).run(u)
    assert_traj_parity(rec, refs, which, min_big=min_big)


def test_ellipsoid_ccs_and_losses_agree_with_oracle():
    """A statistical check of the 50k ellipsoid (N2, force type 6) against the
    oracle on the same 2,000 samples: the CCS within three paired standard
    errors (chaotic trajectories part ways on any rounding change, so the
    records of the two cannot agree one by one), and the number of
    trajectories lost to the reference's energy test within four Poisson
    standard deviations (both lose about 8 %)."""
    n = 2000
    print("case ccs ellipsoid50k vs oracle, %d trajectories" % n, flush=True)
    P = _target("ellipsoid50k")
    u = P.samples(n)
    rec = M.Engine(P).run(u)
    ref = oracle_for(P, N2_CFG).run(u)
    assert np.array_equal(rec["hit"], ref["hit"])
    den = float(((ref["hit"] == 0) | (ref["success"] == 1)).sum())
    dens = float(((rec["hit"] == 0) | (rec["success"] == 1)).sum())
    co, cg = ref["domega"].sum() / den, rec["domega"].sum() / dens
    se = float(np.std(rec["domega"] - ref["domega"]) * np.sqrt(n) / den)
    lost_o = int(((ref["hit"] == 1) & (ref["success"] == 0)).sum())
    lost_g = int(((rec["hit"] == 1) & (rec["success"] == 0)).sum())
    print("ccs gpu %.2f oracle %.2f paired SE %.2f; lost gpu %d oracle %d"
          % (cg, co, se, lost_g, lost_o), flush=True)
    assert abs(cg - co) <= 3 * se + 1e-9 * co
    assert abs(lost_g - lost_o) <= 4 * np.sqrt(max(lost_o, 1)) + 1


def test_shell_ccs_independent_of_summation_order():
    """2 x 10^4 trajectories of the 10^6-atom shell with 64 and with 32 lanes
    per trajectory (the same force field, summed in another order): hit flags
    identical, CCS within three paired standard errors (the shell's net
    +60 e makes a large share of the trajectories chaotic)."""
    print("case shell 32 vs 64 lanes", flush=True)
    P = _target("shell1m")
    u = P.samples(20000)
    a = M.Engine(P, lanes=64).run(u)
    b = M.Engine(P, lanes=32).run(u)
    assert np.array_equal(a["hit"], b["hit"])
    den = float(((a["hit"] == 0) | (a["success"] == 1)).sum())
    ca = a["domega"].sum() / den
    cb = b["domega"].sum() / float(((b["hit"] == 0) | (b["success"] == 1)).sum())
    se = float(np.std(a["domega"] - b["domega"]) * np.sqrt(len(a)) / den)
    print("shell ccs 64 lanes %.2f, 32 lanes %.2f, paired SE %.2f" % (ca, cb, se), flush=True)
    assert abs(ca - cb) <= 3 * se + 1e-9 * ca


def test_assembly_batch_against_engines_and_oracle():
    """Sixteen of the 64 assemblies (numbers 0, 1, 8, 9, ..., 57: 500-200,000
    atoms; even numbers carry He, force type 4, odd ones N2, force type 6) in
    one batch: records byte-identical to per-target runs of the same launch
    shape, and the smallest He and N2 members against the oracle under the
    strict trajectory criterion."""
    print("case assembly batch", flush=True)
    descs = [d for k, d in enumerate(synth.assemblies()) if k % 8 in (0, 1)]
    P, cfgs, us = [], [], []
    for d in descs:
        s, xyz, q = synth.globule(n=d["n"], axes=d["axes"], net_charge=d["net_charge"], seed=d["seed"])
        cfg = {"targetFileName": "memory.xyz", "GasBuffer": d["probe"], "polarizability": "yes",
               "Short-range cutoff": "yes", "Long-range forces": "yes",
               "Long-range cutoff": "yes", "seed": d["seed"]}
        if d["probe"] == "N2":
            cfg["alpha"] = 1.7403
        p = M.Prepared(M.config_from_dict(cfg), symbols=list(s), xyz=xyz, q=q)
        P.append(p)
        cfgs.append(cfg)
        us.append(p.samples(300))
    print("atoms", [p.info.natoms for p in P], flush=True)
    B = M.Batch(P, lanes=64, skin=1.0)
    recs = B.run(us)
    for p, u, r in zip(P, us, recs):
        ref = M.Engine(p, lanes=64, use_lds=0).run(u)
        assert np.array_equal(r.view(np.uint8), ref.view(np.uint8))
    for gas in ("He", "N2"):
        k = min((i for i in range(len(P)) if cfgs[i]["GasBuffer"] == gas), key=lambda i: P[i].info.natoms)
        refs = oracle_refs(P[k], cfgs[k], us[k])
        assert_traj_parity(recs[k], refs, "assembly %d (%s, %d atoms)" % (k, gas, P[k].info.natoms),
                           min_big=5)
