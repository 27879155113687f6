"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle on
the same prepared target and the same seeded samples.

Criteria (per point): |dF_k| <= 1e-12 * sum_pairs |F_k,pair| and likewise for
U (the reference's own fast-math and strict builds reorder these sums, so a
plain relative error is not attainable where the total cancels).  Every site
kernel the engine compiles is a case: the 16 reference kernels of
src/Force.cpp through the force types and site roles that select them.

Per trajectory (common.assert_traj_parity): identical hit flags everywhere;
on every non-chaotic hit (the oracle's tries, fate and chi unchanged at
dt/2 and from an impact parameter moved by 4 ulp) the same tries and fate,
|dchi| <= 1e-8 rad for chi > 1e-6 rad and |d(1 - cos chi)| <= 1e-15
below; no fraction may fail.
"""
import numpy as np
import pytest

from common import (LONG, M, O, C60_XYZ, UBQ_PQR, assert_traj_parity, c60_lj_cfg, globule_cfg,
                    gpu_available, oracle_for, oracle_refs, pair_scales, prepared)

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

REL = 1e-12


def _points(P, n, rng, spread=1.3):
    i = P.info
    h = np.array(i.half_box)
    pts = rng.uniform(-spread, spread, size=(n, 3)) * np.array([i.a, i.b, i.c]) * 0.5
    # a few points near atoms and a few outside the grid
    a = P.atoms()
    k = rng.integers(0, i.natoms, size=n // 4)
    pts[: n // 4] = np.stack([a["x"][k], a["y"][k], a["z"][k]], 1) + rng.normal(0, 2.5, (n // 4, 3))
    pts[-4:] = h * np.array([[1.02, 0, 0], [0, -1.01, 0.2], [0.3, 0.1, 1.05], [-1.5, -1.5, -1.5]])
    return pts


def _flags(ft):
    """input.json flags that select force type ft (src/System.cpp:49-87):
    types 1-6 for N2/CO2; for the one-site probes 3 and 4 are the dipole types."""
    s = "yes" if ft in (2, 4, 6) else "no"
    lr = "no" if ft in (1, 2) else "yes"
    lc = "yes" if ft in (4, 6) else "no"
    return {"Short-range cutoff": s, "Long-range forces": lr, "Long-range cutoff": lc}


def _one_site_flags(ft):
    return {1: {"Short-range cutoff": "no", "Long-range forces": "no", "polarizability": "no"},
            2: {"Short-range cutoff": "yes", "Long-range forces": "no", "polarizability": "no"},
            3: {"Short-range cutoff": "no", "Long-range forces": "yes", "Long-range cutoff": "no",
                "polarizability": "yes"},
            4: {"Short-range cutoff": "yes", "Long-range forces": "yes", "Long-range cutoff": "yes",
                "polarizability": "yes"}}[ft]


def _n2(ft):
    return dict(GasBuffer="N2", polarizability="yes" if ft in (5, 6) else "no", **_flags(ft))


def _co2(ft):
    return dict(GasBuffer="CO2", polarizability="yes" if ft in (5, 6) else "no", **_flags(ft))


CASES = [
    # name, cfg, site role, kernel name, probe charge, terms (lj, coul, dip), cutoff
    ("c60_He_t4", dict(targetFileName=C60_XYZ, GasBuffer="He", **_one_site_flags(4)),
     M.SITE_ATOM, "lennardjones_induced_dipole_LC", 0.0, (True, False, True), True),
    ("ubq_He_t4", dict(targetFileName=UBQ_PQR, GasBuffer="He", **_one_site_flags(4)),
     M.SITE_ATOM, "lennardjones_induced_dipole_LC", 0.0, (True, False, True), True),
    ("ubq_He_t3", dict(targetFileName=UBQ_PQR, GasBuffer="He", **_one_site_flags(3)),
     M.SITE_ATOM, "lennardjones_induced_dipole", 0.0, (True, False, True), False),
    ("ubq_He_t2", dict(targetFileName=UBQ_PQR, GasBuffer="He", **_one_site_flags(2)),
     M.SITE_ATOM, "lennardjones_LC", 0.0, (True, False, False), True),
    ("ubq_He_t1", dict(targetFileName=UBQ_PQR, GasBuffer="He", **_one_site_flags(1)),
     M.SITE_ATOM, "lennardjones", 0.0, (True, False, False), False),
    ("ubq_Ar_t4", dict(targetFileName=UBQ_PQR, GasBuffer="Ar", **_one_site_flags(4)),
     M.SITE_ATOM, "lennardjones_induced_dipole_LC", 0.0, (True, False, True), True),
    ("ubq_N2_t1_N", dict(targetFileName=UBQ_PQR, **_n2(1)),
     M.SITE_N, "lennardjones", -0.4825, (True, False, False), False),
    ("ubq_N2_t2_N", dict(targetFileName=UBQ_PQR, **_n2(2)),
     M.SITE_N, "lennardjones_LC", -0.4825, (True, False, False), True),
    ("ubq_N2_t3_N", dict(targetFileName=UBQ_PQR, **_n2(3)),
     M.SITE_N, "lennardjones_coulomb", -0.4825, (True, True, False), False),
    ("ubq_N2_t3_D", dict(targetFileName=UBQ_PQR, **_n2(3)),
     M.SITE_DUMMY, "coulomb", 0.965, (False, True, False), False),
    ("ubq_N2_t4_N", dict(targetFileName=UBQ_PQR, **_n2(4)),
     M.SITE_N, "lennardjones_coulomb_LC", -0.4825, (True, True, False), True),
    ("ubq_N2_t4_D", dict(targetFileName=UBQ_PQR, **_n2(4)),
     M.SITE_DUMMY, "coulomb_LC", 0.965, (False, True, False), True),
    ("ubq_N2_t5_D", dict(targetFileName=UBQ_PQR, **_n2(5)),
     M.SITE_DUMMY, "coulomb_induced_dipole_iso", 0.965, (False, True, True), False),
    ("ubq_N2_t6_N", dict(targetFileName=UBQ_PQR, **_n2(6)),
     M.SITE_N, "lennardjones_coulomb_LC", -0.4825, (True, True, False), True),
    ("ubq_N2_t6_D", dict(targetFileName=UBQ_PQR, **_n2(6)),
     M.SITE_DUMMY, "coulomb_induced_dipole_iso_LC", 0.965, (False, True, True), True),
    ("ubq_CO2_t1_C", dict(targetFileName=UBQ_PQR, **_co2(1)),
     M.SITE_C, "lennardjones_CO2", 0.6512, ("central", False, False), False),
    ("ubq_CO2_t2_C", dict(targetFileName=UBQ_PQR, **_co2(2)),
     M.SITE_C, "lennardjones_LC_CO2", 0.6512, ("central", False, False), True),
    ("ubq_CO2_t3_C", dict(targetFileName=UBQ_PQR, **_co2(3)),
     M.SITE_C, "lennardjones_coulomb_CO2", 0.6512, ("central", True, False), False),
    ("ubq_CO2_t4_C", dict(targetFileName=UBQ_PQR, **_co2(4)),
     M.SITE_C, "lennardjones_coulomb_LC_CO2", 0.6512, ("central", True, False), True),
    ("ubq_CO2_t5_C", dict(targetFileName=UBQ_PQR, **_co2(5)),
     M.SITE_C, "lennardjones_coulomb_induced_dipole_iso_CO2", 0.6512, ("central", True, True),
     False),
    ("ubq_CO2_t6_C", dict(targetFileName=UBQ_PQR, **_co2(6)),
     M.SITE_C, "lennardjones_coulomb_induced_dipole_iso_LC_CO2", 0.6512, ("central", True, True),
     True),
    ("ubq_CO2_t6_O", dict(targetFileName=UBQ_PQR, **_co2(6)),
     M.SITE_O, "lennardjones_coulomb_LC", -0.3256, (True, True, False), True),
]
# every reference kernel is exercised by at least one case
assert sorted(set(c[3] for c in CASES)) == sorted(O.KERNELS)


@pytest.mark.parametrize("name,cfg,site,kernel,qi,terms,cut", CASES, ids=[c[0] for c in CASES])
def test_site_forces_match_oracle(name, cfg, site, kernel, qi, terms, cut):
    cfg = dict(cfg, seed=2104)
    print("case", name, flush=True)
    P = prepared(cfg)
    orc = oracle_for(P, cfg)
    rng = np.random.default_rng(7)
    pts = _points(P, 200, rng)
    E = M.Engine(P)
    f, U = E.eval_sites(site, pts)
    worst = 0.0
    for k in range(len(pts)):
        fr, Ur = orc.site_force(kernel, pts[k], qi)
        sf, su = pair_scales(P, cfg, pts[k], qi, terms[0], terms[1], terms[2], cut)
        err = max(np.max(np.abs(f[k] - fr) / sf), abs(U[k] - Ur) / su)
        worst = max(worst, err)
        assert err <= REL, (name, k, pts[k], f[k], fr, U[k], Ur)
    print(name, "worst relative error", worst)


def _traj_case(cfg, n, name, key=None, **engine_kw):
    print("case", name, n, engine_kw, flush=True)
    P = prepared(cfg)
    u = P.samples(n)
    refs = oracle_refs(P, cfg, u, key)
    rec = M.Engine(P, **engine_kw).run(u)
    assert_traj_parity(rec, refs, name)
    return P, u, rec, refs[0]


@pytest.mark.parametrize("lanes", [16, 64])
def test_trajectories_c60_he(lanes):
    cfg = dict(targetFileName=C60_XYZ, GasBuffer="He", polarizability="yes", seed=2104, **LONG)
    P, u, rec, ref = _traj_case(cfg, 4000, "c60_He_L%d" % lanes, key="c60_He_4000", lanes=lanes)
    assert np.array_equal(u, oracle_for(P, cfg).samples(4000))


def test_trajectories_ubq_he():
    cfg = dict(targetFileName=UBQ_PQR, GasBuffer="He", polarizability="yes", seed=2104, **LONG)
    _traj_case(cfg, 2000, "ubq_He")


def test_trajectories_ubq_n2():
    cfg = dict(targetFileName=UBQ_PQR, GasBuffer="N2", polarizability="yes", seed=2104, **LONG)
    _traj_case(cfg, 2000, "ubq_N2")


@pytest.mark.parametrize("cutoff", [True, False], ids=["type2", "type1"])
def test_trajectories_c60_lj(workdir, cutoff):
    cfg = c60_lj_cfg(workdir, cutoff=cutoff)
    P = prepared(cfg)
    assert P.info.force_type == (2 if cutoff else 1)
    _traj_case(cfg, 3000, "c60_LJ_t%d" % P.info.force_type)


def test_trajectories_globule_n2(workdir):
    cfg = globule_cfg(workdir)
    P = prepared(cfg)
    assert P.info.force_type == 6
    _traj_case(cfg, 1000, "globule_N2", key="globule_1000")


@pytest.mark.parametrize("gas,n,lanes", [("Ar", 2000, 16), ("CO2", 600, 64), ("co2", 1000, 16)])
def test_trajectories_other_probes(gas, n, lanes):
    """Ar (He integrator with the Ar mass and mixed LJ), the rigid 3-site CO2
    (Ciccotti forces and the reference's vcm quirk) and the 1-site co2."""
    cfg = dict(targetFileName=C60_XYZ if gas != "CO2" else UBQ_PQR, GasBuffer=gas,
               polarizability="yes", seed=2104, **LONG)
    _traj_case(cfg, n, gas, lanes=lanes)


# The force types no configuration of the benchmark uses, each through its
# trajectory kernel: the one-site probe without cutoffs (1) and with the
# all-atom dipole (3), N2 and CO2 types 1-5.
OTHER_TYPES = ([("He", ft) for ft in (1, 3)] + [("N2", ft) for ft in (1, 2, 3, 4, 5)] +
               [("CO2", ft) for ft in (1, 2, 3, 4, 5)])


@pytest.mark.parametrize("gas,ft", OTHER_TYPES, ids=["%s_t%d" % c for c in OTHER_TYPES])
def test_trajectories_other_force_types(gas, ft):
    flags = _one_site_flags(ft) if gas == "He" else (_n2(ft) if gas == "N2" else _co2(ft))
    cfg = dict(targetFileName=C60_XYZ, seed=2104, **dict(flags, GasBuffer=gas))
    P = prepared(cfg)
    assert P.info.force_type == ft
    _traj_case(cfg, 600 if gas == "He" else 300, "c60_%s_t%d" % (gas, ft))


def test_trajectories_global_memory_atoms(workdir):
    """The path large targets take: atoms read from HBM (not staged in LDS),
    32-bit neighbour lists."""
    cfg = globule_cfg(workdir)
    P, u, rec, ref = _traj_case(cfg, 1000, "globule_N2_global", key="globule_1000", use_lds=0)
    # identical to the LDS-staged run bit for bit (same lists, same order)
    lds = M.Engine(P, use_lds=1).run(u)
    assert np.array_equal(lds.view(np.uint8), rec.view(np.uint8))


@pytest.mark.parametrize("use_lds,cap", [(0, 64), (1, 600)])
def test_trajectories_list_overflow(workdir, use_lds, cap):
    """Verlet lists that outgrow their capacity: every list at 64 entries,
    some at 600 of the globule's 1,231 atoms.  The kernel walks the cell
    columns instead, switching between list and walk within a trajectory."""
    cfg = globule_cfg(workdir)
    _traj_case(cfg, 1000, "globule_N2_cap%d" % cap, key="globule_1000", use_lds=use_lds,
               list_cap=cap)


def test_tally_device_matches_host():
    import torch
    cfg = dict(targetFileName=C60_XYZ, GasBuffer="He", polarizability="yes", seed=2104, **LONG)
    P = prepared(cfg)
    E = M.Engine(P)
    u = P.samples(3000)
    rec = E.run(u)
    om, cnt, mean, err = M.tally(rec, 3, 1000)
    om_o, cnt_o, mean_o, err_o = O.tally(rec[["hit", "success", "tries", "steps", "chi",
                                              "domega"]].astype(O.RECORD_DTYPE), 3, 1000)
    assert np.array_equal(om, om_o) and np.array_equal(cnt, cnt_o)
    d_rec = torch.from_numpy(rec.view(np.uint8).copy()).cuda()
    d_om = torch.zeros(3, dtype=torch.float64, device="cuda")
    d_cnt = torch.zeros(9, dtype=torch.int64, device="cuda")
    d_me = torch.zeros(2, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    E.tally_device(d_rec.data_ptr(), 3, 1000, d_om.data_ptr(), d_cnt.data_ptr(), d_me.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(d_om.cpu().numpy(), om)
    assert np.array_equal(d_cnt.cpu().numpy().reshape(3, 3), cnt)
    assert abs(d_me[0].item() - mean) <= 1e-12 * abs(mean)


def test_restarts_run_side_by_side_with_identical_results(workdir):
    """Later tries of a restarting trajectory may run concurrently on idle
    groups; the fold over their outcomes must give exactly the records of
    running them one after another, whatever the schedule (each try builds
    its own neighbour list, so even the rounding is schedule-independent).
    Likewise the order in which trajectories start (likely-long first, or
    index order) leaves every record unchanged."""
    cfg = globule_cfg(workdir)
    print("case globule_N2 restarts", flush=True)
    P = prepared(cfg)
    u = P.samples(4000)
    runs = []
    # the three spawning thresholds in the start order (likely-long first),
    # then once in index order
    for spec, order in ((0, 0), (-1, 0), (10 ** 9, 0), (-1, 1)):
        E = M.Engine(P, spec_below=spec, order=order)
        runs.append(E.run(u))
        w = E.last_work()
        assert w["steps"] >= runs[-1]["steps"].sum()
        E.close()
    for r in runs[1:]:
        assert np.array_equal(r.view(np.uint8), runs[0].view(np.uint8))
    r = runs[0]
    assert (r["tries"] >= 4).sum() >= 5  # the sample holds long restart chains
    # the three longest chains against the oracle: same restarts and fate.  The
    # orbits are chaotic (10^5 steps and more): any last-bit change of the
    # forces (summation order, fused multiply-adds) moves their exit, so their
    # length agrees only roughly
    idx = np.argsort(r["steps"])[-3:]
    ref = oracle_for(P, cfg).run(u[idx])
    assert np.array_equal(ref["tries"], r["tries"][idx])
    assert np.array_equal(ref["success"], r["success"][idx])
    assert np.all(np.abs(ref["steps"] - r["steps"][idx].astype(np.int64)) <= 0.15 * ref["steps"])


def test_trajectories_co2_corrected_mode():
    """The corrected CO2 chi reference ("CO2 vcm fix": "yes") against the
    oracle in the same mode; the fates equal those of the reference mode."""
    cfg = dict(targetFileName=UBQ_PQR, GasBuffer="CO2", polarizability="yes", seed=2104,
               **{"CO2 vcm fix": "yes"}, **LONG)
    P, u, rec, ref = _traj_case(cfg, 600, "CO2_fix")
    base = {k: v for k, v in cfg.items() if k != "CO2 vcm fix"}
    quirk = M.Engine(prepared(base)).run(u)
    assert np.array_equal(quirk["hit"], rec["hit"])
    assert np.array_equal(quirk["success"], rec["success"])
    assert np.array_equal(quirk["tries"], rec["tries"])
