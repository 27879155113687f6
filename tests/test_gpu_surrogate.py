"""The surrogate network's GPU kernels (BF16 MFMA, FP32 accumulation) against
the CPU oracle and its training; the range-split far field (near pair terms
exact below the shell, the remainder from a cubic B-spline table) per point
against the oracle and in trajectories: CCS against the exact method on the
same samples, in single-target and batched runs, and what the far field and
the network gain over the near terms alone."""
import os
import sys
import time

import numpy as np
import pytest

from common import LONG, M, O, UBQ_PQR, globule_cfg, gpu_available, oracle_for, pair_scales, prepared

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as MO  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _net(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(MO.NPARAM, dtype=np.float32)
    for lo, fin, fout in [(MO.W1, 64, 128), (MO.W2, 128, 128), (MO.W3, 128, 128), (MO.W4, 128, 2)]:
        a = np.sqrt(6.0 / (fin + fout))
        p[lo:lo + fin * fout] = rng.uniform(-a, a, fin * fout)
    p[MO.B1:MO.W2] = rng.normal(0, 0.1, 128)
    p[MO.B4:] = rng.normal(0, 0.1, 2)
    freq = rng.normal(0, 2.0, (32, 3))
    centre = np.array([1.0, -2.0, 0.5])
    s, osc = 1 / 30.0, np.array([2.0, 0.5])
    return p, freq, centre, s, osc, rng


def test_mfma_forward_and_position_gradient_match_oracle():
    print("case mlp eval", flush=True)
    p, freq, centre, s, osc, rng = _net()
    xyz = rng.uniform(-30, 30, (3000, 3))
    spec = M.mlp_spec(freq.astype(np.float32), centre, s, osc)
    y, dy = M.mlp_eval(spec, p, xyz)
    f32 = freq.astype(np.float32).astype(np.float64)
    yr, dyr = MO.eval_with_grad(p, xyz.astype(np.float32).astype(np.float64), f32, centre, s, osc, bf=True)
    for o in range(2):
        scale = np.sqrt(np.mean(yr[:, o] ** 2))
        assert np.max(np.abs(y[:, o] - yr[:, o])) <= 1e-2 * scale
        gs = np.sqrt(np.mean(dyr[:, o] ** 2))
        assert np.sqrt(np.mean((dy[:, o] - dyr[:, o]) ** 2)) <= 2e-2 * gs


def test_mfma_loss_gradient_matches_oracle():
    print("case mlp loss grad", flush=True)
    p, freq, centre, s, osc, rng = _net(1)
    # 4,093 points: the batch-wide products end in a partial k tile whose
    # last chunk is not a whole 16-byte vector
    xyz = rng.uniform(-30, 30, (4093, 3))
    e = np.stack([np.sin(xyz[:, 0] / 7.0) * 2.0, np.cos(xyz[:, 1] / 5.0) * 0.5], 1)
    spec = M.mlp_spec(freq.astype(np.float32), centre, s, osc)
    loss, g = M.mlp_loss_grad(spec, p, xyz, e)
    f32 = freq.astype(np.float32).astype(np.float64)
    lr, gr = MO.loss_grad(p, xyz.astype(np.float32).astype(np.float64), e, f32, centre, s, osc, bf=True)
    assert abs(loss - lr) <= 1e-3 * lr
    blocks = [(MO.W1, MO.B1), (MO.B1, MO.W2), (MO.W2, MO.B2), (MO.B2, MO.W3), (MO.W3, MO.B3),
              (MO.B3, MO.W4), (MO.W4, MO.B4), (MO.B4, MO.NPARAM)]
    for lo, hi in blocks:
        err = np.linalg.norm(g[lo:hi] - gr[lo:hi]) / max(np.linalg.norm(gr[lo:hi]), 1e-30)
        assert err <= 2e-2, (lo, hi, err)


def test_mfma_forward_mode_loss_gradient_matches_oracle():
    """Value and tangent rows (the outputs' derivative along a unit vector
    per point, carried forward through the network) on the GPU against the
    oracle's forward-mode restatement with the GPU's BF16 roundings: the
    outputs and derivatives, the loss and every parameter block's gradient."""
    print("case mlp jvp loss grad", flush=True)
    p, freq, centre, s, osc, rng = _net(3)
    n = 3001  # a partial last row tile
    xyz = rng.uniform(-30, 30, (n, 3))
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    e = np.stack([np.sin(xyz[:, 0] / 7.0) * 2.0, np.cos(xyz[:, 1] / 5.0) * 0.5], 1)
    d = np.stack([np.cos(xyz[:, 0] / 7.0) * 2.0 / 7.0 * dirs[:, 0],
                  -np.sin(xyz[:, 1] / 5.0) * 0.5 / 5.0 * dirs[:, 1]], 1)
    w = 4.0
    spec = M.mlp_spec(freq.astype(np.float32), centre, s, osc)
    loss, g, y, dy = M.mlp_jvp_loss_grad(spec, p, xyz, dirs, e, d, w)
    f32 = freq.astype(np.float32).astype(np.float64)
    x32 = xyz.astype(np.float32).astype(np.float64)
    u32 = dirs.astype(np.float32).astype(np.float64)
    lr, gr, yr, dyr = MO.jvp_loss_grad(p, x32, u32, e, d, w, f32, centre, s, osc, bf=True)
    for o in range(2):
        assert np.max(np.abs(y[:, o] - yr[:, o])) <= 1e-2 * np.sqrt(np.mean(yr[:, o] ** 2))
        assert np.sqrt(np.mean((dy[:, o] - dyr[:, o]) ** 2)) <= 2e-2 * np.sqrt(np.mean(dyr[:, o] ** 2))
    assert abs(loss - lr) <= 1e-3 * lr
    blocks = [(MO.W1, MO.B1), (MO.B1, MO.W2), (MO.W2, MO.B2), (MO.B2, MO.W3), (MO.W3, MO.B3),
              (MO.B3, MO.W4), (MO.W4, MO.B4), (MO.B4, MO.NPARAM)]
    for lo, hi in blocks:
        err = np.linalg.norm(g[lo:hi] - gr[lo:hi]) / max(np.linalg.norm(gr[lo:hi]), 1e-30)
        assert err <= 2e-2, (lo, hi, err)


def test_training_fits_a_smooth_field():
    print("case mlp train", flush=True)
    p, freq, centre, s, osc, rng = _net(2)
    freq = rng.normal(0, 30 / 12.0, (32, 3))
    xyz = rng.uniform(-30, 30, (60000, 3))

    def field(x):
        r1 = np.linalg.norm(x - [8, 0, 0], axis=1) + 3.0
        r2 = np.linalg.norm(x + [5, 6, 0], axis=1) + 3.0
        return np.stack([10.0 / r1 - 6.0 / r2, np.exp(-np.sum(x ** 2, 1) / 400.0)], 1)

    e = field(xyz)
    osc = np.sqrt(np.mean(e ** 2, 0))
    spec = M.mlp_spec(freq.astype(np.float32), centre, s, osc)
    l0, _ = M.mlp_loss_grad(spec, p, xyz[:4096], e[:4096])
    p2, hist = M.mlp_train(spec, p, xyz, e, epochs=60, batch=1024, lr=3e-3)
    print("loss %.4g -> %.4g" % (l0, hist[-1]), flush=True)
    assert hist[-1] < hist[0] and hist[-1] < 0.02 * l0
    xv = rng.uniform(-25, 25, (2000, 3))
    y, _ = M.mlp_eval(spec, p2, xv, grad=False)
    ev = field(xv)
    assert np.sqrt(np.mean((y - ev) ** 2)) <= 0.1 * np.sqrt(np.mean(ev ** 2))


def test_training_is_reproducible():
    """The batch-wide weight-gradient products are split over blocks and the
    partial products summed in a fixed order: two trainings from the same
    start give the same weights bit for bit."""
    print("case mlp train twice", flush=True)
    p, freq, centre, s, osc, rng = _net(3)
    xyz = rng.uniform(-30, 30, (20000, 3))
    e = np.stack([np.sin(xyz[:, 0] / 7.0), np.cos(xyz[:, 2] / 9.0)], 1)
    spec = M.mlp_spec(freq.astype(np.float32), centre, s, np.sqrt(np.mean(e ** 2, 0)))
    a, ha = M.mlp_train(spec, p, xyz, e, epochs=3, batch=8192, lr=3e-3)
    b, hb = M.mlp_train(spec, p, xyz, e, epochs=3, batch=8192, lr=3e-3)
    assert not np.array_equal(a, p)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_surrogate_build_is_reproducible(workdir):
    """Two builds of the far-field network for one engine (training points,
    exact values, training) give the same parameters bit for bit."""
    print("case surrogate build twice", flush=True)
    P = prepared(globule_cfg(workdir))
    E = M.Engine(P)
    _, pa, _ = E.build_surrogate(n_train=40000, epochs=2)
    _, pb, _ = E.build_surrogate(n_train=40000, epochs=2)
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))


def _switched_near_terms(P, r, qi, lj, coul, dip, sw1, sw2, rlj=12.0, rc=25.0):
    """Numpy restatement of the near part of the range-split site force:
    LJ (unchanged, r < rlj) plus S(r) times the reference's shifted Coulomb
    potential u = 1/r - 1.5/rc + r^2/(2 rc^3) and the dipole field of
    S u (E = -grad, grad E = -Hess); S = 1 - t^3 (10 - 15 t + 6 t^2),
    t = (r - sw1) / (sw2 - sw1) clamped to [0, 1].  Returns F, U and a
    scale of the summed magnitudes."""
    a = P.atoms()
    K = 332.06348078020324
    d = np.stack([r[0] - a["x"], r[1] - a["y"], r[2] - a["z"]], 1)
    r2 = np.sum(d * d, 1)
    rr = np.sqrt(r2)
    F, U, sc = np.zeros(3), 0.0, 1.0
    if lj:
        m = rr < rlj
        l1 = 4 * a["eps"] * a["sig"] ** 6
        l2 = l1 * a["sig"] ** 6
        U += np.sum((l2 / r2 ** 6 - l1 / r2 ** 3 - (l2 / rlj ** 12 - l1 / rlj ** 6))[m])
        fl = ((12 * l2 / r2 ** 7 - 6 * l1 / r2 ** 4)[:, None] * d)[m]
        F += fl.sum(0)
        sc += np.abs(fl).sum() + np.sum(np.abs(l2 / r2 ** 6)[m]) + np.sum(np.abs(l1 / r2 ** 3)[m])
    m = rr < sw2
    t = np.clip((rr - sw1) / (sw2 - sw1), 0, 1)
    w = sw2 - sw1
    S = 1 - t ** 3 * (10 - 15 * t + 6 * t * t)
    S1 = -30 * t * t * (1 - t) ** 2 / w
    S2 = -60 * t * (1 - t) * (1 - 2 * t) / w ** 2
    u = 1 / rr - 1.5 / rc + rr * rr / (2 * rc ** 3)
    u1 = -1 / rr ** 2 + rr / rc ** 3
    u2 = 2 / rr ** 3 + 1 / rc ** 3
    ph, ph1, ph2 = S * u, S1 * u + S * u1, S2 * u + 2 * S1 * u1 + S * u2
    q = a["q"]
    E, G = np.zeros(3), np.zeros((3, 3))
    A = (-q * ph1 / rr)[m]
    if coul:
        U += qi * K * np.sum((q * ph)[m])
        fc = qi * K * (A[:, None] * d[m])
        F += fc.sum(0)
        sc += np.abs(fc).sum() + np.sum(np.abs(qi * K * q * ph)[m])
    if dip:
        B = (-q * (ph2 - ph1 / rr) / r2)[m]
        dm = d[m]
        E = np.sum(A[:, None] * dm, 0)
        G = np.sum(A)[None, None] * np.eye(3) + np.einsum("k,ki,kj->ij", B, dm, dm)
        alpha = P.info.alpha
        U -= 0.5 * alpha * E @ E
        F += alpha * G @ E
        sE = np.sum(np.abs(A) * np.linalg.norm(dm, axis=1))
        sG = np.sum(np.abs(A) * 3 + np.abs(B) * r2[m])
        sc += alpha * sE * (sE + sG) * 3
    return F, U, sc


@pytest.mark.parametrize("gas,site,qi,terms", [
    ("N2", M.SITE_N, -0.4825, (True, True, False)),
    ("N2", M.SITE_DUMMY, 0.965, (False, True, True)),
    ("He", M.SITE_ATOM, 0.0, (True, False, True))])
def test_split_near_terms_match_closed_form(gas, site, qi, terms):
    """The near part of the range-split evaluation (the far field left out,
    table_source 2) through the engine against a numpy restatement of the
    switched terms, to 1e-12 of the summed magnitudes: pins the switch, its
    derivatives in the field gradient and the dipole epilogue."""
    cfg = dict(targetFileName=UBQ_PQR, GasBuffer=gas, polarizability="yes", seed=2104, **LONG)
    print("case split near", gas, site, flush=True)
    P = prepared(cfg)
    E = M.Engine(P)
    E.build_surrogate(shell=6.0, blend=6.0, grid=0.5, table_source=2)
    E.use_surrogate(True)
    rng = np.random.default_rng(5)
    a = P.atoms()
    k = rng.integers(0, P.info.natoms, 200)
    pts = np.stack([a["x"][k], a["y"][k], a["z"][k]], 1) + rng.normal(0, 5.0, (200, 3))
    f, U = E.eval_sites(site, pts)
    worst = 0.0
    for i in range(len(pts)):
        if np.min(np.sum((np.stack([a["x"], a["y"], a["z"]], 1) - pts[i]) ** 2, 1)) < 1.0:
            continue
        fr, ur, sc = _switched_near_terms(P, pts[i], qi, *terms, sw1=6.0, sw2=12.0)
        err = max(np.max(np.abs(f[i] - fr)), abs(U[i] - ur)) / sc
        worst = max(worst, err)
        assert err <= 1e-12, (i, pts[i], f[i], fr, U[i], ur)
    print("worst", worst, flush=True)


@pytest.mark.parametrize("gas,site,kernel,qi", [
    ("N2", M.SITE_N, "lennardjones_coulomb_LC", -0.4825),
    ("N2", M.SITE_DUMMY, "coulomb_induced_dipole_iso_LC", 0.965),
    ("He", M.SITE_ATOM, "lennardjones_induced_dipole_LC", 0.0)])
def test_split_with_tabulated_far_field_approximates_oracle(gas, site, kernel, qi):
    """Near terms plus the tabulated exact far field (0.5 A grid) against the
    oracle's exact site force at points a probe reaches (2 A or more from
    every atom): the interpolation error stays within 0.5 % of the force
    scale on average and 3 % at worst."""
    cfg = dict(targetFileName=UBQ_PQR, GasBuffer=gas, polarizability="yes", seed=2104, **LONG)
    print("case split table", gas, kernel, flush=True)
    P = prepared(cfg)
    orc = oracle_for(P, cfg)
    E = M.Engine(P)
    E.build_surrogate(shell=6.0, blend=6.0, grid=0.5, exact_table=True)
    E.use_surrogate(True)
    rng = np.random.default_rng(6)
    a = P.atoms()
    A = np.stack([a["x"], a["y"], a["z"]], 1)
    k = rng.integers(0, P.info.natoms, 400)
    pts = A[k] + rng.normal(0, 6.0, (400, 3))
    keep = np.array([np.min(np.sum((A - x) ** 2, 1)) >= 4.0 for x in pts])
    pts = pts[keep]
    f, U = E.eval_sites(site, pts)
    ref = [orc.site_force(kernel, x, qi) for x in pts]
    fr = np.array([r[0] for r in ref])
    df = np.linalg.norm(f - fr, axis=1)
    scale = np.sqrt(np.mean(np.sum(fr ** 2, 1)))
    print("%d points: rms |dF| %.3g, max %.3g, rms |F| %.3g" % (len(pts), np.sqrt(np.mean(df ** 2)),
                                                           df.max(), scale), flush=True)
    assert np.sqrt(np.mean(df ** 2)) <= 5e-3 * scale
    assert df.max() <= 3e-2 * scale


def _c4():
    from massccs_amd import synth
    s, xyz, q = synth.globule(n=50000, axes=(60.0, 45.0, 35.0), net_charge=20.0, seed=2104)
    cfg = {"targetFileName": "memory.xyz", "GasBuffer": "N2", "Short-range cutoff": "yes",
           "LJ-cutoff": 12.0, "Long-range forces": "yes", "Long-range cutoff": "yes",
           "Coul-cutoff": 25.0, "polarizability": "yes", "alpha": 1.7403, "seed": 2104,
           "dt": 10.0, "Temp": 298.0}
    return cfg, s, xyz, q


def _paired(rs, rp):
    den = float(((rp["hit"] == 0) | (rp["success"] == 1)).sum())
    cp = rp["domega"].sum() / den
    cs = rs["domega"].sum() / float(((rs["hit"] == 0) | (rs["success"] == 1)).sum())
    se = float(np.std(rs["domega"] - rp["domega"]) * np.sqrt(len(rp)) / den) / cp
    return cs / cp - 1.0, se, cp, cs


def test_range_split_ccs_within_half_percent_on_the_50k_ellipsoid():
    """The headline configuration (50,000 atoms, N2, force type 6): the
    range-split far field (exact below 6 A, switched off at 12 A, exact far
    field tabulated at 0.5 A) against the exact method on the same samples,
    3 x 10^5 trajectories from each of three sample streams.  A tenth of
    the trajectories are chaotic (rounding alone moves their angle), so the
    comparison is statistical: the pooled CCS error must be within 0.5 %,
    every stream's within 0.5 % + 2 paired standard errors; the surrogate
    must also be faster."""
    cfg, s, xyz, q = _c4()
    errs, ses, t_plain, t_sur = [], [], 0.0, 0.0
    sums = np.zeros(4)
    for seed in (2104, 1, 2):
        print("case c4 split seed", seed, flush=True)
        P = M.Prepared(M.config_from_dict(dict(cfg, seed=seed)), symbols=list(s), xyz=xyz, q=q)
        u = P.samples(300000)
        E = M.Engine(P)
        E.run(u[:5000])
        t0 = time.time()
        rp = E.run(u)
        t_plain += time.time() - t0
        E.build_surrogate(shell=6.0, blend=6.0, grid=0.5, exact_table=True)
        E.use_surrogate(True)
        E.run(u[:5000])
        t0 = time.time()
        rs = E.run(u)
        t_sur += time.time() - t0
        err, se, cp, cs = _paired(rs, rp)
        print("seed %d: ccs plain %.2f split %.2f err %.3f %% (paired SE %.3f %%)"
              % (seed, cp, cs, 100 * err, 100 * se), flush=True)
        assert np.array_equal(rs["hit"], rp["hit"])
        assert abs(err) <= 0.005 + 2 * se
        errs.append(err)
        ses.append(se)
        sums += [rp["domega"].sum(), float(((rp["hit"] == 0) | (rp["success"] == 1)).sum()),
                 rs["domega"].sum(), float(((rs["hit"] == 0) | (rs["success"] == 1)).sum())]
    pooled = (sums[2] / sums[3]) / (sums[0] / sums[1]) - 1.0
    print("pooled error %.3f %%, time plain %.2f s split %.2f s (%.2fx)"
          % (100 * pooled, t_plain, t_sur, t_plain / t_sur), flush=True)
    assert abs(pooled) <= 0.005
    assert t_sur < t_plain


@pytest.mark.parametrize("lanes", [64, 32, 16])
def test_batched_range_split_equals_single_target_runs(workdir, lanes):
    """The batch path with far fields (one persistent launch per probe kind,
    graph replay) gives the records of single-target split runs of the same
    group size byte for byte; members without a dipole field keep the exact
    kernel."""
    print("case batch split, %d lanes" % lanes, flush=True)
    cfgs = [globule_cfg(workdir),
            dict(targetFileName=UBQ_PQR, GasBuffer="He", polarizability="yes", seed=7, **LONG),
            dict(targetFileName=UBQ_PQR, GasBuffer="N2", polarizability="yes", seed=11, **LONG)]
    P = [prepared(c) for c in cfgs]
    us = [p.samples(n) for p, n in zip(P, (800, 1500, 600))]
    B = M.Batch(P, lanes=lanes, skin=1.0)
    reps = B.build_surrogate(shell=6.0, blend=6.0, grid=0.5, exact_table=True)
    assert all(r["grid_nodes"] > 0 for r in reps)
    B.use_surrogate(True)
    recs = B.run(us)
    for p, u, r in zip(P, us, recs):
        E = M.Engine(p, lanes=lanes, use_lds=0)
        E.build_surrogate(shell=6.0, blend=6.0, grid=0.5, exact_table=True)
        E.use_surrogate(True)
        ref = E.run(u)
        assert np.array_equal(r.view(np.uint8), ref.view(np.uint8))
    again = B.run(us)  # graph replay
    for a, b in zip(recs, again):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    B.use_surrogate(False)
    exact = B.run(us)
    ref = M.Engine(P[0], lanes=lanes, use_lds=0).run(us[0])
    assert np.array_equal(exact[0].view(np.uint8), ref.view(np.uint8))


def test_split_site_forces_agree_across_lane_counts(workdir):
    """The far field's 64 B-spline nodes are shared out over the lanes of a
    trajectory's group (one node per lane at 64 lanes, two at 32, four at
    16): the split site forces and energies at 32 and 16 lanes per point
    equal those at 64 to 1e-12 of the summed magnitudes (only the order of
    the sums differs)."""
    print("case split lanes", flush=True)
    cfg = globule_cfg(workdir)
    P = prepared(cfg)
    rng = np.random.default_rng(5)
    a = P.atoms()
    k = rng.integers(0, P.info.natoms, 200)
    pts = np.stack([a["x"][k], a["y"][k], a["z"][k]], 1) + rng.normal(0, 6.0, (200, 3))
    res = {}
    for L in (64, 32, 16):
        E = M.Engine(P, lanes=L)
        E.build_surrogate(shell=6.0, blend=6.0, grid=0.5, exact_table=True)
        E.use_surrogate(True)
        res[L] = {site: E.eval_sites(site, pts) for site in (M.SITE_N, M.SITE_DUMMY)}
    for site, qi, terms in [(M.SITE_N, -0.4825, (True, True, False)), (M.SITE_DUMMY, 0.965, (False, True, True))]:
        f64, u64 = res[64][site]
        for L in (32, 16):
            f, U = res[L][site]
            for j in range(len(pts)):
                sf, su = pair_scales(P, cfg, pts[j], qi, *terms, cut=True)
                scale_f = sf + np.abs(f64[j])
                scale_u = su + abs(u64[j])
                assert np.all(np.abs(f[j] - f64[j]) <= 1e-12 * scale_f), (L, site, j, f[j], f64[j])
                assert abs(U[j] - u64[j]) <= 1e-12 * scale_u, (L, site, j, U[j], u64[j])


def _check_forward_mode_derivative(spec, params, P):
    """Training fits the network's directional derivative, carried forward
    through the network beside its value; on the trained network that
    forward-mode derivative must agree with the reverse-mode gradient (the
    backward pass of mccs_mlp_eval) along the same directions: two routes to
    the same analytic derivative, apart only in BF16 rounding."""
    rng = np.random.default_rng(7)
    a = P.atoms()
    k = rng.integers(0, P.info.natoms, 4000)
    pts = np.stack([a["x"][k], a["y"][k], a["z"][k]], 1) + rng.normal(0, 12.0, (4000, 3))
    d = rng.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    _, g = M.mlp_eval(spec, params, pts)
    zero = np.zeros((4000, 2))
    _, _, _, fwd = M.mlp_jvp_loss_grad(spec, params, pts, d, zero, zero, 1.0)
    fwd = fwd * np.array([spec.out_scale[0], spec.out_scale[1]])
    an = np.einsum("pod,pd->po", g, d)
    rel = np.sqrt(np.mean((fwd - an) ** 2)) / np.sqrt(np.mean(an ** 2))
    print("network: forward-mode vs reverse-mode directional derivative, rms %.2e" % rel, flush=True)
    assert rel <= 0.02


def test_what_the_far_field_and_the_network_gain():
    """On the 50,000-atom ellipsoid (net +20 e, N2): the near terms alone (no
    far field) against the far field from the exact evaluator and from the
    trained network (MFMA training on points the linked-cell evaluator
    sampled), each against the exact method on the same 10^5 samples.
    Dropping the far field must shift the CCS significantly (it carries the
    long-range ion-dipole and ion-quadrupole attraction); the exact table
    must not; the network must remove most of the shift; every split run
    must examine far fewer pairs than the exact one."""
    print("case far field gain", flush=True)
    cfg, s, xyz, q = _c4()
    P = M.Prepared(M.config_from_dict(cfg), symbols=list(s), xyz=xyz, q=q)
    E = M.Engine(P)
    u = P.samples(100000)
    rp = E.run(u)
    res = {}
    for name, kw in [("near only", dict(table_source=2)), ("exact table", dict(exact_table=True)),
                     ("network", dict(n_train=1000000, epochs=150))]:
        rep, params, spec = E.build_surrogate(shell=6.0, blend=6.0, grid=0.5, **kw)
        if name == "network":
            _check_forward_mode_derivative(spec, params, P)
        E.use_surrogate(True)
        rs = E.run(u)
        E.use_surrogate(False)
        err, se, cp, cs = _paired(rs, rp)
        res[name] = (err, se)
        print("%s: ccs %.2f vs %.2f, error %.3f %% (SE %.3f %%), table err psi %.2e grad %.2e, "
              "network err psi %.2e grad %.2e, visits ratio %.3f"
              % (name, cs, cp, 100 * err, 100 * se,
                 rep["val_rmse_u"][0] / rep["val_rms_u"][0], rep["val_rmse_f"][0] / rep["val_rms_f"][0],
                 rep["val_rmse_u"][1] / max(rep["val_rms_u"][1], 1e-300),
                 rep["val_rmse_f"][1] / max(rep["val_rms_f"][1], 1e-300),
                 rs["visits"].sum() / rp["visits"].sum()), flush=True)
        assert rs["visits"].sum() < 0.2 * rp["visits"].sum()
    near, se_near = res["near only"]
    assert abs(near) > 4 * se_near  # the far field matters
    assert abs(res["exact table"][0]) <= 0.005 + 2 * res["exact table"][1]
    assert abs(res["network"][0]) < 0.5 * abs(near)
