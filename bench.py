"""Benchmark of the trajectory-method hot path on one MI355X.

Default workload (the configuration the headline metric is quoted on): the
50,000-atom synthetic ellipsoid (semi-axes 60/45/35 A, net +20 e, C:N:O:H:S
= 0.31:0.08:0.10:0.50:0.01, min spacing 1 A) with the N2 probe, LJ + Coulomb
+ ion-induced dipole with cutoffs 12/25 A (force type 6), 10^5 trajectories,
seed 2104, dt 10 fs, 298 K; samples resident in HBM.

One step = the whole Monte-Carlo CCS calculation of the 10^5 trajectories on
the GPU (setup, velocity-Verlet/RATTLE, site forces, tally).  Two legs:

* plain: the exact FP64 engine (mccs_run_device + mccs_tally_device);
* surrogate+batched (the line's `value`): the same target as a batch of one
  (mccs_batch_run_device: the persistent launch replayed from its hipGraph)
  with the range-split far field -- pairs nearer than 6 A exact, the
  remainder of the electrostatic terms beyond them from a cubic B-spline
  table of the far field -- and mccs_tally_device, in the launch shape
  (lanes per trajectory, Verlet skin, LDS tile) that mccs_batch_autotune
  picks from trial runs over the step's samples before the timed steps.

The line carries both rates, their ratio, the CCS error of the surrogate leg
against the plain leg on the same samples with its paired standard error
(on the timed 10^5 and, untimed, on 4 x 10^5 trajectories of the same
stream, which the 0.5 % criterion is judged on),
the roofline of the surrogate leg's trajectory kernel (counted FP64 work
over its HIP-event time) and the CPU baseline: this project's OpenMP C
restatement of the reference path at the reference's release flags, timed
over the reference's CCS-time window on this host's cores (oracle/).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "trajectories/sec on 1 MI355X, plain vs surrogate+batched at ≤0.5% CCS error"
FP64_PEAK_TFLOPS = 78.6
# FP64 operations per unit of counted work (FMA = 2, sqrt = 1, div = 1):
# every (site, atom) pair examined: difference vector, r^2, cutoff test (9);
# LJ term (18); shifted Coulomb (19); smoothed induced dipole field and
# gradient (39); the dipole epilogue per site evaluation (27).  In the
# range-split kernel the Coulomb / dipole counts are the pairs within the
# switch radius and the far-field interpolation is not counted.
W_VISIT, W_LJ, W_COUL, W_DIP, W_EPI = 9, 18, 19, 39, 27
# Range-split far field of the headline leg: exact below 6 A, switched to
# zero at 12 A, far field on a 0.5 A grid from the exact evaluator.
SUR_SHELL, SUR_BLEND, SUR_GRID = 6.0, 6.0, 0.5
# HBM bytes per launch of the range-split trajectory kernel on the headline
# workload (10^5 trajectories of the 50k ellipsoid, the autotuned shape: 64
# lanes, skin 0.5 A, half-CU tile): rocprofv3 --pmc FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0 (KiB per dispatch, scaled by what the counters read for a
# 1 GiB copy: 524,300 KiB fetched, 1,048,576 KiB written);
# profiles/traj_kernel_ellipsoid50k_split.md
SPLIT_TRAFFIC_BYTES = (8718449 * 2.0 + 1230772 * 1.0) * 1024
# Trajectories of the accuracy check that follows the timed steps: the
# timed 10^5 and the next 3 x 10^5 of the same sample stream (untimed), so
# that the paired standard error of the CCS difference (chaotic trajectories
# make it ~0.5 % at 10^5) falls below the 0.5 % criterion
CHECK_TRAJ = 400000


def n2_cfg(path="memory.xyz"):
    return {"targetFileName": path, "GasBuffer": "N2", "Short-range cutoff": "yes",
            "LJ-cutoff": 12.0, "Long-range forces": "yes", "Long-range cutoff": "yes",
            "Coul-cutoff": 25.0, "polarizability": "yes", "alpha": 1.7403, "seed": 2104,
            "dt": 10.0, "Temp": 298.0}


def ellipsoid50k():
    from massccs_amd import synth
    s, xyz, q = synth.globule(n=50000, axes=(60.0, 45.0, 35.0), net_charge=20.0, seed=2104)
    return n2_cfg(), (s, xyz, q), "50,000-atom ellipsoid 60/45/35 A, net +20 e"


def shell1m():
    from massccs_amd import synth
    s, xyz, q = synth.shell()
    return n2_cfg(), (s, xyz, q), "1,000,000-atom hollow shell r 120-140 A, net +60 e"


def globule_cfg(tmp):
    from massccs_amd import synth
    path = os.path.join(tmp, "globule_1231.xyz")
    s, xyz, q = synth.globule(n=1231, axes=(18.0, 15.0, 13.0), net_charge=4.0, seed=2104)
    synth.write_xyz(path, s, xyz, q, comment="synthetic globule 1231 atoms, net +4e")
    return dict(n2_cfg(path), numberProbe=10000, nIter=10)


def small_cfg(tmp, which):
    """The C60 Lennard-Jones case the reference runs on its CPU (He, force type
    2, 3 x 10^4 trajectories) and the published ubiquitin / He configuration
    (force type 4, 10 x 10^4)."""
    from massccs_amd import synth
    common = {"seed": 2104, "dt": 10.0, "Temp": 298.0}
    if which == "c60":
        xyz = os.path.join(tmp, "c60_analytic.xyz")
        synth.write_xyz(xyz, ["C"] * 60, synth.c60(), comment="C60 truncated icosahedron")
        ff = os.path.join(tmp, "c60_he.ff")
        with open(ff, "w") as fh:
            fh.write("1\nSymbol mass eps sigma\nC 12.011 0.041473 3.02\n")
        return dict(common, targetFileName=xyz, **{"force-field": ff}, GasBuffer="He",
                    **{"Short-range cutoff": "yes", "Long-range forces": "no"},
                    polarizability="no", numberProbe=10000, nIter=3), \
            "plain FP64, C60 (analytic, LJ only) + He, force type 2, 3x10^4 trajectories"
    pqr = os.path.join(HERE, "tests", "golden", "04_1ubq_charge_4e_min.pqr")
    return dict(common, targetFileName=pqr, GasBuffer="He", polarizability="yes",
                **{"Short-range cutoff": "yes", "LJ-cutoff": 12.0, "Long-range forces": "yes",
                   "Long-range cutoff": "yes", "Coul-cutoff": 25.0},
                numberProbe=10000, nIter=10), \
        "plain FP64, ubiquitin 04_1ubq_charge_4e_min.pqr + He, force type 4, 10x10^4 trajectories"


def cpu_baseline(P, cfg, budget_s):
    """oracle/ port at the reference's release flags (never the product)."""
    try:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import cpu_baseline as CB
        return CB.baseline(P, cfg, budget_s=budget_s, repeats=5)
    except Exception as e:  # the baseline is reported, never required
        return {"value": None, "unit": "trajectories/s", "kind": "port", "sample": "failed: %s" % e}


def flops_of(r):
    """Counted FP64 work of a set of records (the trajectories' own work, the
    tries whose outcome they report)."""
    hits = r["hit"] == 1
    evals = (1 + r["steps"] + np.maximum(r["tries"], 0)) * hits
    return float(W_VISIT * r["visits"].sum() + W_LJ * r["lj"].sum() + W_COUL * r["coul"].sum()
                 + W_DIP * r["dip"].sum() + W_EPI * evals.sum())


def paired_error(rs, rp):
    """CCS of records rs relative to rp on the same samples (one iteration of
    all trajectories), and the standard error of that difference from the
    per-trajectory differences (chaotic trajectories make it non-zero even
    for rounding-level changes of the force)."""
    den = float(((rp["hit"] == 0) | (rp["success"] == 1)).sum())
    dens = float(((rs["hit"] == 0) | (rs["success"] == 1)).sum())
    cp = rp["domega"].sum() / den
    cs = rs["domega"].sum() / dens
    se = float(np.std(rs["domega"] - rp["domega"]) * np.sqrt(len(rp)) / den)
    return cs / cp - 1.0, se / cp, cp, cs


class Timer:
    """HIP events on the stream the kernels are launched on."""

    def __init__(self, stream):
        import torch
        self.torch, self.stream, self.ev = torch, stream, []

    def __enter__(self):
        self.e0 = self.torch.cuda.Event(enable_timing=True)
        self.e1 = self.torch.cuda.Event(enable_timing=True)
        self.e0.record(self.stream)
        return self

    def __exit__(self, *a):
        self.e1.record(self.stream)
        self.ev.append((self.e0, self.e1))

    def mean_ms(self):
        return float(np.mean([a.elapsed_time(b) for a, b in self.ev]))

    def total_ms(self):
        return float(np.sum([a.elapsed_time(b) for a, b in self.ev]))


def run_headline(args, target):
    """Plain vs surrogate+batched on one large N2 target (samples in HBM)."""
    import torch
    import massccs_amd as M
    cfg, (s, xyz, q), label = target()
    n = args.traj or 100000
    # a step may run as several launches of `chunk` trajectories (10^6-trajectory
    # runs of the large shell), each followed by a progress line on stderr
    chunk = args.chunk or n
    if n % chunk:
        sys.exit("--chunk must divide the number of trajectories")
    starts = list(range(0, n, chunk))
    RS = M.RECORD_DTYPE.itemsize
    P = M.Prepared(M.config_from_dict(cfg), symbols=list(s), xyz=xyz, q=q)
    stream = torch.cuda.Stream()  # default priority; every launch and event goes on it
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream
    u = torch.from_numpy(P.samples(n)).cuda()
    rec_p = torch.zeros((n, M.RECORD_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    rec_s = torch.zeros_like(rec_p)
    om = torch.zeros(1, dtype=torch.float64, device="cuda")
    cnt = torch.zeros(3, dtype=torch.int64, device="cuda")
    me = torch.zeros(2, dtype=torch.float64, device="cuda")

    # plain leg
    E = M.Engine(P, count_work=True)

    def launches(run, rec, timer, leg, first_only=False):
        for k, a in enumerate(starts[:1] if first_only else starts):
            if timer:
                with timer:
                    run(u.data_ptr() + 72 * a, rec.data_ptr() + RS * a)
            else:
                run(u.data_ptr() + 72 * a, rec.data_ptr() + RS * a)
            if len(starts) > 1:
                torch.cuda.synchronize()
                print("%s leg: launch %d of %d done" % (leg, k + 1, len(starts)), file=sys.stderr,
                      flush=True)

    def plain_step(timer=None, first_only=False):
        launches(lambda up, rp: E.run_device(up, chunk, rp, sptr), rec_p, timer, "plain", first_only)
        E.tally_device(rec_p.data_ptr(), 1, n, om.data_ptr(), cnt.data_ptr(), me.data_ptr(), sptr)

    for _ in range(args.warmup):
        plain_step()
    torch.cuda.synchronize()
    tp = Timer(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plain_step(tp)
    torch.cuda.synchronize()
    wall_p = time.perf_counter() - t0
    rp = rec_p.cpu().numpy().reshape(-1).view(M.RECORD_DTYPE)
    flops_p = flops_of(rp)

    # surrogate+batched leg: the target as a batch of one, range-split far
    # field; its launch shape (lanes per trajectory, Verlet skin, LDS tile)
    # picked by the batch autotuner from trial runs over one step's samples
    # (untimed), unless --sur-lanes fixes the lanes (skin 1 A)
    tuned = args.sur_lanes == 0
    B = M.Batch([P], lanes=args.sur_lanes, count_work=True, use_graph=True,
                skin=0.0 if tuned else 1.0)
    t0 = time.perf_counter()
    rep = B.build_surrogate(shell=SUR_SHELL, blend=SUR_BLEND, grid=SUR_GRID, exact_table=True)[0]
    build_s = time.perf_counter() - t0
    B.use_surrogate(True)
    shape = {"lanes": args.sur_lanes, "skin": 1.0, "tile": None, "autotune_s": None}
    if tuned:
        t0 = time.perf_counter()
        tune = B.autotune([P.samples(chunk)])[0]
        shape = {"lanes": tune["lanes"], "skin": tune["skin"], "tile": tune["tile"],
                 "autotune_s": round(time.perf_counter() - t0, 2), "trials": len(tune["trials"])}
        print("autotuned range-split shape: %s" % shape, file=sys.stderr, flush=True)

    def sur_step(timer=None, first_only=False):
        launches(lambda up, rp: B.run_device([chunk], [up], [rp], sptr), rec_s, timer, "surrogate",
                 first_only)
        E.tally_device(rec_s.data_ptr(), 1, n, om.data_ptr(), cnt.data_ptr(), me.data_ptr(), sptr)

    for w in range(max(1, args.warmup)):  # the first launch captures the graph
        sur_step(first_only=args.warmup == 0)
    torch.cuda.synchronize()
    ts = Timer(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sur_step(ts)
    torch.cuda.synchronize()
    wall_s = time.perf_counter() - t0
    ccs_s = me.cpu().numpy()
    rs = rec_s.cpu().numpy().reshape(-1).view(M.RECORD_DTYPE)
    flops_s = flops_of(rs)
    err, se, cp, cs = paired_error(rs, rp)
    # kernel time per step (the sum of its launches) and per launch
    kern_s = ts.total_ms() / args.steps
    kern_p = tp.total_ms() / args.steps
    # accuracy check over CHECK_TRAJ trajectories (untimed, same engines);
    # a larger timed run is its own check
    nchk = n if n >= CHECK_TRAJ or args.traj else CHECK_TRAJ
    if nchk == n:
        rcp, rcs = rp, rs
    else:
        uc = torch.from_numpy(P.samples(nchk)).cuda()
        rc_p = torch.zeros((nchk, RS), dtype=torch.uint8, device="cuda")
        rc_s = torch.zeros_like(rc_p)
        E.run_device(uc.data_ptr(), nchk, rc_p.data_ptr(), sptr)
        B.run_device([nchk], [uc.data_ptr()], [rc_s.data_ptr()], sptr)
        torch.cuda.synchronize()
        rcp = rc_p.cpu().numpy().reshape(-1).view(M.RECORD_DTYPE)
        rcs = rc_s.cpu().numpy().reshape(-1).view(M.RECORD_DTYPE)
        del uc, rc_p, rc_s
    err4, se4, cp4, cs4 = paired_error(rcs, rcp)
    value = n * args.steps / wall_s
    plain_value = n * args.steps / wall_p
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "trajectories/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall_s / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded target and mt19937 samples)",
        "config": {"workload": "surrogate+batched vs plain FP64, %s + N2, LJ+Coulomb+induced "
                               "dipole with cutoffs (force type 6), %d trajectories" % (label, n),
                   "n_atoms": P.info.natoms, "trajectories_per_step": n, "probe": "N2",
                   "force_type": P.info.force_type, "parallelism": "1 GPU"},
        "plain": {"value": round(plain_value, 1), "ms_per_step": round(1e3 * wall_p / args.steps, 3),
                  "kernel_ms": round(kern_p, 3), "launches_per_step": len(starts), "ccs": cp,
                  "force_evals_per_s": round(3 * int(rp["steps"].sum()) / (kern_p * 1e-3), 1),
                  "roofline_frac": round(flops_p / (kern_p * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4)},
        "speedup_vs_plain": round(value / plain_value, 3),
        "ccs": {"surrogate": cs, "plain": cp, "error_vs_plain": round(err, 5),
                "paired_standard_error": round(se, 5), "device_tally_mean": float(ccs_s[0]),
                "lost": {"surrogate": int(((rs["hit"] == 1) & (rs["success"] == 0)).sum()),
                         "plain": int(((rp["hit"] == 1) & (rp["success"] == 0)).sum())},
                "check": {"trajectories": nchk, "surrogate": cs4, "plain": cp4,
                          "error_vs_plain": round(err4, 5), "paired_standard_error": round(se4, 5)},
                "within_0.5%": bool(abs(err4) <= 0.005)},
        "surrogate": {"shell_A": SUR_SHELL, "switch_end_A": SUR_SHELL + SUR_BLEND,
                      "lanes_per_trajectory": shape["lanes"], "skin_A": shape["skin"],
                      "lds_tile": shape["tile"], "autotune_s": shape["autotune_s"],
                      "grid_A": SUR_GRID, "table": "exact far field at the nodes",
                      "grid_nodes": rep["grid_nodes"], "build_s": round(build_s, 2),
                      "table_rel_err_psi": rep["val_rmse_u"][0] / max(rep["val_rms_u"][0], 1e-300),
                      "table_rel_err_grad": rep["val_rmse_f"][0] / max(rep["val_rms_f"][0], 1e-300),
                      "visits_ratio": round(float(rs["visits"].sum()) / max(1.0, float(rp["visits"].sum())), 4),
                      "note": "value excludes the one-off table build (build_s) and, in both "
                              "legs, the host generation of the samples (the reference's CCS "
                              "time includes it: about 1 ms per 10^5 samples here, resident "
                              "in HBM before the timed steps)"},
        "force_evals_per_s": round(3 * int(rs["steps"].sum()) / (kern_s * 1e-3), 1),
        "roofline": {"bound": "fp64-valu", "achieved": round(flops_s / (kern_s * 1e-3) / 1e12, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(flops_s / (kern_s * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
                     "traffic": int(SPLIT_TRAFFIC_BYTES) if P.info.natoms == 50000 and chunk == 100000 else None,
                     "kernel": "traj_kernel<N2, type 6, range-split, batched>",
                     "kernel_ms": round(kern_s / len(starts), 3), "launches_per_step": len(starts),
                     "flops_per_launch": int(flops_s / len(starts))},
    }
    if args.network:
        E2 = M.Engine(P, count_work=True)
        t0 = time.perf_counter()
        rep2 = E2.build_surrogate(shell=SUR_SHELL, blend=SUR_BLEND, grid=SUR_GRID)[0]
        b2 = time.perf_counter() - t0
        E2.use_surrogate(True)
        uc = torch.from_numpy(P.samples(nchk)).cuda()
        rn_d = torch.zeros((nchk, M.RECORD_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
        E2.run_device(uc.data_ptr(), nchk, rn_d.data_ptr(), sptr)
        torch.cuda.synchronize()
        rn = rn_d.cpu().numpy().reshape(-1).view(M.RECORD_DTYPE)
        e2, se2, _, cn = paired_error(rn, rcp)
        out["network_table"] = {
            "trajectories": nchk, "ccs_error_vs_plain": round(e2, 5), "paired_standard_error": round(se2, 5),
            "build_s": round(b2, 2), "train_s": round(rep2["seconds_train"], 2),
            "net_rel_err_psi": rep2["val_rmse_u"][1] / max(rep2["val_rms_u"][1], 1e-300),
            "net_rel_err_grad": rep2["val_rmse_f"][1] / max(rep2["val_rms_f"][1], 1e-300)}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(P, cfg, budget_s=args.cpu_budget)
    return out


def run_single(args):
    """Plain engine on the globule / ubiquitin / C60 workloads (secondary)."""
    import torch
    import massccs_amd as M
    tmp = tempfile.mkdtemp(prefix="mccs_bench_")
    if args.workload == "globule":
        cfg = globule_cfg(tmp)
        label = ("plain FP64, 1231-atom globule + N2, LJ+Coulomb+induced dipole with cutoffs "
                 "(force type 6), 10x10^4 trajectories")
    else:
        cfg, label = small_cfg(tmp, args.workload)
    P = M.Prepared(M.config_from_dict(cfg))
    n_iter, n_traj = cfg["nIter"], cfg["numberProbe"]
    n = n_iter * n_traj
    E = M.Engine(P, device=0, lanes=args.lanes, count_work=True)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream
    u = torch.from_numpy(P.samples(n)).cuda()
    rec = torch.zeros((n, M.RECORD_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    omega = torch.zeros(n_iter, dtype=torch.float64, device="cuda")
    counts = torch.zeros(3 * n_iter, dtype=torch.int64, device="cuda")
    mean_err = torch.zeros(2, dtype=torch.float64, device="cuda")
    tm = Timer(stream)

    def step(timer=None):
        if timer:
            with timer:
                E.run_device(u.data_ptr(), n, rec.data_ptr(), sptr)
        else:
            E.run_device(u.data_ptr(), n, rec.data_ptr(), sptr)
        E.tally_device(rec.data_ptr(), n_iter, n_traj, omega.data_ptr(), counts.data_ptr(),
                       mean_err.data_ptr(), sptr)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(tm)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = tm.mean_ms()
    r = rec.cpu().numpy().reshape(-1).view(M.RECORD_DTYPE)
    me = mean_err.cpu().numpy()
    flops = flops_of(r)
    nsite = 3 if cfg["GasBuffer"] in ("N2", "CO2") else 1
    out = {
        "metric": METRIC, "value": round(n * args.steps / wall, 1), "unit": "trajectories/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded target and mt19937 samples)",
        "config": {"workload": label, "n_atoms": P.info.natoms, "trajectories_per_step": n,
                   "probe": cfg["GasBuffer"], "force_type": P.info.force_type,
                   "lanes_per_trajectory": args.lanes or "auto", "parallelism": "1 GPU"},
        "force_evals_per_s": round(nsite * int(r["steps"].sum()) / (kern_ms * 1e-3), 1),
        "ccs": {"mean": float(me[0]), "err": float(me[1])},
        "roofline": {"bound": "fp64-valu", "achieved": round(flops / (kern_ms * 1e-3) / 1e12, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(flops / (kern_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
                     "traffic": None,
                     "kernel": "traj_kernel<%s, type %d>" % (cfg["GasBuffer"], P.info.force_type),
                     "kernel_ms": round(kern_ms, 3), "flops_per_launch": int(flops)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(P, cfg, budget_s=args.cpu_budget)
    return out


def run_batch(args):
    """The 64-assembly batch: one persistent launch per probe kind, autotuned
    launch shapes, replayed from a hipGraph; with --surrogate the members'
    range-split far fields."""
    import massccs_amd as M
    from massccs_amd import synth
    descs = synth.assemblies()
    per = args.traj or 20000
    P, us = [], []
    for d in descs:
        s, xyz, q = synth.globule(n=d["n"], axes=d["axes"], net_charge=d["net_charge"],
                                  seed=d["seed"])
        cfg = {"targetFileName": "memory.xyz", "GasBuffer": d["probe"], "polarizability": "yes",
               "Short-range cutoff": "yes", "Long-range forces": "yes",
               "Long-range cutoff": "yes", "seed": d["seed"]}
        if d["probe"] == "N2":
            cfg["alpha"] = 1.7403
        p = M.Prepared(M.config_from_dict(cfg), symbols=list(s), xyz=xyz, q=q)
        P.append(p)
        us.append(p.samples(per))
    B = M.Batch(P)
    build = None
    if args.surrogate:
        t0 = time.perf_counter()
        B.build_surrogate(shell=SUR_SHELL, blend=SUR_BLEND, grid=SUR_GRID, exact_table=True)
        build = time.perf_counter() - t0
        B.use_surrogate(True)
    t0 = time.perf_counter()
    tune = B.autotune([u[: max(64, per // 16)] for u in us])
    t_tune = time.perf_counter() - t0
    for _ in range(args.warmup):
        B.run(us)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        recs = B.run(us)
    wall = time.perf_counter() - t0
    n = per * len(P)
    r = np.concatenate(recs)
    return {
        "metric": METRIC, "value": round(n * args.steps / wall, 1), "unit": "trajectories/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (64 seeded assemblies and mt19937 samples)",
        "config": {"workload": "batched FP64%s, 64 assemblies of 514-199,788 atoms (He type 4 / "
                               "N2 type 6 alternating), %d trajectories each"
                               % (" with range-split far fields" if args.surrogate else "", per),
                   "targets": len(P), "atoms_total": int(sum(p.info.natoms for p in P)),
                   "trajectories_per_step": n, "parallelism": "1 GPU"},
        "steps_per_trajectory": {"min": int(r["steps"][r["hit"] == 1].min()),
                                 "max": int(r["steps"].max()),
                                 "mean": float(r["steps"][r["hit"] == 1].mean())},
        "autotune": {"seconds": round(t_tune, 2),
                     "choices": [{"class": c["class"], "lanes": c["lanes"], "skin": c["skin"],
                                  "tile": c["tile"]} for c in tune]},
        "surrogate_build_s": build, "flops": flops_of(r), "roofline": None, "cpu_baseline": None,
    }


# BF16 MFMA products per network row (64 features -> 3 x 128 SiLU): forward
# 64x128 + 2 x 128x128; backward data products 2 x 128x128 (the layer-1
# input gradient is not needed in training); weight gradients 64x128 + 2 x
# 128x128; 2 flops per multiply-add.
MLP_FWD_FLOPS = 2 * (64 * 128 + 2 * 128 * 128)
MLP_TRAIN_FLOPS_PER_ROW = MLP_FWD_FLOPS + 2 * 2 * 128 * 128 + MLP_FWD_FLOPS
BF16_PEAK_TFLOPS = 2500.0


def run_network(args):
    """The learned far field of the 50k ellipsoid: 2 x 10^6 points sampled by
    the linked-cell evaluator, trained on the GPU (value and central-difference
    directional derivative: three network rows per point, batch 8192) and
    tabulated at every grid node; reports the BF16 MFMA rate of training."""
    import massccs_amd as M
    cfg, (s, xyz, q), label = ellipsoid50k()
    P = M.Prepared(M.config_from_dict(cfg), symbols=list(s), xyz=xyz, q=q)
    E = M.Engine(P, count_work=True)
    epochs = args.epochs
    n_train = 2_000_000
    t0 = time.perf_counter()
    rep = E.build_surrogate(n_train=n_train, epochs=epochs, shell=SUR_SHELL, blend=SUR_BLEND,
                            grid=SUR_GRID)[0]
    build_s = time.perf_counter() - t0
    rows = 3 * n_train * epochs
    train_flops = rows * MLP_TRAIN_FLOPS_PER_ROW
    table_flops = rep["grid_nodes"] * MLP_FWD_FLOPS
    t_train = rep["seconds_train"]
    rate = train_flops / t_train / 1e12
    return {
        "metric": "far-field network training on 1 MI355X (BF16 MFMA, FP32 accumulation)",
        "value": round(rows / t_train, 1), "unit": "network rows/s (forward + backward + Adam)",
        "n_gpus": 1, "steps": epochs, "warmup": 0, "ms_per_step": round(1e3 * t_train / epochs, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic (exact far field of the seeded target at sampled points)",
        "config": {"workload": "network far field, %s + N2: %d points x %d epochs, batch 8192, "
                               "3 rows per point" % (label, n_train, epochs),
                   "n_atoms": P.info.natoms, "grid_nodes": rep["grid_nodes"], "parallelism": "1 GPU"},
        "seconds": {"sample": round(rep["seconds_sample"], 3), "train": round(t_train, 3),
                    "table": round(rep["seconds_table"], 3), "build": round(build_s, 3)},
        "roofline": {"bound": "mfma", "achieved": round(rate, 2), "peak": BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(rate / BF16_PEAK_TFLOPS, 5), "traffic": None,
                     "kernel": "whole training (every kernel of the step in the time)",
                     "flops_train": int(train_flops), "flops_table": int(table_flops)},
        "network_rel_err": {"psi": rep["val_rmse_u"][1] / max(rep["val_rms_u"][1], 1e-300),
                            "grad": rep["val_rmse_f"][1] / max(rep["val_rms_f"][1], 1e-300)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lanes", type=int, default=0, help="lanes per trajectory: 0 auto, 16, 32 or 64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=8.0,
                    help="seconds of CPU work per baseline run (5 runs, median)")
    ap.add_argument("--workload", default="ellipsoid50k",
                    choices=["ellipsoid50k", "shell1m", "globule", "batch64", "c60", "ubq_he",
                             "network"],
                    help="ellipsoid50k: the headline configuration (plain vs surrogate+batched); "
                         "the others are secondary measurements")
    ap.add_argument("--traj", type=int, default=0, help="trajectories (per target for batch64)")
    ap.add_argument("--chunk", type=int, default=0,
                    help="headline workloads: trajectories per launch (must divide --traj; 0 = all)")
    ap.add_argument("--network", action="store_true",
                    help="headline workloads: also build the far field from the trained network "
                         "and report its CCS error")
    ap.add_argument("--surrogate", action="store_true", help="batch64: range-split far fields")
    ap.add_argument("--epochs", type=int, default=300, help="network: training epochs")
    ap.add_argument("--sur-lanes", type=int, default=0, choices=[0, 16, 32, 64],
                    help="headline workloads: lanes per trajectory of the range-split leg "
                         "(0: the batch autotuner picks lanes, skin and LDS tile)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy a 1 GiB float32 tensor three times: the calibration of "
                         "the FETCH_SIZE / WRITE_SIZE counters for the traffic figure")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 << 28, "copies": 3}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the trajectory engine is measured on a single "
              "MI355X; trajectories would shard by index range across GPUs, which is not run here",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch64":
        out = run_batch(args)
    elif args.workload == "network":
        out = run_network(args)
    elif args.workload in ("ellipsoid50k", "shell1m"):
        out = run_headline(args, ellipsoid50k if args.workload == "ellipsoid50k" else shell1m)
    else:
        out = run_single(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
