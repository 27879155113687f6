"""Benchmark of the trajectory-method hot path on one MI355X.

Workload: the 1,231-atom synthetic globule (ellipsoid 18/15/13 A, min spacing
1 A, C:N:O:H:S = 0.31:0.08:0.10:0.50:0.01, net +4 e) with the N2 probe,
LJ + Coulomb + ion-induced dipole with cutoffs 12/25 A (force type 6),
10 x 10^4 trajectories, seed 2104, dt 10 fs, 298 K.

One step = the whole Monte-Carlo CCS calculation on the GPU: every
trajectory of the 10^5 samples (setup, velocity-Verlet/RATTLE, site forces)
plus the tally, with the samples already resident in HBM.

Printed: one JSON line with trajectories/s, the roofline of the trajectory
kernel (FP64 work counted by the kernel itself, over its HIP-event duration)
and the reference CPU program timed on this host's cores.
"""
import argparse
import json
import os
import re
import shutil
import subprocess
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
# gradient (39); the dipole epilogue per site evaluation (27).
W_VISIT, W_LJ, W_COUL, W_DIP, W_EPI = 9, 18, 19, 39, 27
# HBM bytes per launch of this workload's trajectory kernel, from two
# rocprofv3 --pmc runs: FETCH_SIZE 240,881 KiB x 2.0 + WRITE_SIZE 6,716,231 KiB
# x 1.0, each scaled by what the counter reads for a 1 GiB copy
# (bench.py --copy-probe); 1.36 GB at 2 waves per SIMD, the rest comes with
# the 3-wave launch shape (spill scratch, more resident groups); not measured
# by this script.
TRAFFIC_BYTES = (240881 * 2 + 6716231) * 1024


def workload_cfg(tmp):
    from massccs_amd import synth
    path = os.path.join(tmp, "globule_1231.xyz")
    s, xyz, q = synth.globule(n=1231, axes=(18.0, 15.0, 13.0), net_charge=4.0, seed=2104)
    synth.write_xyz(path, s, xyz, q, comment="synthetic globule 1231 atoms, net +4e")
    return {"targetFileName": path, "GasBuffer": "N2", "Short-range cutoff": "yes",
            "LJ-cutoff": 12.0, "Long-range forces": "yes", "Long-range cutoff": "yes",
            "Coul-cutoff": 25.0, "polarizability": "yes", "alpha": 1.7403, "seed": 2104,
            "dt": 10.0, "Temp": 298.0, "numberProbe": 10000, "nIter": 10}


def small_cfg(tmp, which):
    """Other single-target workloads: the C60 Lennard-Jones case the reference
    runs on its CPU (He, force type 2, 3 x 10^4 trajectories) and the
    published ubiquitin / He configuration (force type 4, 10 x 10^4)."""
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


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def cpu_flags():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("flags"):
                    return set(ln.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def cpu_baseline(cfg, budget_s=15.0):
    """The reference program (oracle/_ref, built from the reference sources with
    its release flags) on this host's cores, on a bounded sample of the same
    workload: trajectories/s from its own "CCS time" stamp."""
    ref_dir = os.path.join(HERE, "oracle", "_ref")
    v4 = {"avx512f", "avx512bw", "avx512dq", "avx512vl", "avx512cd"} <= cpu_flags()
    exe = os.path.join(ref_dir, "massccs_ref_v4" if v4 else "massccs_ref")
    if not os.path.exists(exe):
        return None
    cores = host_cores()
    tmp = tempfile.mkdtemp(prefix="mccs_cpu_")
    try:
        shutil.copy(cfg["targetFileName"], tmp)

        def run(n):
            c = dict(cfg, targetFileName=os.path.basename(cfg["targetFileName"]),
                     numberProbe=int(n), nIter=1, nthreads=cores)
            with open(os.path.join(tmp, "in.json"), "w") as fh:
                json.dump(c, fh)
            out = subprocess.run([exe, "in.json"], cwd=tmp, capture_output=True, text=True,
                                 timeout=600).stdout
            t = float(re.search(r"CCS time: ([0-9.eE+-]+) s", out).group(1))
            return t
        n0 = 64 * cores
        t0 = run(n0)
        n1 = int(min(100000, max(n0, n0 * budget_s / max(t0, 1e-3))))
        t1 = run(n1)
        return {"value": n1 / t1, "unit": "trajectories/s", "cores": cores, "kind": "reference",
                "sample": "%d trajectories (first %d of the same seeded stream, nIter 1), "
                          "reference CLI built %s -O3 -ffast-math -fopenmp, %d OpenMP threads, "
                          "CCS time %.2f s" % (n1, n1, "-march=x86-64-v4" if v4 else
                                               "-march=x86-64-v3", cores, t1)}
    except Exception as e:  # the baseline is reported, never required
        return {"value": None, "unit": "trajectories/s", "cores": cores, "kind": "reference",
                "sample": "failed: %s" % e}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def big_target_cfg(which):
    """In-memory synthetic targets of the large configurations (N2, type 6)."""
    from massccs_amd import synth
    if which == "shell1m":
        s, xyz, q = synth.shell()
        label = "1,000,000-atom hollow shell r 120-140 A, net +60 e"
    else:
        s, xyz, q = synth.globule(n=50000, axes=(60.0, 45.0, 35.0), net_charge=20.0, seed=2104)
        label = "50,000-atom ellipsoid 60/45/35 A, net +20 e"
    cfg = {"targetFileName": "memory.xyz", "GasBuffer": "N2", "Short-range cutoff": "yes",
           "LJ-cutoff": 12.0, "Long-range forces": "yes", "Long-range cutoff": "yes",
           "Coul-cutoff": 25.0, "polarizability": "yes", "alpha": 1.7403, "seed": 2104,
           "dt": 10.0, "Temp": 298.0}
    return cfg, (s, xyz, q), label


def run_big(args, which):
    """Plain FP64 engine on one large target (samples resident in HBM)."""
    import torch
    import massccs_amd as M
    cfg, (s, xyz, q), label = big_target_cfg(which)
    n = args.traj or (10 ** 6 if which == "shell1m" else 10 ** 5)
    P = M.Prepared(M.config_from_dict(cfg), symbols=list(s), xyz=xyz, q=q)
    E = M.Engine(P, count_work=True, lanes=args.lanes)
    u = torch.from_numpy(P.samples(n)).cuda()
    rec = torch.zeros((n, M.RECORD_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    for _ in range(args.warmup):
        E.run_device(u.data_ptr(), min(n, 20000), rec.data_ptr(), stream.cuda_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    k_ev = []
    for _ in range(args.steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        E.run_device(u.data_ptr(), n, rec.data_ptr(), stream.cuda_stream)
        e1.record(stream)
        k_ev.append((e0, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    r = rec.cpu().numpy().reshape(-1).view(M.RECORD_DTYPE)
    om, cnt, mean, err = M.tally(r, 1, n)
    flops = _flops(r)
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in k_ev]))  # mean over the timed launches
    sur = None
    if args.surrogate:
        t0 = time.perf_counter()
        rep, _, _ = E.build_surrogate()
        t_build = time.perf_counter() - t0
        E.use_surrogate(True)
        E.run_device(u.data_ptr(), min(n, 20000), rec.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            E.run_device(u.data_ptr(), n, rec.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        wall_s = time.perf_counter() - t0
        E.use_surrogate(False)
        rs = rec.cpu().numpy().reshape(-1).view(M.RECORD_DTYPE)
        _, _, mean_s, _ = M.tally(rs, 1, n)
        sur = {"value": round(n * args.steps / wall_s, 1), "unit": "trajectories/s",
               "speedup_vs_plain": round(wall / wall_s, 3),
               "ccs": mean_s, "ccs_error_vs_plain": round(mean_s / mean - 1, 5),
               "lost": int(((rs["hit"] == 1) & (rs["success"] == 0)).sum()),
               "build_seconds": round(t_build, 2),
               "visits_ratio": round(float(rs["visits"].sum()) / max(1.0, float(r["visits"].sum())), 4),
               "validation": {"u_rel": [rep["val_rmse_u"][k] / rep["val_rms_u"][k] for k in range(2)],
                              "f_rel": [rep["val_rmse_f"][k] / rep["val_rms_f"][k] for k in range(2)]},
               "grid_nodes": rep["grid_nodes"]}
    return {
        "metric": METRIC, "value": round(n * args.steps / wall, 1), "unit": "trajectories/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded target and mt19937 samples)",
        "config": {"workload": "plain FP64, %s + N2, force type 6, %d trajectories" % (label, n),
                   "n_atoms": P.info.natoms, "trajectories_per_step": n, "probe": "N2",
                   "force_type": P.info.force_type, "parallelism": "1 GPU"},
        "force_evals_per_s": round(3 * int(r["steps"].sum()) / (kern_ms * 1e-3), 1),
        "ccs": {"mean": mean, "lost": int(((r["hit"] == 1) & (r["success"] == 0)).sum())},
        "roofline": {"bound": "mfma", "achieved": round(flops / (kern_ms * 1e-3) / 1e12, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(flops / (kern_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
                     "traffic": None, "kernel": "traj_kernel<N2, type 6>",
                     "kernel_ms": round(kern_ms, 3), "flops_per_launch": int(flops)},
        "cpu_baseline": None,
        "surrogate": sur,
    }


def _flops(r):
    hits = r["hit"] == 1
    evals = (1 + r["steps"] + np.maximum(r["tries"], 0)) * hits
    return float(W_VISIT * r["visits"].sum() + W_LJ * r["lj"].sum() + W_COUL * r["coul"].sum()
                 + W_DIP * r["dip"].sum() + W_EPI * evals.sum())


def run_batch(args):
    """The 64-assembly batch: one persistent launch per probe kind, autotuned
    launch shapes, replayed from a hipGraph."""
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
    steps = np.array([int(x["steps"].sum()) for x in recs])
    return {
        "metric": METRIC, "value": round(n * args.steps / wall, 1), "unit": "trajectories/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (64 seeded assemblies and mt19937 samples)",
        "config": {"workload": "batched FP64, 64 assemblies of 514-199,788 atoms (He type 4 / "
                               "N2 type 6 alternating), %d trajectories each" % per,
                   "targets": len(P), "atoms_total": int(sum(p.info.natoms for p in P)),
                   "trajectories_per_step": n, "parallelism": "1 GPU"},
        "force_evals_per_s": None,
        "steps_per_trajectory": {"min": int(r["steps"][r["hit"] == 1].min()),
                                 "max": int(r["steps"].max()),
                                 "mean": float(r["steps"][r["hit"] == 1].mean())},
        "autotune": {"seconds": round(t_tune, 2),
                     "choices": [{"class": c["class"].split("/")[0] + "/" + c["class"].split("/")[1],
                                  "lanes": c["lanes"], "skin": c["skin"]} for c in tune]},
        "flops": _flops(r), "roofline": None, "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lanes", type=int, default=0, help="lanes per trajectory: 0 auto, 16, 32 or 64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="globule",
                    choices=["globule", "batch64", "ellipsoid50k", "shell1m", "c60", "ubq_he"],
                    help="globule: the measured configuration (1,231 atoms, N2, 10^5 "
                         "trajectories); the others are the batch and large-target runs")
    ap.add_argument("--traj", type=int, default=0, help="trajectories (per target for batch64)")
    ap.add_argument("--surrogate", action="store_true",
                    help="ellipsoid50k/shell1m: also build the far-field network and time the "
                         "surrogate path against the plain kernel on the same samples")
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
              "MI355X; trajectories would shard by index range across GPUs, which is not run here", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch64":
        print(json.dumps(run_batch(args)))
        return
    if args.workload in ("ellipsoid50k", "shell1m"):
        print(json.dumps(run_big(args, args.workload)))
        return

    import torch
    import massccs_amd as M

    tmp = tempfile.mkdtemp(prefix="mccs_bench_")
    if args.workload == "globule":
        cfg = workload_cfg(tmp)
        label = ("plain FP64, 1231-atom globule + N2, LJ+Coulomb+induced dipole with cutoffs "
                 "(force type 6), 10x10^4 trajectories")
    else:
        cfg, label = small_cfg(tmp, args.workload)
    P = M.Prepared(M.config_from_dict(cfg))
    info = P.info
    n_iter, n_traj = cfg["nIter"], cfg["numberProbe"]
    n = n_iter * n_traj
    E = M.Engine(P, device=0, lanes=args.lanes, count_work=True)

    u = torch.from_numpy(P.samples(n)).cuda()
    rec = torch.zeros((n, M.RECORD_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    omega = torch.zeros(n_iter, dtype=torch.float64, device="cuda")
    counts = torch.zeros(3 * n_iter, dtype=torch.int64, device="cuda")
    mean_err = torch.zeros(2, dtype=torch.float64, device="cuda")
    # a stream of our own (default priority): the kernels and the HIP events
    # that time them are enqueued on it
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream

    def step():
        E.run_device(u.data_ptr(), n, rec.data_ptr(), sptr)
        E.tally_device(rec.data_ptr(), n_iter, n_traj, omega.data_ptr(), counts.data_ptr(),
                       mean_err.data_ptr(), sptr)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    k_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        E.run_device(u.data_ptr(), n, rec.data_ptr(), sptr)
        e1.record(stream)
        E.tally_device(rec.data_ptr(), n_iter, n_traj, omega.data_ptr(), counts.data_ptr(),
                       mean_err.data_ptr(), sptr)
        k_ms.append((e0, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in k_ms]))

    executed = E.last_work()  # every try run, including side-by-side tries left unused
    r = rec.cpu().numpy().reshape(-1).view(M.RECORD_DTYPE)
    me = mean_err.cpu().numpy()
    flops = _flops(r)
    achieved = flops / (kern_ms * 1e-3) / 1e12
    value = n * args.steps / wall
    nsite = 3 if cfg["GasBuffer"] in ("N2", "CO2") else 1
    site_evals = nsite * int(r["steps"].sum())
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "trajectories/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded globule target and mt19937 samples)",
        "config": {"workload": label,
                   "n_atoms": info.natoms, "trajectories_per_step": n, "probe": cfg["GasBuffer"],
                   "force_type": info.force_type, "lanes_per_trajectory": args.lanes or "auto",
                   "parallelism": "1 GPU"},
        "force_evals_per_s": round(site_evals / (kern_ms * 1e-3), 1),
        "ccs": {"mean": float(me[0]), "err": float(me[1])},
        "steps_per_launch": {"of_records": int(r["steps"].sum()),
                             "executed": int(executed["steps"])},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": TRAFFIC_BYTES if args.workload == "globule" else None,
                     "kernel": "traj_kernel<%s, type %d>" % (cfg["GasBuffer"], info.force_type),
                     "kernel_ms": round(kern_ms, 3), "flops_per_launch": int(flops)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg)
    shutil.rmtree(tmp, ignore_errors=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
