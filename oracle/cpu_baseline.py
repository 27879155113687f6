"""CPU baseline of the benchmark: this project's own OpenMP restatement of the
reference path (oracle.c) built with the reference's release flags
(liboracle_fast_v3.so / liboracle_fast_v4.so, -O3 -ffast-math -fopenmp,
CMakeLists.txt:66-90) and timed over the reference's "CCS time" window
(src/System.cpp:160-311): the serial mt19937 sample generation, the
trajectory loop (OpenMP, schedule(dynamic), src/System.cpp:204-285) and the
tally.  Test and measurement infrastructure: never part of the product.

The target preparation (orientation, ellipsoid, linked cells) comes from the
product's host code, which is bit-identical to the reference's; the
reference excludes it from its CCS time as well.
"""
import os
import statistics
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.join(os.path.dirname(HERE), "tests")
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)

import oracle_py as O  # noqa: E402


def cpu_flags():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("flags"):
                    return set(ln.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def variant():
    """The timing build that -march=native would select on this host: v4 where
    the AVX-512 foundation set is present, else v3."""
    v4 = {"avx512f", "avx512bw", "avx512dq", "avx512vl", "avx512cd"} <= cpu_flags()
    return "v4" if v4 else "v3"


def oracle_prep(P):
    """Oracle context inputs from a prepared target (massccs_amd.Prepared)."""
    i = P.info
    a = P.atoms()
    return dict(natoms=i.natoms, force_type=i.force_type, grid=[i.nx, i.ny, i.nz, i.ncells],
                alpha=i.alpha, mu=i.mu, abc=[i.a, i.b, i.c], bmax=i.bmax,
                halfbox=list(i.half_box), **a)


def ccs_time(orc, n_iter, n_traj, threads):
    """One CCS-time window: samples, trajectories, tally.  Returns
    (seconds, omega per iteration, mean, err)."""
    t0 = time.perf_counter()
    u = orc.samples(n_iter * n_traj)
    rec = orc.run(u, nthreads=threads)
    om, cnt, mean, err = O.tally(rec, n_iter, n_traj)
    return time.perf_counter() - t0, om, mean, err


def measure(P, cfg, n_iter, n_traj, threads=None, repeats=5, var=None):
    """Median CCS time of `repeats` runs of n_iter x n_traj trajectories."""
    threads = threads or host_cores()
    var = var or variant()
    orc = O.Oracle(oracle_prep(P), cfg, variant=var)
    runs = []
    for k in range(repeats):
        # progress on stderr: a long CPU run is not mistaken for a hang
        print("cpu baseline: run %d of %d, %d trajectories on %d threads"
              % (k + 1, repeats, n_iter * n_traj, threads), file=sys.stderr, flush=True)
        runs.append(ccs_time(orc, n_iter, n_traj, threads))
    t = statistics.median(r[0] for r in runs)
    return {"seconds": t, "all_seconds": [r[0] for r in runs], "omega": runs[0][1].tolist(),
            "mean": runs[0][2], "err": runs[0][3], "threads": threads, "variant": var,
            "trajectories": n_iter * n_traj}


def baseline(P, cfg, budget_s=15.0, repeats=5, threads=None):
    """The bench line's cpu_baseline: a bounded prefix of the workload's own
    sample stream (about budget_s of CPU work per run), median of `repeats`
    runs, on all host cores."""
    threads = threads or host_cores()
    var = variant()
    orc = O.Oracle(oracle_prep(P), cfg, variant=var)
    # A small prefix is timed by its longest trajectory (the threads idle
    # while one orbiting trajectory finishes), which would make the rate
    # depend on the prefix: the size is estimated twice, from a probe and
    # then from a first run of the estimate, and is at least 32 trajectories
    # per thread.
    n0 = 4 * threads
    print("cpu baseline: probe of %d trajectories" % n0, file=sys.stderr, flush=True)
    t0 = ccs_time(orc, 1, n0, threads)[0]
    n1 = int(min(10 ** 6, max(32 * threads, n0 * budget_s / max(t0, 1e-3))))
    print("cpu baseline: sizing run of %d trajectories" % n1, file=sys.stderr, flush=True)
    t1 = ccs_time(orc, 1, n1, threads)[0]
    n = int(min(10 ** 6, max(32 * threads, n1 * budget_s / max(t1, 1e-3))))
    m = measure(P, cfg, 1, n, threads=threads, repeats=repeats, var=var)
    return {"value": n / m["seconds"], "unit": "trajectories/s", "cores": threads, "kind": "port",
            "sample": "first %d trajectories of the same seeded sample stream (nIter 1); "
                      "median of %d CCS-time windows (samples + trajectories + tally) of %s "
                      "(liboracle_fast_%s.so: oracle.c at -O3 -ffast-math -fopenmp -march=x86-64-%s, "
                      "OpenMP schedule(dynamic)) on %d threads of %s; runs %s s"
                      % (n, repeats, "this project's C restatement of the reference path", var, var,
                         threads, cpu_model(), ", ".join("%.2f" % x for x in m["all_seconds"])),
            "model": cpu_model(), "nproc": os.cpu_count()}
