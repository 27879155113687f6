"""Calibrates the CPU baseline (oracle/cpu_baseline.py, this project's OpenMP
restatement built with the reference's release flags) against the reference
program itself, built from its sources by oracle/Makefile into oracle/_ref/.
Runs only where /root/reference was compiled (this container); the GPU host
never receives the reference binaries.

For every input both programs run the whole 10 x 10^4-trajectory job on the
same threads; printed: per-iteration Omega of both at the reference's printed
precision (%g), the CCS, and the median "CCS time" of `repeats` runs each.

Usage: python oracle/calibrate.py [--threads 8] [--repeats 3] [names...]
"""
import argparse
import json
import os
import re
import shutil
import statistics
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

import cpu_baseline as CB  # noqa: E402


def inputs(tmp):
    from common import PUBLISHED, globule_cfg
    from massccs_amd import synth
    d = {k: dict(v) for k, v in PUBLISHED.items()}
    d["globule_N2"] = globule_cfg(tmp)
    # the 50,000-atom ellipsoid / N2 (the benchmark's headline target), on a
    # 8,000-trajectory prefix of its sample stream (the long orbits of its
    # chaotic trajectories dominate the time of small samples)
    path = os.path.join(tmp, "ellipsoid_50k.xyz")
    s, xyz, q = synth.globule(n=50000, axes=(60.0, 45.0, 35.0), net_charge=20.0, seed=2104)
    synth.write_xyz(path, s, xyz, q, comment="synthetic ellipsoid 50000 atoms, net +20e")
    d["ellipsoid50k_N2"] = {"targetFileName": path, "GasBuffer": "N2", "Short-range cutoff": "yes",
                            "LJ-cutoff": 12.0, "Long-range forces": "yes", "Long-range cutoff": "yes",
                            "Coul-cutoff": 25.0, "polarizability": "yes", "alpha": 1.7403,
                            "seed": 2104, "dt": 10.0, "Temp": 298.0, "numberProbe": 8000,
                            "nIter": 1}
    return d


def run_reference(exe, cfg, threads, tmp):
    files = [cfg["targetFileName"]] + ([cfg["force-field"]] if "force-field" in cfg else [])
    for f in files:
        dst = os.path.join(tmp, os.path.basename(f))
        if not os.path.exists(dst):
            shutil.copy(f, dst)
    c = dict(cfg, targetFileName=os.path.basename(cfg["targetFileName"]), nthreads=threads)
    with open(os.path.join(tmp, "in.json"), "w") as fh:
        json.dump(c, fh)
    out = subprocess.run([exe, "in.json"], cwd=tmp, capture_output=True, text=True,
                         check=True).stdout
    t = float(re.search(r"CCS time: (\S+) s", out).group(1))
    om = re.findall(r"^omega: (\S+)$", out, re.M)
    mean = re.search(r"average value of CCS = (\S+) Ang", out).group(1)
    return t, om, mean


def main():
    import massccs_amd as M
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=CB.host_cores())
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("names", nargs="*")
    a = ap.parse_args()
    var = CB.variant()
    exe = os.path.join(HERE, "_ref", "massccs_ref_v4" if var == "v4" else "massccs_ref")
    if not os.path.exists(exe):
        sys.exit("oracle/_ref not built (make -C oracle ref)")
    tmp = tempfile.mkdtemp(prefix="mccs_cal_")
    try:
        cfgs = inputs(tmp)
        for name in a.names or list(cfgs):
            cfg = cfgs[name]
            ref = [run_reference(exe, cfg, a.threads, tmp) for _ in range(a.repeats)]
            P = M.Prepared(M.config_from_dict(cfg))
            port = CB.measure(P, cfg, cfg["nIter"], cfg["numberProbe"], threads=a.threads,
                              repeats=a.repeats, var=var)
            om_port = ["%g" % w for w in port["omega"]]
            same = [x == y for x, y in zip(ref[0][1], om_port)]
            t_ref = statistics.median(r[0] for r in ref)
            print(json.dumps({"input": name, "threads": a.threads, "variant": var,
                              "omega_reference": ref[0][1], "omega_port": om_port,
                              "omega_identical": sum(same), "of": len(same),
                              "ccs_reference": ref[0][2], "ccs_port": "%g" % port["mean"],
                              "ccs_time_reference_s": round(t_ref, 3),
                              "ccs_time_port_s": round(port["seconds"], 3),
                              "time_ratio_port_over_reference": round(port["seconds"] / t_ref, 3),
                              "runs_reference_s": [round(r[0], 3) for r in ref],
                              "runs_port_s": [round(x, 3) for x in port["all_seconds"]]}),
                  flush=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
