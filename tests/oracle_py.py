"""ctypes view of the CPU oracle (oracle/liboracle.so) and of the reference
harness (oracle/_ref/ref_harness).  Test infrastructure only: these are the
checkers the product is compared against, never part of the product."""
import ctypes as C
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
# timing builds (the reference's release flags): the CPU baseline of bench.py
ORACLE_FAST_SO = {"v3": os.path.join(REPO, "oracle", "liboracle_fast_v3.so"),
                  "v4": os.path.join(REPO, "oracle", "liboracle_fast_v4.so")}
REF_DIR = os.path.join(REPO, "oracle", "_ref")
HARNESS = os.path.join(REF_DIR, "ref_harness")
GOLDEN = os.path.join(REPO, "tests", "golden")

GAS_FLAGS = {"He": 1, "N2": 2, "CO2": 3, "Ar": 4, "co2": 5}

# kernel ids of oracle.h, by the reference member name they restate
KERNELS = [
    "lennardjones", "lennardjones_LC", "lennardjones_coulomb",
    "lennardjones_coulomb_LC", "coulomb", "coulomb_LC",
    "lennardjones_induced_dipole", "lennardjones_induced_dipole_LC",
    "coulomb_induced_dipole_iso", "coulomb_induced_dipole_iso_LC",
    "lennardjones_CO2", "lennardjones_LC_CO2", "lennardjones_coulomb_CO2",
    "lennardjones_coulomb_LC_CO2",
    "lennardjones_coulomb_induced_dipole_iso_CO2",
    "lennardjones_coulomb_induced_dipole_iso_LC_CO2",
]
KERNEL_ID = {k: i for i, k in enumerate(KERNELS)}


class OracleParams(C.Structure):
    _fields_ = [
        ("natoms", C.c_int),
        ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
        ("z", C.POINTER(C.c_double)), ("q", C.POINTER(C.c_double)),
        ("eps", C.POINTER(C.c_double)), ("sig", C.POINTER(C.c_double)),
        ("eps_c", C.POINTER(C.c_double)), ("sig_c", C.POINTER(C.c_double)),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("cell", C.c_double), ("lj_cutoff", C.c_double),
        ("coul_cutoff", C.c_double), ("alpha", C.c_double),
        ("long_range_flag", C.c_int), ("gas", C.c_int),
        ("force_type", C.c_int), ("mu", C.c_double),
        ("temperature", C.c_double), ("dt", C.c_double),
        ("a", C.c_double), ("b", C.c_double), ("c", C.c_double),
        ("bmax", C.c_double), ("half_box", C.c_double * 3), ("co2_vcm_fix", C.c_int),
        ("max_total_steps", C.c_int),
    ]


class OracleRecord(C.Structure):
    _fields_ = [("hit", C.c_int32), ("success", C.c_int32),
                ("tries", C.c_int32), ("steps", C.c_int32),
                ("chi", C.c_double), ("domega", C.c_double)]


RECORD_DTYPE = np.dtype([("hit", "<i4"), ("success", "<i4"), ("tries", "<i4"),
                         ("steps", "<i4"), ("chi", "<f8"), ("domega", "<f8")])

_libs = {}


def lib(variant="strict"):
    """The oracle library: "strict" (the checker, bit-identical to the strict
    reference build) or a timing build "v3" / "v4" (-O3 -ffast-math, the
    reference's release flags, for x86-64-v3 / AVX-512 hosts)."""
    if variant not in _libs:
        path = ORACLE_SO if variant == "strict" else ORACLE_FAST_SO[variant]
        if not os.path.exists(path):
            subprocess.run(["make", "-C", os.path.join(REPO, "oracle"),
                            "oracle" if variant == "strict" else "fast"],
                           check=True, capture_output=True)
        L = C.CDLL(path)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(OracleParams)]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_site_force.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                        C.c_double, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
        L.oracle_samples.argtypes = [C.c_uint32, C.c_int, C.c_double, C.c_int64,
                                     C.POINTER(C.c_double)]
        L.oracle_mt_raw.argtypes = [C.c_uint32, C.c_int64, C.POINTER(C.c_uint32)]
        L.oracle_vel_generator.restype = C.c_double
        L.oracle_vel_generator.argtypes = [C.c_double, C.c_double, C.c_double]
        L.oracle_run.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int64,
                                 C.c_void_p, C.c_int]
        L.oracle_setup.restype = C.c_int
        L.oracle_setup.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 5
        L.oracle_tally.argtypes = [C.c_void_p, C.c_int64, C.c_int64,
                                   C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.oracle_stencil_sums.argtypes = [C.c_void_p, C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64)]
        for n in ("oracle_rattle_first", "oracle_rattle_second"):
            getattr(L, n).restype = None
        _libs[variant] = L
    return _libs[variant]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def input_defaults(cfg):
    """Run parameters the reference derives from input.json
    (src/Input.cpp:49-263, defaults src/headers/Constants.h:68-83)."""
    gas = cfg.get("GasBuffer", "He")
    gflag = GAS_FLAGS[gas]
    yes = lambda k, d: cfg.get(k, d) == "yes"
    short = yes("Short-range cutoff", "yes")
    longf = yes("Long-range forces", "no")
    longc = yes("Long-range cutoff", "yes")
    pol = yes("polarizability", "yes" if gflag == 1 else "no")
    return dict(gas=gflag, short=short, long_flag=longf, long_cut=longc, pol=pol,
                lj_cutoff=float(cfg.get("LJ-cutoff", 12.0)),
                coul_cutoff=float(cfg.get("Coul-cutoff", 25.0)),
                skin=float(cfg.get("skin", 0.01)),
                dt=float(cfg.get("dt", 10.0)), T=float(cfg.get("Temp", 298.0)),
                seed=int(cfg.get("seed", 20162104)),
                n_probe=int(cfg.get("numberProbe", 10000)),
                n_iter=int(cfg.get("nIter", 10)))


class Oracle:
    """Owns an oracle context built from a prepared target."""

    def __init__(self, prep, cfg, variant="strict", max_total_steps=0):
        d = input_defaults(cfg)
        self.L = lib(variant)
        self.prep, self.cfg, self.d = prep, cfg, d
        self._keep = []
        arr = lambda k: self._hold(np.ascontiguousarray(prep[k], dtype=np.float64))
        p = OracleParams()
        p.natoms = prep["natoms"]
        for k in ("x", "y", "z", "q", "eps", "sig"):
            setattr(p, k, _dp(arr(k)))
        if "eps_c" in prep:
            p.eps_c, p.sig_c = _dp(arr("eps_c")), _dp(arr("sig_c"))
        p.nx, p.ny, p.nz = prep["grid"][:3]
        p.cell = d["lj_cutoff"] + d["skin"]
        p.lj_cutoff, p.coul_cutoff = d["lj_cutoff"], d["coul_cutoff"]
        p.alpha = prep["alpha"]
        p.long_range_flag = int(d["long_flag"])
        p.gas, p.force_type = d["gas"], prep["force_type"]
        p.mu, p.temperature, p.dt = prep["mu"], d["T"], d["dt"]
        p.a, p.b, p.c = prep["abc"]
        p.bmax = prep["bmax"]
        for i in range(3):
            p.half_box[i] = prep["halfbox"][i]
        p.co2_vcm_fix = 1 if cfg.get("CO2 vcm fix", "no") == "yes" else 0
        p.max_total_steps = int(max_total_steps)
        self.params = p
        self.ctx = self.L.oracle_create(C.byref(p))

    def _hold(self, a):
        self._keep.append(a)
        return a

    def __del__(self):
        if getattr(self, "ctx", None):
            self.L.oracle_destroy(self.ctx)
            self.ctx = None

    def site_force(self, kernel, r, qi):
        r = np.ascontiguousarray(r, dtype=np.float64)
        f = np.zeros(3)
        U = C.c_double()
        kid = KERNEL_ID[kernel] if isinstance(kernel, str) else kernel
        self.L.oracle_site_force(self.ctx, kid, _dp(r), qi, _dp(f), C.byref(U))
        return f, U.value

    def samples(self, n, seed=None):
        u = np.zeros((n, 9))
        multi = 1 if self.d["gas"] in (2, 3) else 0
        s = self.d["seed"] if seed is None else seed
        self.L.oracle_samples(s, multi, self.prep["bmax"], n, _dp(u))
        return u

    def run(self, u, nthreads=0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(len(u), dtype=RECORD_DTYPE)
        self.L.oracle_run(self.ctx, _dp(u), len(u), out.ctypes.data_as(C.c_void_p), nthreads)
        return out

    def setup(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        pos, vel, rcm, vcm = np.zeros(9), np.zeros(9), np.zeros(3), np.zeros(3)
        hit = self.L.oracle_setup(self.ctx, _dp(u), _dp(pos), _dp(vel), _dp(rcm), _dp(vcm))
        return hit, pos, vel, rcm, vcm

    def stencil_sums(self):
        a, b = C.c_int64(), C.c_int64()
        self.L.oracle_stencil_sums(self.ctx, C.byref(a), C.byref(b))
        return a.value, b.value


def tally(rec, n_iter, n_traj):
    rec = np.ascontiguousarray(rec, dtype=RECORD_DTYPE)
    om = np.zeros(n_iter)
    cnt = np.zeros(3 * n_iter, dtype=np.int64)
    mean, err = C.c_double(), C.c_double()
    lib().oracle_tally(rec.ctypes.data_as(C.c_void_p), n_iter, n_traj, _dp(om),
                       cnt.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(mean), C.byref(err))
    return om, cnt.reshape(n_iter, 3), mean.value, err.value


# ---------------------------------------------------------------- harness
def ref_available():
    return os.path.exists(HARNESS)


def _with_json(cfg, target_files, fn):
    tmp = tempfile.mkdtemp(prefix="mccs_ref_")
    try:
        for f in target_files:
            os.symlink(os.path.abspath(f), os.path.join(tmp, os.path.basename(f)))
        c = dict(cfg)
        c["targetFileName"] = os.path.basename(cfg["targetFileName"])
        if "force-field" in c:
            c["force-field"] = os.path.basename(c["force-field"])
        c["numberProbe"] = 0
        c["nIter"] = 1
        with open(os.path.join(tmp, "in.json"), "w") as fh:
            json.dump(c, fh)
        return fn(tmp)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def _files(cfg):
    fs = [cfg["targetFileName"]]
    if "force-field" in cfg:
        fs.append(cfg["force-field"])
    return fs


def harness_prep(cfg):
    def go(tmp):
        out = subprocess.run([HARNESS, "prep", "in.json"], cwd=tmp, check=True,
                             capture_output=True, text=True).stdout
        return parse_prep(out)
    return _with_json(cfg, _files(cfg), go)


def harness_force(cfg, lines):
    def go(tmp):
        with open(os.path.join(tmp, "pts.txt"), "w") as fh:
            for k, s, r in lines:
                fh.write("%s %d %.17g %.17g %.17g\n" % (k, s, r[0], r[1], r[2]))
        out = subprocess.run([HARNESS, "force", "in.json", "pts.txt"], cwd=tmp, check=True,
                             capture_output=True, text=True).stdout
        res = []
        for ln in out.splitlines():
            t = ln.split()
            res.append((t[0], int(t[1]), np.array([float(v) for v in t[5:8]]), float(t[8])))
        return res
    return _with_json(cfg, _files(cfg), go)


def harness_traj(cfg, first, count):
    def go(tmp):
        out = subprocess.run([HARNESS, "traj", "in.json", str(first), str(count)], cwd=tmp,
                             check=True, capture_output=True, text=True).stdout
        rows = [ln.split() for ln in out.splitlines() if ln.startswith("traj")]
        return (np.array([int(r[2]) for r in rows]), np.array([int(r[3]) for r in rows]),
                np.array([float(r[4]) for r in rows]))
    return _with_json(cfg, _files(cfg), go)


def parse_prep(text):
    d = {"atoms": [], "cells": []}
    for ln in text.splitlines():
        t = ln.split()
        if not t:
            continue
        k, v = t[0], t[1:]
        if k == "equipotential":  # the reference's destructor message
            continue
        if k == "atom":
            d["atoms"].append([float(x) for x in v])
        elif k == "cell":
            d["cells"].append([int(x) for x in v])
        elif k in ("force_type", "natoms"):
            d[k] = int(v[0])
        elif k in ("grid", "stencil_sums"):
            d[k] = [int(x) for x in v]
        elif len(v) == 1:
            d[k] = float(v[0])
        else:
            d[k] = [float(x) for x in v]
    A = np.array(d.pop("atoms"))
    for i, k in enumerate(("x", "y", "z", "q", "m", "eps", "sig")):
        d[k] = A[:, i].copy()
    if A.shape[1] > 7:
        d["eps_c"], d["sig_c"] = A[:, 7].copy(), A[:, 8].copy()
    return d
