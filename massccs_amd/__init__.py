"""massccs_amd: MI355X trajectory-method engine for ion-neutral collision
cross sections (a drop-in for the MassCCS trajectory loop).

The product is libmassccs_amd.so (C ABI in include/massccs.h) and the
`massccs` command line built next to this file.  This module is a thin ctypes
view of that C ABI for Python callers (tests, benchmarks, scripts); it loads
the in-tree library and raises if it has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MASSCCS_LIB selects another build of the library (e.g. a different launch shape)
LIB_PATH = os.environ.get("MASSCCS_LIB") or os.path.join(HERE, "libmassccs_amd.so")
CLI_PATH = os.path.join(HERE, "massccs")

GAS = {"He": 1, "N2": 2, "CO2": 3, "Ar": 4, "co2": 5}
SITE_ATOM, SITE_N, SITE_DUMMY, SITE_O, SITE_C = 0, 1, 2, 3, 4

ERRORS = {1: "invalid argument", 2: "I/O error", 3: "parse error", 4: "unsupported",
          5: "device error", 6: "out of memory"}


class MassCCSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%s)" % (msg, ERRORS.get(code, code)))
        self.code = code


class Config(C.Structure):
    """mccs_config of include/massccs.h (fields of the reference's Input)."""
    _fields_ = [
        ("target_file", C.c_char * 1024), ("force_field", C.c_char * 1024),
        ("n_probe", C.c_uint32), ("n_iter", C.c_uint32), ("seed", C.c_uint32),
        ("nthreads", C.c_uint32),
        ("dt", C.c_double), ("temperature", C.c_double), ("skin", C.c_double),
        ("lj_cutoff", C.c_double), ("coul_cutoff", C.c_double), ("alpha", C.c_double),
        ("gas", C.c_int32), ("equipotential", C.c_int32), ("short_range_cutoff", C.c_int32),
        ("long_range_flag", C.c_int32), ("long_range_cutoff", C.c_int32),
        ("polarizability", C.c_int32), ("user_ff", C.c_int32), ("alpha_set", C.c_int32),
        ("co2_vcm_fix", C.c_int32),
    ]


class PreparedInfo(C.Structure):
    _fields_ = [
        ("gas", C.c_int32), ("force_type", C.c_int32), ("nsites", C.c_int32),
        ("natoms", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("ncells", C.c_int32), ("filled_cells", C.c_int32), ("max_atoms_cell", C.c_int32),
        ("mass", C.c_double), ("charge", C.c_double), ("mu", C.c_double), ("alpha", C.c_double),
        ("a", C.c_double), ("b", C.c_double), ("c", C.c_double), ("bmax", C.c_double),
        ("half_box", C.c_double * 3), ("cell", C.c_double), ("corner", C.c_double * 3),
        ("radius", C.c_double), ("stencil1_atoms", C.c_int64), ("stencil2_atoms", C.c_int64),
    ]


class EngineOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("lanes", C.c_int32), ("block", C.c_int32),
                ("use_lds", C.c_int32), ("count_work", C.c_int32), ("spec_below", C.c_int32),
                ("list_cap", C.c_int32), ("order", C.c_int32)]


class BatchOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("lanes", C.c_int32), ("count_work", C.c_int32),
                ("use_graph", C.c_int32), ("skin", C.c_double), ("order", C.c_int32)]


class MlpSpec(C.Structure):
    """mccs_mlp_spec: Fourier frequencies, input centre/scale, output scales."""
    _fields_ = [("freq", (C.c_float * 3) * 32), ("centre", C.c_double * 3),
                ("in_scale", C.c_double), ("out_scale", C.c_double * 2)]


class SurrogateOpts(C.Structure):
    _fields_ = [("n_train", C.c_uint64), ("epochs", C.c_int32), ("batch", C.c_int32),
                ("lr", C.c_double), ("shell", C.c_double), ("blend", C.c_double),
                ("grid", C.c_double), ("feature_length", C.c_double), ("seed", C.c_uint32),
                ("table_source", C.c_int32)]


class SurrogateReport(C.Structure):
    _fields_ = [("train_loss", C.c_double), ("val_rmse_u", C.c_double * 2),
                ("val_rms_u", C.c_double * 2), ("val_rmse_f", C.c_double * 2),
                ("val_rms_f", C.c_double * 2), ("seconds_sample", C.c_double),
                ("seconds_train", C.c_double), ("seconds_table", C.c_double),
                ("n_train", C.c_uint64), ("grid_nodes", C.c_uint64), ("spec", MlpSpec)]


MLP_NPARAM = 41602
NOALLOC_MAX_TRAJECTORIES = 65536  # MCCS_NOALLOC_MAX_TRAJECTORIES

RECORD_DTYPE = np.dtype([("hit", "<i4"), ("success", "<i4"), ("tries", "<i4"), ("steps", "<i4"),
                         ("chi", "<f8"), ("domega", "<f8"), ("visits", "<u8"), ("lj", "<u8"),
                         ("coul", "<u8"), ("dip", "<u8")])

# every symbol declared in include/massccs.h
EXPORTS = [
    "mccs_last_error", "mccs_config_default", "mccs_config_load_json", "mccs_prepare",
    "mccs_prepare_atoms", "mccs_prepared_destroy", "mccs_prepared_get_info",
    "mccs_prepared_get_atoms", "mccs_prepared_report", "mccs_samples", "mccs_engine_create",
    "mccs_engine_destroy", "mccs_run", "mccs_run_device", "mccs_eval_sites", "mccs_tally",
    "mccs_tally_device", "mccs_version", "mccs_engine_last_work", "mccs_batch_create",
    "mccs_batch_destroy", "mccs_batch_run", "mccs_batch_autotune", "mccs_batch_shapes",
    "mccs_batch_last_work", "mccs_mlp_eval", "mccs_mlp_loss_grad", "mccs_mlp_jvp_loss_grad",
    "mccs_mlp_train",
    "mccs_engine_build_surrogate", "mccs_engine_use_surrogate", "mccs_run_device_noalloc",
    "mccs_batch_build_surrogate", "mccs_batch_use_surrogate", "mccs_batch_run_device",
    "mccs_cells_device", "mccs_inertia_device",
]

_lib = None
_dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))


def lib():
    """The loaded C library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MassCCSError(5, "libmassccs_amd.so is not built (run __graft_entry__.build() or "
                              "make -C massccs_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    L.mccs_last_error.restype = C.c_char_p
    L.mccs_version.restype = C.c_char_p
    st = C.c_int
    L.mccs_config_default.argtypes = [C.POINTER(Config)]
    L.mccs_config_load_json.argtypes = [C.c_char_p, C.POINTER(Config)]
    L.mccs_prepare.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
    L.mccs_prepare_atoms.argtypes = [C.POINTER(Config), C.c_int64, C.POINTER(C.c_char_p),
                                     C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_void_p)]
    L.mccs_prepared_destroy.argtypes = [C.c_void_p]
    L.mccs_prepared_destroy.restype = None
    L.mccs_prepared_get_info.argtypes = [C.c_void_p, C.POINTER(PreparedInfo)]
    L.mccs_prepared_get_atoms.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 9
    L.mccs_prepared_report.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.mccs_samples.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_double)]
    L.mccs_engine_create.argtypes = [C.c_void_p, C.POINTER(EngineOpts), C.POINTER(C.c_void_p)]
    L.mccs_engine_destroy.argtypes = [C.c_void_p]
    L.mccs_engine_destroy.restype = None
    L.mccs_run.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_uint64, C.c_void_p]
    L.mccs_run_device.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
    L.mccs_run_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                          C.c_void_p]
    L.mccs_eval_sites.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.c_uint64,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.mccs_engine_last_work.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
    L.mccs_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(BatchOpts),
                                    C.POINTER(C.c_void_p)]
    L.mccs_batch_destroy.argtypes = [C.c_void_p]
    L.mccs_batch_destroy.restype = None
    L.mccs_batch_run.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_void_p)]
    L.mccs_batch_autotune.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_void_p),
                                      C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.mccs_batch_shapes.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
    L.mccs_batch_last_work.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
    fp = C.POINTER(C.c_float)
    L.mccs_mlp_eval.argtypes = [C.c_int32, C.POINTER(MlpSpec), fp, C.POINTER(C.c_double),
                                C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.mccs_mlp_loss_grad.argtypes = [C.c_int32, C.POINTER(MlpSpec), fp, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.c_uint64, C.POINTER(C.c_double), fp]
    dp = C.POINTER(C.c_double)
    L.mccs_mlp_jvp_loss_grad.argtypes = [C.c_int32, C.POINTER(MlpSpec), fp, dp, dp, dp, dp, C.c_double,
                                         C.c_uint64, dp, fp, dp, dp]
    L.mccs_mlp_train.argtypes = [C.c_int32, C.POINTER(MlpSpec), fp, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.c_uint64, C.c_int32, C.c_int32,
                                 C.c_double, C.c_uint32, C.POINTER(C.c_double)]
    L.mccs_engine_build_surrogate.argtypes = [C.c_void_p, C.POINTER(SurrogateOpts),
                                              C.POINTER(SurrogateReport), fp]
    L.mccs_engine_use_surrogate.argtypes = [C.c_void_p, C.c_int32]
    L.mccs_batch_build_surrogate.argtypes = [C.c_void_p, C.POINTER(SurrogateOpts),
                                             C.POINTER(SurrogateReport)]
    L.mccs_batch_use_surrogate.argtypes = [C.c_void_p, C.c_int32]
    L.mccs_cells_device.argtypes = [C.c_int32, C.c_int64, C.c_void_p, C.c_double,
                                    C.POINTER(C.c_double), C.c_int32, C.c_int32, C.c_int32,
                                    C.c_void_p, C.c_void_p, C.c_void_p]
    L.mccs_inertia_device.argtypes = [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.c_void_p]
    L.mccs_batch_run_device.argtypes = [C.c_void_p, C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_void_p]
    L.mccs_tally.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_double),
                             C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.mccs_tally_device.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p]
    for n in EXPORTS:
        if n not in ("mccs_last_error", "mccs_version", "mccs_prepared_destroy",
                     "mccs_engine_destroy", "mccs_batch_destroy"):
            getattr(L, n).restype = st
    _lib = L
    return L


def _check(status):
    if status != 0:
        raise MassCCSError(status, lib().mccs_last_error().decode(errors="replace"))


def config_from_json(path):
    c = Config()
    _check(lib().mccs_config_load_json(os.fsencode(path), C.byref(c)))
    return c


def config_default(target_file=""):
    c = Config()
    _check(lib().mccs_config_default(C.byref(c)))
    c.target_file = os.fsencode(target_file)
    return c


def config_from_dict(d):
    """Config from a dict holding input.json keys (reference semantics)."""
    import json
    import tempfile
    with tempfile.NamedTemporaryFile("w", suffix=".json", delete=False) as fh:
        json.dump(d, fh)
        name = fh.name
    try:
        return config_from_json(name)
    finally:
        os.unlink(name)


def _symbol_pointers(symbols, n):
    """NUL-terminated element symbols and a char* per atom for
    mccs_prepare_atoms.  ASCII symbols (every element symbol) are packed
    without a Python loop: the UCS-4 code points of a str array become the
    bytes of one fixed-width buffer; anything else is encoded one by one."""
    a = np.asarray(symbols)
    if len(a) != n:
        raise ValueError("one symbol per atom expected")
    cp = None
    if a.dtype.kind == "U" and a.dtype.itemsize > 0:
        cp = np.ascontiguousarray(a).view(np.uint32).reshape(n, -1)
        if cp.size and int(cp.max()) >= 128:
            cp = None
    elif a.dtype.kind == "S" and a.dtype.itemsize > 0:
        cp = np.ascontiguousarray(a).view(np.uint8).reshape(n, -1)
    if cp is not None:
        w = cp.shape[1] + 1
        buf = np.zeros((n, w), dtype=np.uint8)
        buf[:, :w - 1] = cp
        ptrs = np.uint64(buf.ctypes.data) + np.uint64(w) * np.arange(n, dtype=np.uint64)
        return (buf, ptrs), ptrs.ctypes.data_as(C.POINTER(C.c_char_p))
    arr = (C.c_char_p * n)(*[os.fsencode(x) for x in symbols])
    return arr, arr


class Prepared:
    """Prepared target (MoleculeTarget + LinkedCell + ellipsoid) on the host."""

    def __init__(self, cfg, symbols=None, xyz=None, q=None):
        self.cfg = cfg
        h = C.c_void_p()
        if symbols is None:
            _check(lib().mccs_prepare(C.byref(cfg), C.byref(h)))
        else:
            xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
            n = len(xyz)
            buf, arr = _symbol_pointers(symbols, n)  # buf holds the strings during the call
            qa = None if q is None else np.ascontiguousarray(q, dtype=np.float64)
            _check(lib().mccs_prepare_atoms(C.byref(cfg), n, arr, _dp(xyz),
                                            None if qa is None else _dp(qa), C.byref(h)))
            del buf
        self.h = h
        self._destroy = lib().mccs_prepared_destroy  # bound now: usable at interpreter exit

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def info(self):
        i = PreparedInfo()
        _check(lib().mccs_prepared_get_info(self.h, C.byref(i)))
        return i

    def atoms(self):
        n = self.info.natoms
        out = {k: np.zeros(n) for k in ("x", "y", "z", "q", "m", "eps", "sig", "eps_c", "sig_c")}
        _check(lib().mccs_prepared_get_atoms(self.h, *[_dp(out[k]) for k in
                                                       ("x", "y", "z", "q", "m", "eps", "sig",
                                                        "eps_c", "sig_c")]))
        return out

    def report(self):
        n = C.c_size_t()
        _check(lib().mccs_prepared_report(self.h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        _check(lib().mccs_prepared_report(self.h, buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def samples(self, n, first=0):
        u = np.zeros((n, 9))
        _check(lib().mccs_samples(self.h, first, n, _dp(u)))
        return u


class Engine:
    """Device engine on one GPU (target resident in HBM)."""

    def __init__(self, prepared, device=0, lanes=0, block=0, use_lds=-1, count_work=True,
                 spec_below=-1, list_cap=0, order=0):
        o = EngineOpts(device, lanes, block, use_lds, 1 if count_work else 0, spec_below, list_cap,
                       order)
        h = C.c_void_p()
        _check(lib().mccs_engine_create(prepared.h, C.byref(o), C.byref(h)))
        self.h, self.prepared = h, prepared
        self._destroy = lib().mccs_engine_destroy  # bound now: usable at interpreter exit

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    __del__ = close

    def run(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        rec = np.zeros(len(u), dtype=RECORD_DTYPE)
        _check(lib().mccs_run(self.h, _dp(u), len(u), rec.ctypes.data_as(C.c_void_p)))
        return rec

    def run_device(self, u_ptr, n, rec_ptr, stream=None):
        _check(lib().mccs_run_device(self.h, C.c_void_p(u_ptr), n, C.c_void_p(rec_ptr),
                                     C.c_void_p(stream) if stream else None))

    def run_device_noalloc(self, u_ptr, n, rec_ptr, stream=None):
        """mccs_run_device_noalloc: the same run with the work queue in the
        library's static device workspace (at most NOALLOC_MAX_TRAJECTORIES)."""
        _check(lib().mccs_run_device_noalloc(self.h, C.c_void_p(u_ptr), n, C.c_void_p(rec_ptr),
                                             C.c_void_p(stream) if stream else None))

    def tally_device(self, rec_ptr, n_iter, n_traj, omega_ptr, counts_ptr, mean_err_ptr,
                     stream=None):
        _check(lib().mccs_tally_device(self.h, C.c_void_p(rec_ptr), n_iter, n_traj,
                                       C.c_void_p(omega_ptr), C.c_void_p(counts_ptr),
                                       C.c_void_p(mean_err_ptr),
                                       C.c_void_p(stream) if stream else None))

    def last_work(self):
        """Executed totals of the last run: visits, lj, coul, dip, steps."""
        out = (C.c_uint64 * 5)()
        _check(lib().mccs_engine_last_work(self.h, out))
        return dict(zip(("visits", "lj", "coul", "dip", "steps"), list(out)))

    def build_surrogate(self, n_train=2_000_000, epochs=300, batch=8192, lr=5e-3, shell=8.0,
                        blend=4.0, grid=0.5, feature_length=0.0, seed=2104, exact_table=False,
                        table_source=None):
        """Range-split far field of this engine's target (N2 type 6, He/Ar/co2
        type 4): pairs nearer than shell keep the exact terms, the switch
        reaches zero at shell + blend, and the remainder psi is tabulated on a
        cubic B-spline grid -- from the network trained on the exact far field
        (default) or, with exact_table, from the exact field at the nodes
        (table_source 2: no far field, the ablation).
        Returns (report dict, trained parameters, spec)."""
        src = table_source if table_source is not None else (1 if exact_table else 0)
        o = SurrogateOpts(n_train, epochs, batch, lr, shell, blend, grid, feature_length, seed, src)
        r = SurrogateReport()
        params = np.zeros(MLP_NPARAM, dtype=np.float32)
        _check(lib().mccs_engine_build_surrogate(self.h, C.byref(o), C.byref(r),
                                                 params.ctypes.data_as(C.POINTER(C.c_float))))
        rep = {k: (list(getattr(r, k)) if k.startswith("val_") else getattr(r, k))
               for k, _ in SurrogateReport._fields_ if k != "spec"}
        return rep, params, r.spec

    def use_surrogate(self, on=True):
        _check(lib().mccs_engine_use_surrogate(self.h, 1 if on else 0))

    def eval_sites(self, site, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        f = np.zeros_like(xyz)
        U = np.zeros(len(xyz))
        _check(lib().mccs_eval_sites(self.h, site, _dp(xyz), len(xyz), _dp(f), _dp(U)))
        return f, U


class Batch:
    """Several prepared targets run together (one persistent launch per probe
    kind / force type, replayed from a hipGraph)."""

    def __init__(self, prepared_list, device=0, lanes=0, count_work=True, use_graph=True,
                 skin=0.0, order=0):
        self.targets = list(prepared_list)
        arr = (C.c_void_p * len(self.targets))(*[p.h.value for p in self.targets])
        o = BatchOpts(device, lanes, 1 if count_work else 0, 1 if use_graph else 0, skin, order)
        h = C.c_void_p()
        _check(lib().mccs_batch_create(arr, len(self.targets), C.byref(o), C.byref(h)))
        self.h = h
        self._destroy = lib().mccs_batch_destroy  # bound now: usable at interpreter exit

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    __del__ = close

    def _ptrs(self, samples):
        us = [np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 9) for u in samples]
        assert len(us) == len(self.targets)
        n = (C.c_uint64 * len(us))(*[len(u) for u in us])
        up = (C.c_void_p * len(us))(*[u.ctypes.data for u in us])
        return us, n, up

    def run(self, samples):
        """samples: one (n_t, 9) array per target; returns one record array each."""
        us, n, up = self._ptrs(samples)
        recs = [np.zeros(len(u), dtype=RECORD_DTYPE) for u in us]
        rp = (C.c_void_p * len(us))(*[r.ctypes.data for r in recs])
        _check(lib().mccs_batch_run(self.h, n, up, rp))
        return recs

    def run_device(self, n, sample_ptrs, record_ptrs, stream=None):
        """Device-resident run: n[t] trajectories of target t from the device
        samples at sample_ptrs[t] into the device records at record_ptrs[t],
        enqueued on stream (a HIP stream handle, None = the batch's own)."""
        k = len(self.targets)
        nn = (C.c_uint64 * k)(*[int(x) for x in n])
        sp = (C.c_void_p * k)(*sample_ptrs)
        rp = (C.c_void_p * k)(*record_ptrs)
        _check(lib().mccs_batch_run_device(self.h, nn, sp, rp,
                                           C.c_void_p(stream) if stream else None))

    def autotune(self, samples):
        """Trial runs over the given samples (a prefix of each target's);
        returns the tuner's report (JSON)."""
        import json
        us, n, up = self._ptrs(samples)
        ln = C.c_size_t()
        buf = C.create_string_buffer(1 << 20)
        _check(lib().mccs_batch_autotune(self.h, n, up, buf, len(buf), C.byref(ln)))
        return json.loads(buf.value.decode())

    def build_surrogate(self, n_train=2_000_000, epochs=300, batch=8192, lr=5e-3, shell=8.0,
                        blend=4.0, grid=0.5, feature_length=0.0, seed=2104, exact_table=False):
        """Far fields of every member that carries a dipole field (see
        Engine.build_surrogate); returns one report dict per target."""
        o = SurrogateOpts(n_train, epochs, batch, lr, shell, blend, grid, feature_length, seed,
                          1 if exact_table else 0)
        reps = (SurrogateReport * len(self.targets))()
        _check(lib().mccs_batch_build_surrogate(self.h, C.byref(o), reps))
        return [{k: (list(getattr(r, k)) if k.startswith("val_") else getattr(r, k))
                 for k, _ in SurrogateReport._fields_ if k != "spec"} for r in reps]

    def use_surrogate(self, on=True):
        _check(lib().mccs_batch_use_surrogate(self.h, 1 if on else 0))

    def shapes(self):
        k = len(self.targets)
        lanes = (C.c_int32 * k)()
        skin = (C.c_double * k)()
        _check(lib().mccs_batch_shapes(self.h, lanes, skin))
        return list(lanes), list(skin)

    def last_work(self):
        out = (C.c_uint64 * 5)()
        _check(lib().mccs_batch_last_work(self.h, out))
        return dict(zip(("visits", "lj", "coul", "dip", "steps"), list(out)))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def mlp_spec(freq, centre, in_scale, out_scale):
    sp = MlpSpec()
    for f in range(32):
        for d in range(3):
            sp.freq[f][d] = float(freq[f][d])
    for d in range(3):
        sp.centre[d] = float(centre[d])
    sp.in_scale = float(in_scale)
    sp.out_scale[0], sp.out_scale[1] = float(out_scale[0]), float(out_scale[1])
    return sp


def mlp_eval(spec, params, xyz, grad=True, device=0):
    """Network outputs (n, 2) and gradients (n, 2, 3) at points (n, 3), on the GPU."""
    params = np.ascontiguousarray(params, dtype=np.float32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    y = np.zeros((len(xyz), 2))
    dy = np.zeros((len(xyz), 2, 3)) if grad else None
    _check(lib().mccs_mlp_eval(device, C.byref(spec), _fp(params), _dp(xyz), len(xyz), _dp(y),
                               _dp(dy) if grad else None))
    return y, dy


def mlp_loss_grad(spec, params, xyz, e, device=0):
    params = np.ascontiguousarray(params, dtype=np.float32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    e = np.ascontiguousarray(e, dtype=np.float64).reshape(-1, 2)
    g = np.zeros(MLP_NPARAM, dtype=np.float32)
    loss = C.c_double()
    _check(lib().mccs_mlp_loss_grad(device, C.byref(spec), _fp(params), _dp(xyz), _dp(e),
                                    len(xyz), C.byref(loss), _fp(g)))
    return loss.value, g


def mlp_jvp_loss_grad(spec, params, xyz, dirs, e, d, w, device=0):
    """Loss of values and forward-mode directional derivatives and its
    parameter gradient; also the scaled outputs y and derivatives dy."""
    params = np.ascontiguousarray(params, dtype=np.float32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    e = np.ascontiguousarray(e, dtype=np.float64).reshape(-1, 2)
    d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1, 2)
    g = np.zeros(MLP_NPARAM, dtype=np.float32)
    y = np.zeros((len(xyz), 2))
    dy = np.zeros((len(xyz), 2))
    loss = C.c_double()
    _check(lib().mccs_mlp_jvp_loss_grad(device, C.byref(spec), _fp(params), _dp(xyz), _dp(dirs),
                                        _dp(e), _dp(d), float(w), len(xyz), C.byref(loss), _fp(g),
                                        _dp(y), _dp(dy)))
    return loss.value, g, y, dy


def mlp_train(spec, params, xyz, e, epochs=10, batch=4096, lr=2e-3, seed=1, device=0):
    params = np.array(params, dtype=np.float32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    e = np.ascontiguousarray(e, dtype=np.float64).reshape(-1, 2)
    hist = np.zeros(epochs)
    _check(lib().mccs_mlp_train(device, C.byref(spec), _fp(params), _dp(xyz), _dp(e), len(xyz),
                                epochs, batch, lr, seed, _dp(hist)))
    return params, hist


def cells_device(n, xyz_ptr, cell, corner, nx, ny, nz, start_ptr, order_ptr, device=0,
                 stream=None):
    """mccs_cells_device over device buffers (pointers): the linked-cell
    binning and stable reorder of n atoms on the GPU."""
    c = (C.c_double * 3)(*corner)
    _check(lib().mccs_cells_device(device, n, C.c_void_p(xyz_ptr), cell, c, nx, ny, nz,
                                   C.c_void_p(start_ptr), C.c_void_p(order_ptr),
                                   C.c_void_p(stream) if stream else None))


def inertia_device(n, xyz_ptr, m_ptr, device=0, stream=None):
    """mccs_inertia_device: (centre of mass, total mass, 3 x 3 inertia tensor)."""
    com = (C.c_double * 3)()
    mass = C.c_double()
    I = (C.c_double * 9)()
    _check(lib().mccs_inertia_device(device, n, C.c_void_p(xyz_ptr), C.c_void_p(m_ptr), com,
                                     C.byref(mass), I, C.c_void_p(stream) if stream else None))
    return np.array(com), mass.value, np.array(I).reshape(3, 3)


def tally(rec, n_iter, n_traj):
    rec = np.ascontiguousarray(rec, dtype=RECORD_DTYPE)
    om = np.zeros(n_iter)
    cnt = np.zeros(3 * n_iter, dtype=np.int64)
    mean, err = C.c_double(), C.c_double()
    _check(lib().mccs_tally(rec.ctypes.data_as(C.c_void_p), n_iter, n_traj, _dp(om),
                            cnt.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(mean),
                            C.byref(err)))
    return om, cnt.reshape(n_iter, 3), mean.value, err.value


def build(verbose=False):
    """Compile libmassccs_amd.so and the massccs CLI for gfx950 (in-tree)."""
    import subprocess
    r = subprocess.run(["make", "-C", os.path.join(HERE, "csrc"), "-j8"],
                       capture_output=not verbose, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + (r.stdout or "") + (r.stderr or ""))
