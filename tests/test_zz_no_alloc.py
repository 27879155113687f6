"""mccs_run_device_noalloc (the run entry point that allocates nothing: its
work queue lives in the library's static device workspace) against
mccs_run_device on the same engine and samples, through the C ABI: records
byte-identical on two calls in a row, the second with a smaller input, and
sizes beyond the workspace rejected."""
import ctypes as C

import numpy as np
import pytest

from common import C60_XYZ, LONG, M, globule_cfg, gpu_available, prepared

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _run(E, u, noalloc):
    import torch
    du = torch.from_numpy(np.ascontiguousarray(u)).cuda()
    rec = torch.zeros((len(u), M.RECORD_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    L = M.lib()
    fn = L.mccs_run_device_noalloc if noalloc else L.mccs_run_device
    M._check(fn(E.h, C.c_void_p(du.data_ptr()), len(u), C.c_void_p(rec.data_ptr()), None))
    torch.cuda.synchronize()
    return rec.cpu().numpy().reshape(-1).view(M.RECORD_DTYPE)


@pytest.mark.parametrize("which", ["globule_N2", "c60_He"])
def test_noalloc_run_equals_run_device(workdir, which):
    if which == "globule_N2":
        cfg, n1, n2 = globule_cfg(workdir), 6000, 1500
    else:
        cfg = dict(targetFileName=C60_XYZ, GasBuffer="He", polarizability="yes", seed=2104,
                   **LONG)
        n1, n2 = M.NOALLOC_MAX_TRAJECTORIES, 3000
    print("case noalloc", which, n1, n2, flush=True)
    P = prepared(cfg)
    E = M.Engine(P)
    u = P.samples(n1)
    for n in (n1, n2):  # the second call runs fewer trajectories over the same workspace
        ref = _run(E, u[:n], False)
        got = _run(E, u[:n], True)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), (which, n)
        w = E.last_work()  # the no-allocation run's own executed work
        assert w["steps"] >= int(got["steps"].sum()) > 0
        print(which, n, "records identical; steps", int(got["steps"].sum()), flush=True)


def test_noalloc_rejects_more_than_its_workspace():
    import torch
    cfg = dict(targetFileName=C60_XYZ, GasBuffer="He", polarizability="yes", seed=2104, **LONG)
    E = M.Engine(prepared(cfg))
    n = M.NOALLOC_MAX_TRAJECTORIES + 1
    du = torch.zeros(9 * n, dtype=torch.float64, device="cuda")
    rec = torch.zeros((n, M.RECORD_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    st = M.lib().mccs_run_device_noalloc(E.h, C.c_void_p(du.data_ptr()), n,
                                         C.c_void_p(rec.data_ptr()), None)
    assert st == 1  # MCCS_ERR_INVALID, nothing launched
    assert b"at most 65536" in M.lib().mccs_last_error()
