"""Target preparation steps on the GPU (mccs_cells_device,
mccs_inertia_device) against the host preparation, which stays
authoritative: on the 10^6-atom shell the linked-cell binning and stable
reorder give the same integers as the reference's counting sort
(src/LinkedCell.cpp:93-225, restated with numpy's stable argsort), and the
inertia tensor agrees to rounding."""
import time

import numpy as np
import pytest

from common import M, gpu_available, synth

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

CFG = {"targetFileName": "memory.xyz", "GasBuffer": "N2", "Short-range cutoff": "yes",
       "Long-range forces": "yes", "Long-range cutoff": "yes", "polarizability": "yes",
       "alpha": 1.7403, "seed": 2104}


def _cells_numpy(xyz, cell, corner, nx, ny, nz):
    ijk = np.floor((xyz - np.asarray(corner)) / cell).astype(np.int64)
    cid = ijk[:, 2] + nz * ijk[:, 1] + nz * ny * ijk[:, 0]
    order = np.argsort(cid, kind="stable")
    start = np.searchsorted(cid[order], np.arange(nx * ny * nz + 1), side="left")
    return order.astype(np.int32), start.astype(np.int32)


def test_cells_and_inertia_on_the_shell():
    import torch
    print("case device preparation, 10^6-atom shell", flush=True)
    s, xyz, q = synth.shell()
    t0 = time.perf_counter()
    P = M.Prepared(M.config_from_dict(CFG), symbols=list(s), xyz=xyz, q=q)
    t_host = time.perf_counter() - t0
    i = P.info
    a = P.atoms()
    X = np.stack([a["x"], a["y"], a["z"]], 1)
    n = len(X)
    grid = (i.cell, list(i.corner), i.nx, i.ny, i.nz)
    rng = np.random.default_rng(4)
    perm = rng.permutation(n)
    start = torch.zeros(i.ncells + 1, dtype=torch.int32, device="cuda")
    order = torch.zeros(n, dtype=torch.int32, device="cuda")
    for label, src in [("host order", X), ("shuffled", X[perm])]:
        d = torch.from_numpy(np.ascontiguousarray(src)).cuda()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        M.cells_device(n, d.data_ptr(), *grid, start.data_ptr(), order.data_ptr())
        t_dev = time.perf_counter() - t0
        o_ref, s_ref = _cells_numpy(src, *grid)
        assert np.array_equal(order.cpu().numpy(), o_ref), label
        assert np.array_equal(start.cpu().numpy(), s_ref), label
        if label == "host order":
            # the host's own reorder is what a stable sort of its output keeps
            assert np.array_equal(o_ref, np.arange(n, dtype=np.int32))
        print("%s: %d atoms in %d cells sorted on the GPU in %.4f s (host preparation %.2f s)"
              % (label, n, i.ncells, t_dev, t_host), flush=True)
    dm = torch.from_numpy(np.ascontiguousarray(a["m"])).cuda()
    dX = torch.from_numpy(np.ascontiguousarray(X)).cuda()
    com, mass, I = M.inertia_device(n, dX.data_ptr(), dm.data_ptr())
    m = a["m"]
    com_ref = (m[:, None] * X).sum(0) / m.sum()
    R = X - com_ref
    r2 = np.sum(R * R, 1)
    I_ref = np.einsum("k,kab->ab", m, r2[:, None, None] * np.eye(3)[None] - R[:, :, None] * R[:, None, :])
    assert abs(mass - m.sum()) <= 1e-12 * m.sum()
    assert np.all(np.abs(com - com_ref) <= 1e-9 * np.max(np.abs(X)))
    assert np.all(np.abs(I - I_ref) <= 1e-11 * np.max(np.abs(I_ref)))


def test_cells_reject_an_atom_outside_the_grid():
    import torch
    X = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 3.0], [500.0, 0.0, 0.0]])
    d = torch.from_numpy(X).cuda()
    start = torch.zeros(2 * 2 * 2 + 1, dtype=torch.int32, device="cuda")
    order = torch.zeros(3, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(M.MassCCSError) as e:
        M.cells_device(3, d.data_ptr(), 12.01, [-12.01, -12.01, -12.01], 2, 2, 2, start.data_ptr(),
                       order.data_ptr())
    assert e.value.code == 1
