"""The surrogate network's GPU kernels (BF16 MFMA, FP32 accumulation) against
the CPU oracle, its training, and the surrogate in trajectories: CCS against
the exact method on the same samples, and what the network gains."""
import os
import sys
import time

import numpy as np
import pytest

from common import M, globule_cfg, gpu_available, prepared

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
    xyz = rng.uniform(-30, 30, (4096, 3))
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


def _ccs_pair(E, u):
    E.use_surrogate(False)
    E.run(u[:2000])
    t0 = time.time()
    plain = E.run(u)
    t_plain = time.time() - t0
    E.use_surrogate(True)
    E.run(u[:2000])
    t0 = time.time()
    sur = E.run(u)
    t_sur = time.time() - t0
    E.use_surrogate(False)
    _, _, c_plain, _ = M.tally(plain, 1, len(u))
    _, _, c_sur, _ = M.tally(sur, 1, len(u))
    print("ccs plain %.3f surrogate %.3f (%.3f %%); time %.3f s -> %.3f s; visits %.3g -> %.3g; "
          "lost %d -> %d" % (c_plain, c_sur, 100 * (c_sur / c_plain - 1), t_plain, t_sur,
                             plain["visits"].sum(), sur["visits"].sum(),
                             ((plain["hit"] == 1) & (plain["success"] == 0)).sum(),
                             ((sur["hit"] == 1) & (sur["success"] == 0)).sum()), flush=True)
    return plain, sur, c_plain, c_sur


def test_tabulated_far_field_path_matches_exact(workdir):
    """The trajectory kernel's tabulated far-field path (B-spline grid, blend
    band) fed with the exact evaluator's values: CCS agrees with the exact
    method, so what the network adds is its own fitting error only."""
    print("case tabulated exact far field", flush=True)
    P = prepared(globule_cfg(workdir))
    E = M.Engine(P)
    E.build_surrogate(exact_table=True, n_train=100000, epochs=1, grid=0.25)
    u = P.samples(20000)
    plain, sur, c_plain, c_sur = _ccs_pair(E, u)
    assert abs(c_sur / c_plain - 1) <= 0.005
    assert sur["visits"].sum() < plain["visits"].sum()
    again = E.run(u[:3000])
    assert np.array_equal(again.view(np.uint8), plain[:3000].view(np.uint8))


def test_network_surrogate_in_trajectories(workdir):
    """The trained network in place of the exact far field: its held-out
    errors, and CCS and cost against the exact method on the same samples."""
    print("case network surrogate globule N2", flush=True)
    P = prepared(globule_cfg(workdir))
    E = M.Engine(P)
    rep, params, spec = E.build_surrogate()
    print(rep, flush=True)
    for r in range(2):
        assert rep["val_rmse_u"][r] <= 0.5 * rep["val_rms_u"][r]
        assert rep["val_rmse_f"][r] <= 0.6 * rep["val_rms_f"][r]
    u = P.samples(20000)
    plain, sur, c_plain, c_sur = _ccs_pair(E, u)
    assert np.array_equal(plain["hit"], sur["hit"])
    assert sur["visits"].sum() < plain["visits"].sum()
