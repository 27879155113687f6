"""Batched runs (several targets, one persistent launch per probe kind and
force type, replayed from a hipGraph) against per-target runs and the oracle."""
import numpy as np
import pytest

from common import LONG, M, C60_XYZ, UBQ_PQR, globule_cfg, gpu_available, oracle_for, prepared

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _targets(workdir):
    cfgs = [
        dict(targetFileName=C60_XYZ, GasBuffer="He", polarizability="yes", seed=2104, **LONG),
        dict(targetFileName=UBQ_PQR, GasBuffer="He", polarizability="yes", seed=7, **LONG),
        globule_cfg(workdir),
        dict(targetFileName=UBQ_PQR, GasBuffer="N2", polarizability="yes", seed=11, **LONG),
    ]
    return cfgs, [prepared(c) for c in cfgs]


def test_batch_equals_per_target_runs(workdir):
    print("case batch vs engines", flush=True)
    cfgs, P = _targets(workdir)
    ns = [1500, 600, 400, 300]
    us = [p.samples(n) for p, n in zip(P, ns)]
    B = M.Batch(P, lanes=64, skin=1.0)
    recs = B.run(us)
    for p, u, r in zip(P, us, recs):
        ref = M.Engine(p, lanes=64, use_lds=0).run(u)
        assert np.array_equal(r.view(np.uint8), ref.view(np.uint8))
    # the graph replays give the same bytes, with other device work in between
    for _ in range(2):
        M.Engine(P[0], lanes=16).run(us[0][:100])
        again = B.run(us)
        for a, b in zip(recs, again):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    w = B.last_work()
    assert w["steps"] >= sum(int(r["steps"].sum()) for r in recs)


def test_batch_autotune_and_parity(workdir):
    print("case batch autotune", flush=True)
    cfgs, P = _targets(workdir)
    ns = [2000, 500, 300, 200]
    us = [p.samples(n) for p, n in zip(P, ns)]
    B = M.Batch(P)
    rep = B.autotune([u[: max(64, len(u) // 4)] for u in us])
    assert len(rep) == 2  # He type 4 and N2 type 6
    for c in rep:
        # lanes x skin x LDS tile (half or whole CU per block)
        assert len(c["trials"]) == 18 and c["lanes"] in (16, 32, 64) and c["skin"] in (0.5, 1.0, 2.0)
        assert c["tile"] in (1, 2) and {t["tile"] for t in c["trials"]} == {1, 2}
        best = min(t["ms"] for t in c["trials"])
        assert any(t["ms"] == best and t["lanes"] == c["lanes"] and t["skin"] == c["skin"]
                   and t["tile"] == c["tile"] for t in c["trials"])
    lanes, skin = B.shapes()
    assert set(lanes) <= {16, 32, 64}
    recs = B.run(us)
    # C60/He against the oracle on the same samples
    ref = oracle_for(P[0], cfgs[0]).run(us[0])
    assert np.array_equal(ref["hit"], recs[0]["hit"])
    same = ref["tries"] == recs[0]["tries"]
    assert same.mean() >= 0.99
    big = (ref["chi"] > 1e-6) & same & (ref["success"] == 1)
    dchi = np.abs(ref["chi"] - recs[0]["chi"])
    assert (dchi[big] <= 1e-8).mean() >= 0.99
    # a second batch of the same targets starts from the cached shapes
    B2 = M.Batch(P)
    assert B2.shapes() == (lanes, skin)
