"""Equipotential ellipsoid on the GPU against the reference's own axes
(tests/golden/equi_*.json, written by the reference harness with its OpenMP
reductions on one thread)."""
import json
import os

import numpy as np
import pytest

from common import GOLDEN, gpu_available, prepared

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

CASES = sorted(f[len("equi_"):-len(".json")] for f in os.listdir(GOLDEN) if f.startswith("equi_"))


@pytest.mark.parametrize("name", CASES)
def test_equipotential_axes_match_reference(name):
    print("case equipotential", name, flush=True)
    with open(os.path.join(GOLDEN, "equi_%s.json" % name)) as fh:
        g = json.load(fh)
    cfg = dict(g["config"])
    cfg["targetFileName"] = os.path.join(GOLDEN, cfg["targetFileName"])
    P = prepared(cfg)
    i = P.info
    abc = np.array([i.a, i.b, i.c])
    ref = np.array(g["abc"])
    print(name, abc.tolist(), ref.tolist(), flush=True)
    # the potentials are summed in another order than the reference's, which
    # could move a crossing by one 0.1 A ray step; on these targets every
    # crossing agrees and the axes differ only in the fit's last bits
    assert np.all(np.abs(abc - ref) <= 1e-12 * ref)
    assert [i.nx, i.ny, i.nz, i.ncells] == g["grid"]
    rep = P.report()
    assert "Equipotential Ellipsoid: " in rep.splitlines()
    assert any(ln.startswith("ellipsoid axis length: ") for ln in rep.splitlines())
