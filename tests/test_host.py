"""Host side of the product (no GPU needed): input parsing with the
reference's keys and error behaviour, target preparation bit-identical to the
reference (checksums written from the reference harness), the sample stream,
the tally, the report text, and the C ABI's exported symbols."""
import ctypes as C
import hashlib
import json
import os
import re
import subprocess

import numpy as np
import pytest

from common import GOLDEN, LONG, M, O, PUBLISHED, gpu_available, prepared, synth

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "massccs.h")
PREPS = sorted(f[len("prep_"):-len(".json")] for f in os.listdir(GOLDEN) if f.startswith("prep_"))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def test_c_abi_exports_every_declared_symbol():
    names = re.findall(r"^[a-z_ ]*?\*?\s*(mccs_\w+)\s*\(", open(HEADER).read(), re.M)
    assert len(names) >= 18
    L = M.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(M.EXPORTS)


@pytest.mark.parametrize("name", PREPS)
def test_preparation_bit_identical_to_reference(name):
    with open(os.path.join(GOLDEN, "prep_%s.json" % name)) as fh:
        g = json.load(fh)
    cfg = dict(g["config"])
    cfg["targetFileName"] = os.path.join(GOLDEN, cfg["targetFileName"])
    ref = g["prep"]
    P = prepared(cfg)
    i, a = P.info, P.atoms()
    for k in ("x", "y", "z", "q", "m", "eps", "sig"):
        assert digest(a[k]) == ref["sha256"][k], k
    assert i.force_type == ref["force_type"] and i.natoms == ref["natoms"]
    assert [i.a, i.b, i.c] == ref["abc"] and i.bmax == ref["bmax"]
    assert i.mu == ref["mu"] and i.mass == ref["mass"] and i.charge == ref["Q"]
    assert i.alpha == ref["alpha"] and list(i.half_box) == ref["halfbox"]
    assert [i.nx, i.ny, i.nz, i.ncells] == ref["grid"]
    # cells: (index, count, first atom) of every filled cell
    counts = np.diff(np.r_[0, np.cumsum(np.bincount(_cell_of(P), minlength=i.ncells))])
    filled = [[c, int(counts[c])] for c in range(i.ncells) if counts[c]]
    assert filled == [[c[0], c[1]] for c in ref["cells"]]


def _cell_of(P):
    i, a = P.info, P.atoms()
    ix = np.floor((a["x"] - i.corner[0]) / i.cell).astype(int)
    iy = np.floor((a["y"] - i.corner[1]) / i.cell).astype(int)
    iz = np.floor((a["z"] - i.corner[2]) / i.cell).astype(int)
    return iz + i.nz * iy + i.nz * i.ny * ix


def test_atoms_are_in_cell_order():
    P = prepared(PUBLISHED["ubq_He"])
    c = _cell_of(P)
    assert np.all(np.diff(c) >= 0)


def test_samples_match_oracle_stream():
    for key in ("c60_He", "c60_N2"):
        P = prepared(PUBLISHED[key])
        from common import oracle_for
        orc = oracle_for(P, PUBLISHED[key])
        u = P.samples(500)
        assert np.array_equal(u, orc.samples(500))
        # any window of the stream, as mt19937::discard provides it
        assert np.array_equal(P.samples(100, first=250), u[250:350])


def test_tally_matches_oracle():
    rng = np.random.default_rng(1)
    rec = np.zeros(3000, dtype=M.RECORD_DTYPE)
    rec["hit"] = rng.random(3000) < 0.9
    rec["success"] = rec["hit"] & (rng.random(3000) < 0.99)
    rec["domega"] = rec["success"] * rng.uniform(0, 5000, 3000)
    om, cnt, mean, err = M.tally(rec, 3, 1000)
    r2 = np.zeros(3000, dtype=O.RECORD_DTYPE)
    for k in ("hit", "success", "domega"):
        r2[k] = rec[k]
    om2, cnt2, mean2, err2 = O.tally(r2, 3, 1000)
    assert np.array_equal(om, om2) and np.array_equal(cnt, cnt2)
    assert mean == mean2 and err == err2


def test_report_matches_reference_log_header():
    """The report lines the reference prints before the trajectories, against
    its published C60/He log (ccs_He.log), for every line whose content the
    current reference code still prints the same way (the log's rotation-matrix
    lines come from an earlier version of the reference and differ from what
    its current code prints; the coordinates themselves are pinned by the
    checksums above)."""
    P = prepared(PUBLISHED["c60_He"])
    rep = P.report()
    for line in ["Inertia matrix a.u. Ang^2: ", "{  6010.37  -0.00132121  0.0799572  }",
                 "Molecule radius: 3.57355 Ang ",
                 "target mass: 720.66 amu", "reduce mass: 3.98049 amu",
                 "charge state: 0.999997 e", "maximal distances: 3.52258  3.51188  3.57355  Ang",
                 "Ellipsoid axis length: 53.5626  53.5519  53.6135  Ang",
                 "maximal impact parameter: 53.6135 Ang", "Nx: 9 Ny: 9 Nz: 9",
                 "Numbers of cells: 729", "Filled cells: 1"]:
        assert line in rep.splitlines(), line


def test_config_defaults_follow_reference_code(tmp_path):
    p = tmp_path / "in.json"
    p.write_text(json.dumps({"targetFileName": "x.xyz"}))
    c = M.config_from_json(str(p))
    assert (c.n_probe, c.n_iter, c.seed) == (10000, 10, 20162104)
    assert (c.dt, c.temperature, c.skin, c.lj_cutoff, c.coul_cutoff) == (10.0, 298.0, 0.01, 12.0, 25.0)
    assert c.gas == 1 and c.short_range_cutoff == 1 and c.long_range_flag == 0
    assert c.long_range_cutoff == 1 and c.polarizability == 1 and c.alpha == 0.204956
    p.write_text(json.dumps({"targetFileName": "x.xyz", "GasBuffer": "N2"}))
    c = M.config_from_json(str(p))
    assert c.polarizability == 0 and c.alpha_set == 0


@pytest.mark.parametrize("doc,code", [
    ('{"numberProbe": 10}', 1),                                  # targetFileName missing
    ('{"targetFileName": "a.xyz", "GasBuffer": "Xe"}', 1),       # unknown gas
    ('{"targetFileName": "a.xyz", "polarizability": "maybe"}', 1),
    ('{"targetFileName": "a.xyz", "numberProbe": -3}', 3),
    ('{"targetFileName": "a.xyz", ', 3),                         # malformed json
])
def test_config_errors_are_status_codes(tmp_path, doc, code):
    p = tmp_path / "in.json"
    p.write_text(doc)
    with pytest.raises(M.MassCCSError) as e:
        M.config_from_json(str(p))
    assert e.value.code == code


def test_target_file_errors(tmp_path):
    base = {"GasBuffer": "He"}
    with pytest.raises(M.MassCCSError) as e:
        prepared(dict(base, targetFileName=str(tmp_path / "missing.xyz")))
    assert e.value.code == 2
    bad = tmp_path / "bad.xyz"
    bad.write_text("3\ncomment\nC 0 0 0\nC 1 0 0\n")  # fewer lines than atoms
    with pytest.raises(M.MassCCSError) as e:
        prepared(dict(base, targetFileName=str(bad)))
    assert e.value.code == 3
    unk = tmp_path / "unk.xyz"
    unk.write_text("1\ncomment\nXe 0 0 0\n")  # not in the He force field
    with pytest.raises(M.MassCCSError) as e:
        prepared(dict(base, targetFileName=str(unk)))
    assert e.value.code == 1
    with pytest.raises(M.MassCCSError) as e:
        prepared(dict(base, targetFileName=str(tmp_path / "t.pdb")))
    assert e.value.code in (1, 2)


def test_prepare_from_memory_equals_file(tmp_path):
    s, xyz, q = synth.globule(n=300, axes=(9.0, 8.0, 7.0), net_charge=2.0, seed=5)
    path = str(tmp_path / "g.xyz")
    synth.write_xyz(path, s, xyz, q)
    cfg = dict(targetFileName=path, GasBuffer="N2", polarizability="yes", **LONG)
    P1 = prepared(cfg)
    back = np.loadtxt(path, skiprows=2, usecols=(1, 2, 3, 4))
    P2 = M.Prepared(M.config_from_dict(cfg), symbols=list(s), xyz=back[:, :3], q=back[:, 3])
    for k, v in P1.atoms().items():
        assert np.array_equal(v, P2.atoms()[k]), k
    assert P1.info.bmax == P2.info.bmax
    # symbols as a str array (packed without a Python loop) and as bytes
    # (encoded one by one) give the same target
    for sym in (np.asarray(s, dtype=str), [x.encode() for x in s]):
        P3 = M.Prepared(M.config_from_dict(cfg), symbols=sym, xyz=back[:, :3], q=back[:, 3])
        for k, v in P1.atoms().items():
            assert np.array_equal(v, P3.atoms()[k]), k
    with pytest.raises(ValueError):
        M.Prepared(M.config_from_dict(cfg), symbols=list(s)[:-1], xyz=back[:, :3], q=back[:, 3])


def test_synthetic_generators():
    from scipy.spatial.distance import pdist
    d = np.sort(pdist(synth.c60()))
    assert np.allclose(d[:30], 1.40, atol=1e-9) and np.allclose(d[30:90], 1.45, atol=1e-9)
    assert d[90] > 2.0
    s, xyz, q = synth.globule()
    assert len(xyz) == 1231 and pdist(xyz).min() >= 1.0
    assert abs(q.sum() - 4.0) < 1e-9 and np.all((xyz / [18, 15, 13]) ** 2 @ np.ones(3) <= 1 + 1e-12)
    fr = {e: np.mean(s == e) for e in synth.ELEMENTS}
    for e, f in zip(synth.ELEMENTS, synth.GLOBULE_FRACTIONS):
        assert abs(fr[e] - f) < 3.0 / 1231  # rounding remainder goes to the largest class


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU failure path")
def test_engine_fails_loudly_without_gpu():
    P = prepared(PUBLISHED["c60_He"])
    with pytest.raises(M.MassCCSError) as e:
        M.Engine(P)
    assert e.value.code == 5


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU failure path")
def test_cli_reports_missing_gpu(tmp_path):
    cfg = dict(PUBLISHED["c60_He"], numberProbe=10, nIter=1)
    p = tmp_path / "in.json"
    p.write_text(json.dumps(cfg))
    r = subprocess.run([M.CLI_PATH, str(p)], capture_output=True, text=True)
    assert r.returncode != 0 and "GPU" in r.stderr
    assert "INPUT:: Simulation Parameters" in r.stdout


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU failure path")
def test_equipotential_needs_the_gpu():
    cfg = dict(PUBLISHED["c60_He"], Equipotential="yes")
    with pytest.raises(M.MassCCSError) as e:
        prepared(cfg)
    assert e.value.code == 5
