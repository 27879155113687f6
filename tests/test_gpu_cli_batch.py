"""massccs --batch: several inputs in one batched GPU run, reported as the
reference data set's table; each CCS equals a single-target run of the same
launch shape, and the He rows equal the oracle's tally of the same samples."""
import csv
import json
import subprocess

import numpy as np
import pytest

from common import LONG, M, O, C60_XYZ, UBQ_PQR, gpu_available, oracle_for, prepared

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def test_cli_batch_table(tmp_path):
    print("case cli batch", flush=True)
    base = dict(polarizability="yes", seed=2104, numberProbe=1500, nIter=2, **LONG)
    cfgs = {"ubq_He.json": dict(base, targetFileName=UBQ_PQR, GasBuffer="He"),
            "ubq_N2.json": dict(base, targetFileName=UBQ_PQR, GasBuffer="N2", numberProbe=400),
            "c60_He.json": dict(base, targetFileName=C60_XYZ, GasBuffer="He")}
    for name, c in cfgs.items():
        (tmp_path / name).write_text(json.dumps(c))
    lst = tmp_path / "inputs.txt"
    lst.write_text("# three runs, two targets\n" + "\n".join(str(tmp_path / k) for k in cfgs) + "\n")
    out = tmp_path / "table.csv"
    r = subprocess.run([M.CLI_PATH, "--batch", str(lst), str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    rows = list(csv.DictReader(open(out)))
    assert [x["id"] for x in rows] == ["c60", "04_1ubq_charge_4e_min"]  # sorted by mass
    assert rows[0]["ccs_N2"] == "" and rows[1]["ccs_N2"] != ""
    for row, key, gas in [(rows[0], "c60_He.json", "He"), (rows[1], "ubq_He.json", "He"),
                          (rows[1], "ubq_N2.json", "N2")]:
        c = cfgs[key]
        P = prepared(c)
        rec = M.Engine(P, lanes=64, use_lds=0).run(P.samples(c["numberProbe"] * c["nIter"]))
        _, _, mean, err = M.tally(rec, c["nIter"], c["numberProbe"])
        got = float(row["ccs_" + gas])
        print(key, got, mean, flush=True)
        assert abs(got / mean - 1) <= 1e-5
        assert abs(float(row["err_" + gas]) / err - 1) <= 1e-4
    assert int(rows[1]["natoms"]) == prepared(cfgs["ubq_He.json"]).info.natoms
    # the He rows against the oracle's tally of the same samples (every He
    # trajectory of these targets agrees with the oracle: fates identical, chi
    # within 1e-9 rad), to the table's printed precision
    for row, key in [(rows[0], "c60_He.json"), (rows[1], "ubq_He.json")]:
        c = cfgs[key]
        P = prepared(c)
        ref = oracle_for(P, c).run(P.samples(c["numberProbe"] * c["nIter"]))
        _, _, mean, err = O.tally(ref, c["nIter"], c["numberProbe"])
        print(key, "table", row["ccs_He"], "oracle", mean, err, flush=True)
        assert abs(float(row["ccs_He"]) / mean - 1) <= 1e-5
        assert abs(float(row["err_He"]) / err - 1) <= 1e-4
