"""The CLI on the GPU: 3-line stdout, gzip CSV frontier, JSONL schema v1
(src/pmdd/util/output_utils.cpp:87-374 of the reference)."""
import gzip
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import instance_path
from pmdd_amd import DecisionDiagram, abi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,ptype,method,sd", [
    ("KP_p-3_n-20_ins-1.dat", "1", "1", "1"),
    ("KP_p-3_n-20_ins-1.dat", "1", "3", "0"),
    ("bp-100-20-3-10-10192.dat", "2", "3", "0"),
    ("bp-100-20-3-10-10192.dat", "2", "2", "0"),
    ("tsp-nobj3-ncities10-seed12870.dat", "3", "1", "0"),
])
def test_cli_outputs(tmp_path, name, ptype, method, sd):
    fr = tmp_path / "f.csv.gz"
    stats = tmp_path / "s.jsonl"
    r = subprocess.run([abi.CLI_BIN, instance_path(name), ptype, method, sd, "--backend", "gpu",
                        "--frontier-out", str(fr), "--stats-out", str(stats)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().split("\n")
    assert len(lines) == 3
    dd = DecisionDiagram.from_file(instance_path(name), int(ptype))
    ref = oracle.enumerate_dd(dd, 1, 0, threads=4)
    assert int(lines[0]) == len(ref)
    fields = lines[2].split("\t")
    assert len(fields) == (4 if ptype == "3" else 13)
    if ptype != "3":
        assert fields[0] == method and fields[1] == sd
    # frontier CSV: lexicographically ascending rows "v0,...,v{d-1}"
    rows = gzip.decompress(fr.read_bytes()).decode().strip().split("\n")
    got = np.array([[int(x) for x in row.split(",")] for row in rows], dtype=np.int32)
    assert got.shape == ref.shape and (got == ref).all()
    rec = json.loads(stats.read_text().strip().split("\n")[-1])
    assert rec["schema_version"] == 1
    assert rec["identity"]["backend"] == "gpu" and rec["identity"]["method"] == int(method)
    assert rec["outputs"]["num_solutions"] == len(ref)
    for key in ("timing", "memory", "work", "dominance", "structure", "metrics", "status"):
        assert key in rec
    assert rec["status"]["status_state"] == "ok"
    # the kernel times are filled, as the reference fills them whenever it
    # is given stats (its total is expansion plus dominance time)
    k = rec["timing"]["kernel"]
    assert k["kernel_total_s"] > 0
    assert abs(k["kernel_total_s"] - (k["kernel_expand_td_s"] + k["kernel_dominance_s"])) < 3e-6
    assert rec["structure"]["is_tsp_branch"] == (ptype == "3")
    if method == "3":
        assert rec["structure"]["layer_coupling"] > 0
