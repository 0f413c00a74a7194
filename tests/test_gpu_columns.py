"""Parity of the column solver with robber visibility p >= 2 (the reference's
solvers/k_cops_visibility.cpp with 2p columns) on the device.

Expected values: the compiled reference's cache blobs and stdout facts
(tests/golden/ref*), its recorded CSV rows (console_outputs/test_results.csv,
p = 2 and 3), and the oracle's pass sweep on seeded random graphs. Integer
work: every comparison is exact.
"""
import csv
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, REPO, matrix

import copsandrobbers_amd as cra
from copsandrobbers_amd import workloads

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "ref_index.json")) as f:
    INDEX = json.load(f)

PN = sorted(k for k, v in INDEX.items() if v["solver"] == "v1" and v["p"] >= 2)


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def solver():
    s = cra.Solver(0)
    yield s
    s.close()


def run(solver, k, p, path=None, adj=None):
    if path is not None:
        solver.load_graph_file(path, k, p)
    else:
        solver.load_graph_matrix(adj, k, p)
    solver.build_aux_graph()
    base = solver.initialize_captures()
    passes = solver.main_loop()
    verdict = solver.find_final_result()
    return base, passes, verdict


def positions(v):
    return ", ".join(str(x) for x in v["positions"])


@pytest.mark.parametrize("name", PN)
def test_column_tables_match_reference(solver, name):
    e = INDEX[name]
    say("columns case", name)
    base, passes, verdict = run(solver, e["k"], e["p"], path=matrix(e["graph"]))
    states = solver.states()
    assert states.size == solver.info().num_states
    if e["arrays"]:
        g = np.load(os.path.join(GOLDEN, "ref", e["arrays"]))
        assert np.array_equal(states, g["states"])
    assert hashlib.sha256(states.tobytes()).hexdigest() == e["states_sha256"]
    assert passes == e["passes"]
    assert base == e["base_caps"]
    assert solver.count_transitions() == e["edge_pointers"]
    if e["result"] == "WIN":
        assert verdict is not None
        assert positions(verdict) == e["positions"]
        assert verdict["plys"] == e["plys"]
    else:
        assert verdict is None


@pytest.mark.parametrize("name", ["v1__grid9__k2__p2", "v1__peterson__k3__p3"])
def test_column_cache_blob_matches_reference_bytes(solver, name, tmp_path):
    e = INDEX[name]
    say("columns blob", name)
    run(solver, e["k"], e["p"], path=matrix(e["graph"]))
    out = tmp_path / "blob.bin"
    solver.output_data(out)
    data = out.read_bytes()
    assert len(data) == e["blob_bytes"]
    assert hashlib.sha256(data).hexdigest() == e["blob_sha256"]


def csv_rows():
    with open(os.path.join(GOLDEN, "csv", "test_results.csv")) as f:
        return [r for r in csv.DictReader(f) if r["status"] == "ok" and r["p"] != "1"]


ROWS = csv_rows()


def test_recorded_csv_rows(solver):
    for row in ROWS:
        say("columns csv", row["graph"], row["k"], row["p"])
        base, passes, verdict = run(solver, int(row["k"]), int(row["p"]), path=matrix(row["graph"]))
        assert str(base) == row["base_caps"]
        assert str(len(passes)) == row["passes"]
        if row["result"] == "WIN":
            assert verdict is not None
            assert positions(verdict) == row["positions"]
            assert str(verdict["plys"]) == row["capture_plys"]
        else:
            assert verdict is None


def test_random_graphs_match_oracle(solver, oracle):
    rng = workloads.SplitMix64(41)
    cases = []
    for n, pe, seed, k, p in ((9, 0.35, 1, 2, 2), (14, 0.2, 2, 2, 3), (18, 0.15, 3, 3, 2), (25, 0.1, 4, 2, 4),
                              (70, 0.05, 5, 1, 3), (130, 0.03, 6, 1, 2), (12, 0.3, 7, 4, 2)):
        adj, _ = workloads.er_graph(n, pe, seed)
        cases.append((f"er{n}", adj, k, p))
    loops, _ = workloads.er_graph(10, 0.3, 8)
    loops = loops.copy()
    loops[np.arange(10), np.arange(10)] = 1
    cases.append(("loops10", loops, 2, 3))
    n = 12
    di = np.zeros((n, n), np.uint8)
    for u in range(n):
        di[u, (u + 1) % n] = 1
        v = rng.below(n)
        if v != u:
            di[u, v] = 1
    cases.append(("directed12", di, 2, 2))
    for name, adj, k, p in cases:
        say("columns random", name, k, p)
        base, passes, verdict = run(solver, k, p, adj=adj)
        want = oracle.v1_solve(adj, k, p)
        assert np.array_equal(solver.states(), want.states), name
        assert passes == want.pass_counts, name
        assert base == want.base_caps, name


def test_cli_reproduces_recorded_rows_p2_p3(tmp_path):
    exe = os.path.join(REPO, "copsandrobbers_amd", "bin", "k_cops_visibility")
    from copsandrobbers_amd import harness
    for row in ROWS[::5]:
        say("columns cli", row["graph"], row["k"], row["p"])
        res = harness.run_solver(exe, matrix(row["graph"]), int(row["k"]), int(row["p"]), timeout=120,
                              cwd=str(tmp_path))
        assert res["status"] == "ok", res
        for col in ("result", "capture_plys", "passes", "base_caps", "positions"):
            assert str(res[col]) == row[col], (row["graph"], col)
