"""Parity of the limited-visibility device solver (CR_VARIANT_MACRO) with the
reference's solvers/k_cops_visibility_v2.cpp and with the oracle.

Expected values: the compiled reference's v2 cache blobs and stdout facts
(tests/golden/ref*), its recorded CSV (console_outputs/test_results_v2.csv),
and the oracle's memoised trajectory search on seeded random graphs. Integer
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

V2 = sorted(k for k, v in INDEX.items() if v["solver"] == "v2")


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def solver():
    s = cra.Solver(0)
    yield s
    s.close()


def run_macro(solver, k, p, path=None, adj=None):
    if path is not None:
        solver.load_graph_file(path, k, p, cra.VARIANT_MACRO)
    else:
        solver.load_graph_matrix(adj, k, p, cra.VARIANT_MACRO)
    solver.build_aux_graph()
    base = solver.initialize_captures()
    passes = solver.main_loop()
    verdict = solver.find_final_result()
    return base, passes, verdict


def positions(v):
    return ", ".join(str(x) for x in v["positions"])


@pytest.mark.parametrize("name", V2)
def test_macro_tables_match_reference(solver, name):
    e = INDEX[name]
    say("v2 case", name)
    base, passes, verdict = run_macro(solver, e["k"], e["p"], path=matrix(e["graph"]))
    states = solver.states()
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
    m = solver.macro_stats()
    assert m["steps"] == e["p"] and m["nodes_per_step"][0] == states.size


@pytest.mark.parametrize("name", ["v2__grid9__k2__p2", "v2__peterson__k3__p3", "v2__outerplanar32__k2__p3"])
def test_macro_cache_blob_matches_reference_bytes(solver, name, tmp_path):
    e = INDEX[name]
    say("v2 blob", name)
    run_macro(solver, e["k"], e["p"], path=matrix(e["graph"]))
    out = tmp_path / "blob.bin"
    solver.output_data(out)
    data = out.read_bytes()
    assert len(data) == e["blob_bytes"]
    assert hashlib.sha256(data).hexdigest() == e["blob_sha256"]


def csv_rows(fn):
    with open(os.path.join(GOLDEN, "csv", fn)) as f:
        return [r for r in csv.DictReader(f) if r["status"] == "ok"]


ROWS = csv_rows("test_results_v2.csv")
FAST = [r for r in ROWS if float(r["mainloop_s"] or 0) <= 5]
SLOW = [r for r in ROWS if float(r["mainloop_s"] or 0) > 5]


def check_row(solver, row):
    base, passes, verdict = run_macro(solver, int(row["k"]), int(row["p"]), path=matrix(row["graph"]))
    assert str(base) == row["base_caps"]
    assert str(len(passes)) == row["passes"]
    if row["result"] == "WIN":
        assert verdict is not None
        assert positions(verdict) == row["positions"]
        assert str(verdict["plys"]) == row["capture_plys"]
    else:
        assert verdict is None


def test_recorded_csv_rows(solver):
    for row in FAST:
        say("v2 csv", row["graph"], row["k"], row["p"])
        check_row(solver, row)


@pytest.mark.parametrize("row", SLOW, ids=[f"{r['graph']}-k{r['k']}-p{r['p']}" for r in SLOW])
def test_recorded_csv_rows_slow_in_reference(solver, row):
    """Rows the reference needed 5-204 s for."""
    say("v2 csv (slow in the reference)", row["graph"], row["k"], row["p"])
    check_row(solver, row)


def random_cases():
    """Seeded graphs: undirected ER graphs over all four row widths, one with
    self-loops (stay moves) and one directed graph (every vertex keeps an
    out-edge)."""
    cases = []
    for n, p_edge, seed, k, p in ((7, 0.4, 3, 2, 2), (12, 0.25, 5, 2, 3), (16, 0.2, 8, 3, 2), (20, 0.15, 2, 2, 4),
                                  (30, 0.1, 4, 2, 3), (70, 0.06, 6, 1, 3), (70, 0.05, 9, 2, 2),
                                  (140, 0.03, 1, 1, 2), (200, 0.02, 3, 1, 3)):
        adj, _ = workloads.er_graph(n, p_edge, seed)
        cases.append((f"er{n}_s{seed}", adj, k, p))
    loops, _ = workloads.er_graph(14, 0.2, 11)
    loops = loops.copy()
    loops[np.arange(14), np.arange(14)] = 1
    cases.append(("er14_loops", loops, 2, 3))
    rng = workloads.SplitMix64(21)
    n = 15
    di = np.zeros((n, n), np.uint8)
    for u in range(n):
        di[u, (u + 1) % n] = 1
        for _ in range(2):
            v = rng.below(n)
            if v != u:
                di[u, v] = 1
    cases.append(("directed15", di, 2, 2))
    return cases


def test_random_graphs_match_oracle(solver, oracle):
    for name, adj, k, p in random_cases():
        say("v2 random", name, k, p)
        base, passes, verdict = run_macro(solver, k, p, adj=adj)
        want = oracle.v2_solve(adj, k, p)
        assert np.array_equal(solver.states(), want.states), name
        assert passes == want.pass_counts, name
        assert base == want.base_caps, name


def test_shapes_in_sequence_reuse_buffers(solver):
    """One solver, several shapes in a row (buffers grow and are reused), then
    the first shape again: identical tables."""
    say("v2 sequence")
    seq = [("grid36", 2, 3), ("peterson", 3, 1), ("planar32", 2, 2), ("grid36", 2, 3)]
    tables = []
    for graph, k, p in seq:
        run_macro(solver, k, p, path=matrix(graph))
        tables.append(solver.states())
    assert np.array_equal(tables[0], tables[-1])
    solver.initialize_captures()
    solver.main_loop()
    assert np.array_equal(solver.states(), tables[-1])


def test_cli_v2_reproduces_recorded_rows(tmp_path):
    exe = os.path.join(REPO, "copsandrobbers_amd", "bin", "k_cops_visibility_v2")
    from copsandrobbers_amd import harness
    for row in FAST[::6]:
        say("v2 cli", row["graph"], row["k"], row["p"])
        res = harness.run_solver(exe, matrix(row["graph"]), int(row["k"]), int(row["p"]), timeout=120,
                              cwd=str(tmp_path))
        assert res["status"] == "ok", res
        for col in ("result", "capture_plys", "passes", "base_caps", "positions"):
            assert str(res[col]) == row[col], (row["graph"], col)
