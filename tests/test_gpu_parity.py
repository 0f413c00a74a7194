"""Parity of the HIP solver (through the C ABI) with the reference and the oracle.

Expected values: the compiled reference's cache blobs and stdout facts
(tests/golden/ref*), the reference's recorded CSV rows, and the CPU oracle on
the same inputs. Integer work: every comparison is exact.
"""
import csv
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, REPO, matrix

import copsandrobbers_amd as cra

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "ref_index.json")) as f:
    INDEX = json.load(f)

P1 = sorted(k for k, v in INDEX.items() if v["solver"] == "v1" and v["p"] == 1)


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def solver():
    s = cra.Solver(0)
    yield s
    s.close()


def run_steps(solver, path, k, p=1, variant=cra.VARIANT_COLUMNS):
    solver.load_graph_file(path, k, p, variant)
    solver.build_aux_graph()
    base = solver.initialize_captures()
    passes = solver.main_loop()
    verdict = solver.find_final_result()
    return base, passes, verdict


def positions(v):
    return ", ".join(str(x) for x in v["positions"])


@pytest.mark.parametrize("name", P1)
def test_plain_tables_match_reference(solver, name):
    e = INDEX[name]
    say("case", name)
    base, passes, verdict = run_steps(solver, matrix(e["graph"]), e["k"])
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


@pytest.mark.parametrize("name", ["v1__peterson__k3__p1", "v1__grid36__k2__p1", "v1__test__k2__p1",
                                  "v1__line100__k1__p1"])
def test_cache_blob_matches_reference_bytes(solver, name, tmp_path):
    e = INDEX[name]
    say("blob", name)
    run_steps(solver, matrix(e["graph"]), e["k"])
    out = tmp_path / "blob.bin"
    solver.output_data(out)
    data = out.read_bytes()
    assert len(data) == e["blob_bytes"]
    assert hashlib.sha256(data).hexdigest() == e["blob_sha256"]


def test_levels_and_retro_edges_match_oracle(solver, oracle):
    for graph, k in (("grid36", 2), ("planar32", 3), ("dodecahedron", 3), ("scotlandyard-all", 2)):
        say("retro", graph, k)
        n, adj = oracle.load_matrix(matrix(graph))
        run_steps(solver, matrix(graph), k)
        states, levels, e_cop, e_rob = oracle.v1_levelsync(adj, k, nthreads=8)
        assert np.array_equal(solver.states(), states)
        assert np.array_equal(solver.level_counts(), levels)
        # configurations holding states of each depth (the level lists)
        st = states.reshape(-1, n, 2)
        dep = np.where(st & 1, st >> 1, 127)
        rows = solver.level_rows()
        for d in range(1, 40):
            for t in range(2):
                assert rows[d, t] == int((dep[:, :, t] == d).any(axis=1).sum()), (graph, d, t)
        got = solver.retro_edges()
        assert got == (e_cop, e_rob, int((states & 1).sum()))


def test_move_counts_of_dense_graphs_match_oracle(solver, oracle):
    """Move products too large for the LDS dedup set (degree 22, k = 3: a
    set of 2 * 22^3 rounded up to 32768 slots, 256 KiB) are counted through
    the device-memory set: transitions and retro-edges equal the oracle's."""
    n = 24
    adj = np.ones((n, n), np.uint8) - np.eye(n, dtype=np.uint8)
    for i in range(0, n, 2):  # minus a perfect matching
        adj[i, i + 1] = adj[i + 1, i] = 0
    say("dense move counts")
    solver.load_graph_matrix(adj, 3, 1)
    solver.build_aux_graph()
    solver.initialize_captures()
    solver.main_loop()
    states = solver.states()
    assert solver.count_transitions() == int(oracle.transitions(adj, 3)[0][-1])
    assert solver.retro_edges() == oracle.retro_edges(adj, 3, states)


def test_er200_k2_matches_oracle(solver, oracle):
    path = matrix("er200_p0.03_s1")
    n, adj = oracle.load_matrix(path)
    say("er200 k=2")
    base, passes, verdict = run_steps(solver, path, 2)
    states, levels, e_cop, e_rob = oracle.v1_levelsync(adj, 2, nthreads=8)
    assert np.array_equal(solver.states(), states)
    assert passes == oracle.pass_counts_from_depths(n, 2, states)
    best, plys = oracle.Solved(states, base, passes, 2, n, 2).verdict()
    assert (verdict is None) == (best == -1)


def csv_rows(fn):
    with open(os.path.join(GOLDEN, "csv", fn)) as f:
        return [r for r in csv.DictReader(f) if r["status"] == "ok" and r["p"] == "1"]


ROWS = csv_rows("test_results.csv")


def test_recorded_csv_rows_plain_game(solver):
    for row in ROWS:
        say("csv", row["graph"], row["k"])
        base, passes, verdict = run_steps(solver, matrix(row["graph"]), int(row["k"]))
        assert str(base) == row["base_caps"]
        assert str(len(passes)) == row["passes"]
        if row["result"] == "WIN":
            assert positions(verdict) == row["positions"]
            assert str(verdict["plys"]) == row["capture_plys"]
        else:
            assert verdict is None


def test_cli_reproduces_recorded_rows(tmp_path):
    exe = os.path.join(REPO, "copsandrobbers_amd", "bin", "k_cops_visibility")
    from copsandrobbers_amd import harness
    for row in ROWS[::3]:
        say("cli", row["graph"], row["k"])
        res = harness.run_solver(exe, matrix(row["graph"]), int(row["k"]), 1, timeout=120, cwd=str(tmp_path))
        assert res["status"] == "ok", res
        for col in ("result", "capture_plys", "passes", "base_caps", "positions"):
            assert str(res[col]) == row[col], (row["graph"], col)


DROPIN = os.path.join(REPO, "oracle", "_ref", "k_cops_visibility_dropin")


@pytest.mark.skipif(not os.path.exists(DROPIN), reason="drop-in not built (needs the reference tree at build time)")
def test_dropin_in_reference_supervisor_reproduces_recorded_rows(tmp_path):
    """The reference's own app/main.cpp compiled with the plug-in as its solver
    translation unit (oracle/build_ref.sh) gives the recorded CSV facts."""
    from copsandrobbers_amd import harness
    for row in ROWS[1::4]:
        say("dropin", row["graph"], row["k"])
        res = harness.run_solver(DROPIN, matrix(row["graph"]), int(row["k"]), 1, timeout=120, cwd=str(tmp_path))
        assert res["status"] == "ok", res
        for col in ("result", "capture_plys", "passes", "base_caps", "positions"):
            assert str(res[col]) == row[col], (row["graph"], col)


def test_error_paths(solver, tmp_path):
    say("errors")
    _, adj = cra.parse_matrix(open(matrix("peterson"), "rb").read())
    cases = [
        (lambda: solver.load_graph_matrix(adj, 0), -6),
        (lambda: solver.load_graph_matrix(adj, 9), -6),
        (lambda: solver.load_graph_matrix(adj, 2, 0), -7),
        (lambda: solver.load_graph_matrix(np.zeros((0, 0), np.uint8), 1), -3),
        (lambda: solver.load_graph_matrix(np.ones((256, 256), np.uint8), 1), -4),
    ]
    iso = adj.copy()
    iso[3, :] = 0
    iso[:, 3] = 0
    cases.append((lambda: solver.load_graph_matrix(iso, 1), -5))
    empty = tmp_path / "empty.txt"
    empty.write_bytes(b"")
    cases.append((lambda: solver.load_graph_file(str(empty), 1), -3))
    cases.append((lambda: solver.load_graph_file(str(tmp_path / "missing.txt"), 1), -3))
    for fn, code in cases:
        with pytest.raises(cra.SolverError) as exc:
            fn()
        assert exc.value.code == code
    fresh = cra.Solver(0)
    with pytest.raises(cra.SolverError) as exc:
        fresh.build_aux_graph()
    assert exc.value.code == -10
    fresh.close()


def test_repeat_solves_are_identical(solver):
    """Solving twice in one solver (re-initialised tables) gives equal bytes."""
    say("repeat")
    run_steps(solver, matrix("grid36"), 3)
    a = solver.states()
    solver.initialize_captures()
    solver.main_loop()
    assert np.array_equal(solver.states(), a)


PULL_SETTINGS = [{"pull_scale": 0, "pull_tiles": 1, "pull_tile_rows": 4},
                 {"pull_scale": 0, "pull_tiles": 1, "pull_tile_rows": 3, "pull_tile_chunk": 5, "pull_tile_blocks": 24},
                 {"pull_scale": 0, "pull_tiles": 1, "pull_tile1": 0},
                 {"pull_scale": 0, "pull_tiles": 1, "pull_tile1": 3},
                 {"pull_gate": 1, "levels_per_sync": 2},
                 {"robber_open": 0},
                 {"robber_open": 1000},
                 {"row_blocks": 3, "log_blocks": 5, "depth_log": 40, "depth_log_min": 0},
                 {"scan_tile": 4, "levels_per_sync": 3},
                 {"push_dirty": 0, "pull_scale": 10**9},
                 {"pull_scale": 0, "pull_tiles": 0},
                 {"pull_scale": 10**9},
                 {"depth_log": 0},
                 {"pull_scale": 0, "depth_log": 40, "depth_log_min": 0},
                 {"pull_scale": 0, "depth_log": 40, "depth_log_min": 0, "fuse_finalize": 0},
                 {"depth_log": 40, "depth_log_min": 0},
                 {"pull_scale": 0, "depth_log": 2, "depth_log_min": 0, "levels_per_sync": 3},
                 {"pred_csr": 1, "push_dirty": 0, "pull_scale": 10**9},
                 {"pred_csr": 1, "push_dirty": 1, "pull_scale": 10**9},
                 {"push_filter": 0, "pull_scale": 10**9},
                 {"push_blocks": 8, "pull_scale": 10**9},
                 {"pull_scale": 0, "pull_pair": 0},
                 {"end_pass": 0, "levels_per_sync": 3},
                 {"end_pass": 0, "depth_log": 40, "depth_log_min": 0},
                 {"pull_scale": 0, "pull_pair": 2, "pair_rows": 1, "pair_blocks": 8},
                 {"pull_scale": 0, "pull_pair": 2, "pair_rows": 3, "depth_log": 40, "depth_log_min": 0},
                 {"pull_scale": 0, "pull_pair": 2, "pair_rows": 8, "depth_log": 40, "depth_log_min": 0,
                  "fuse_finalize": 0}]
PULL_IDS = ["pull_tiles4", "pull_tiles3_chunk5", "pull_tile1_off", "pull_tile1_3", "pull_gate1", "robber_by_new", "robber_by_open", "small_grids", "scan_tiles4", "push_no_marks", "pull_rows", "push_only", "direct_bytes", "pull_log_fused",
            "pull_log_unfused", "depth_log", "small_log", "push_csr", "push_csr_marks", "push_unfiltered",
            "push_8_blocks", "pull_stage_by_stage", "two_launch_level_end", "two_launch_level_end_log", "pair_rows1", "pair_rows3_log", "pair_rows8_log_unfused"]
DEFAULTS = {"pull_scale": -1, "pull_tiles": 1, "pull_tile_rows": 4, "pull_tile_chunk": 16, "pull_tile_blocks": 8192, "depth_log": 40, "depth_log_min": 1 << 24,
            "levels_per_sync": 8, "fuse_finalize": 1, "pull_tile1": 9, "pull_gate": 4, "robber_open": 2, "row_blocks": 32768, "log_blocks": 4096, "scan_tile": 8192, "push_dirty": 1,
            "pred_csr": 0, "push_filter": 1, "push_blocks": 2048, "pull_pair": 1, "pair_rows": 4, "pair_blocks": 1024,
            "end_pass": 1}
PULL_CASES = ["v1__cycle100__k2__p1", "v1__dodecahedron__k3__p1", "v1__grid36__k3__p1", "v1__grid9__k3__p1",
              "v1__line3__k2__p1", "v1__outerplanar8__k3__p1", "v1__peterson__k3__p1", "v1__planar32__k3__p1",
              "v1__scotlandyard-all__k2__p1", "v1__test__k2__p1"]


@pytest.mark.parametrize("setting", PULL_SETTINGS, ids=PULL_IDS)
def test_pull_kernels_and_push_agree_with_reference(solver, setting):
    """Every level by the factorised pull (LDS-tiled stages with tiles of
    several sizes, or one row per lane), every level by the push, depth bytes
    written directly, or through a depth log so small that it is applied at
    every synchronisation and overflows into direct writes, gives the
    reference's tables and pass counts."""
    try:
        solver.set_tuning(**setting)
        for name in PULL_CASES:
            e = INDEX[name]
            say("direction", setting, name)
            _, passes, _ = run_steps(solver, matrix(e["graph"]), e["k"])
            assert hashlib.sha256(solver.states().tobytes()).hexdigest() == e["states_sha256"], name
            assert passes == e["passes"], name
    finally:
        solver.set_tuning(**DEFAULTS)


@pytest.mark.parametrize("setting", PULL_SETTINGS[:3] + PULL_SETTINGS[-4:], ids=PULL_IDS[:3] + PULL_IDS[-4:])
def test_pull_stages_for_four_to_six_cops_match_oracle(solver, oracle, setting):
    """k = 4..6 run pull stages with up to five cops still to move; the tables
    equal the oracle's."""
    try:
        solver.set_tuning(**setting)
        for graph, k in (("planar8", 4), ("peterson", 4), ("outerplanar8", 5), ("line10", 5), ("grid9", 6)):
            say("pull k", setting, graph, k)
            n, adj = oracle.load_matrix(matrix(graph))
            run_steps(solver, matrix(graph), k)
            want = oracle.v1_solve(adj, k, 1)
            assert np.array_equal(solver.states(), want.states), (graph, k)
    finally:
        solver.set_tuning(**DEFAULTS)
