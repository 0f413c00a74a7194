"""Stay rules, the DP-table and perfect-game outputs, edge-list input and the
cop-number search on the device, against the reference's data files and the
oracle (itself pinned on those files in test_rounds.py)."""
import csv
import gzip
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, matrix
from test_rounds import EXACT, GAMES, PRODUCER, check_game, recorded_games

import copsandrobbers_amd as cra
from copsandrobbers_amd import workloads

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def solver():
    s = cra.Solver(0)
    yield s
    s.set_rules(False, False)
    s.close()


def solve(solver, k, p=1, path=None, adj=None, rules=(False, False)):
    solver.set_rules(*rules)
    if path is not None:
        solver.load_graph_file(path, k, p)
    else:
        solver.load_graph_matrix(adj, k, p)
    solver.build_aux_graph()
    base = solver.initialize_captures()
    passes = solver.main_loop()
    verdict = solver.find_final_result()
    return base, passes, verdict


def test_dp_table_matches_the_reference_file(solver, tmp_path):
    say("dp table grid36")
    solve(solver, 2, path=matrix("grid36"), rules=(True, True))
    out = tmp_path / "dp.json"
    solver.write_dp_table(out)
    with gzip.open(os.path.join(GOLDEN, "dp_tables", "grid36_2cops_dp_table.json.gz"), "rb") as f:
        assert out.read_bytes() == f.read()


def test_stay_rule_tables_match_the_oracle(solver, oracle):
    cases = [("grid9", 2, 1), ("peterson", 2, 1), ("planar32", 2, 1), ("dodecahedron", 3, 1), ("grid9", 2, 2),
             ("line10", 2, 3)]
    di = np.zeros((11, 11), np.uint8)
    for u in range(11):
        di[u, (u + 2) % 11] = 1
        di[u, (u * 3 + 1) % 11] = 1
    for graph, k, p in cases:
        for rules in ((True, False), (False, True), (True, True)):
            say("stay rules", graph, k, p, rules)
            n, adj = oracle.load_matrix(matrix(graph))
            base, passes, _ = solve(solver, k, p, adj=adj, rules=rules)
            want = oracle.v1_solve(adj, k, p, cop_stay=int(rules[0]), robber_stay=int(rules[1]))
            assert np.array_equal(solver.states(), want.states), (graph, k, p, rules)
            assert passes == want.pass_counts and base == want.base_caps
    # a directed graph where staying matters, and an isolated vertex that stay moves make legal
    iso, _ = workloads.er_graph(8, 0.4, 2)
    iso = np.pad(iso, ((0, 1), (0, 1)))
    for adj, k, rules in ((di, 2, (True, True)), (di, 1, (False, True)), (iso, 2, (True, True))):
        say("stay rules special", adj.shape[0], k, rules)
        solve(solver, k, adj=adj, rules=rules)
        want = oracle.v1_solve(adj, k, 1, cop_stay=int(rules[0]), robber_stay=int(rules[1]))
        assert np.array_equal(solver.states(), want.states)


@pytest.mark.parametrize("stem,graph,k", [g for g in recorded_games() if g[0] in PRODUCER])
def test_perfect_games(solver, oracle, stem, graph, k, tmp_path):
    """For both producers' orders the device walk equals the oracle's on graphs
    the oracle solves quickly; every recorded game is legal and value-optimal
    under the device table and its producer's rules, as long as the device's
    game in that order, and seven are the same game (six the same bytes:
    cycle100's file was written with another indentation)."""
    say("game", stem)
    n, adj = oracle.load_matrix(matrix(graph))
    solve(solver, k, adj=adj, rules=(True, True))
    states = solver.states()
    small = oracle.config_count(n, k) * n <= 600000
    want = oracle.v1_solve(adj, k, 1, cop_stay=1, robber_stay=1) if small else None
    if small:
        assert np.array_equal(states, want.states)
    for order in ("py", "cpp"):
        mine = solver.extract_game(order)
        check_game(oracle, adj, k, states, mine, order=order)
        if small:
            assert mine == oracle.perfect_game(adj, k, want.states, order=order), order
        out = tmp_path / f"game_{order}.json"
        solver.write_perfect_game(out, order)
        assert out.read_text() == json.dumps(mine, indent=4)
    order = PRODUCER[stem]
    mine = solver.extract_game(order)
    with open(os.path.join(GAMES, stem + "_perfect_game.json")) as f:
        raw = f.read()
    recorded = json.loads(raw)
    assert check_game(oracle, adj, k, states, recorded, order=order) == \
        check_game(oracle, adj, k, states, mine, order=order)
    assert len(mine) == len(recorded)
    if stem in EXACT:
        assert mine == recorded
        if raw == json.dumps(recorded, indent=4):  # the json.dump(indent=4) layout
            assert (tmp_path / f"game_{order}.json").read_text() == raw


def test_cop_numbers_match_the_recorded_rows(solver):
    rows = {}
    with open(os.path.join(GOLDEN, "csv", "test_results.csv")) as f:
        for r in csv.DictReader(f):
            if r["status"] == "ok" and r["p"] == "1":
                rows.setdefault(r["graph"], {})[int(r["k"])] = r["result"]
    solver.set_rules(False, False)
    for graph, res in sorted(rows.items()):
        say("cop number", graph)
        wins = [k for k in sorted(res) if res[k] == "WIN"]
        solver.load_graph_file(matrix(graph), 1)
        k, verdict = solver.cop_number(max(res))
        assert k == (wins[0] if wins else 0), graph
        assert (verdict is not None) == bool(wins)


def test_edge_file_gives_the_matrix_file_table(solver):
    say("edge file")
    for graph, k in (("peterson", 3), ("dodecahedron", 2)):
        solver.set_rules(False, False)
        solver.load_edge_file(os.path.join(GOLDEN, "edges", graph + ".txt"), k)
        solver.build_aux_graph()
        solver.initialize_captures()
        passes = solver.main_loop()
        table = solver.states()
        _, want, _ = solve(solver, k, path=matrix(graph))
        assert passes == want and np.array_equal(table, solver.states())


def test_round_outputs_need_the_two_column_table(solver, tmp_path):
    solver.set_rules(False, False)
    solver.load_graph_file(matrix("grid9"), 2, 2)
    solver.build_aux_graph()
    solver.initialize_captures()
    solver.main_loop()
    with pytest.raises(cra.SolverError) as exc:
        solver.write_dp_table(tmp_path / "x.json")
    assert exc.value.code == -11
    solver.set_rules(True, False)
    solver.load_graph_file(matrix("grid9"), 2, 2, cra.VARIANT_MACRO)
    with pytest.raises(cra.SolverError) as exc:
        solver.build_aux_graph()
    assert exc.value.code == -11
    solver.set_rules(False, False)
