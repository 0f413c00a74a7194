"""Pins the CPU oracle against the reference's own outputs (CPU only).

Golden data: tests/golden/ref/*.npz are the cache blobs the compiled
reference wrote (solvers/k_cops_visibility*.cpp through CacheManager.h), and
tests/golden/csv/*.csv are the reference's recorded regression results
(console_outputs/). Closed-form cases pin what the golden data leaves open.
"""
import csv
import hashlib
import itertools
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, matrix

with open(os.path.join(GOLDEN, "ref_index.json")) as f:
    INDEX = json.load(f)

V1_INLINE = sorted(k for k, v in INDEX.items() if v["solver"] == "v1" and v["arrays"])
V2_INLINE = sorted(k for k, v in INDEX.items() if v["solver"] == "v2" and v["arrays"])
MATRIX_NAMES = sorted(os.path.splitext(f)[0] for f in os.listdir(os.path.join(GOLDEN, "matrices")))


def golden(name):
    return np.load(os.path.join(GOLDEN, "ref", INDEX[name]["arrays"]))


def load(oracle, graph):
    return oracle.load_matrix(matrix(graph))


# ---------------------------------------------------------------- parsing
@pytest.mark.parametrize("text,n,edges", [
    (b"010\n101\n010\n-", 3, {(0, 1), (1, 0), (1, 2), (2, 1)}),
    (b"01\r\n10\r\n-\n", 2, {(0, 1), (1, 0)}),          # CR does not add width
    (b"0 1\n1 0\n", 3, {(0, 1), (1, 0)}),               # spaces count in the width
    (b"01\n\n\n10\n", 2, {(0, 1), (1, 0)}),             # blank lines skipped
    (b"011\n10-1\n111\n", 3, {(0, 1), (0, 2), (1, 0)}),  # '-' ends parsing
    (b"-\n", 0, set()),
    (b"", 0, set()),
    (b"0101\n1010\n", 4, {(0, 1), (0, 3), (1, 0), (1, 2)}),  # missing rows stay 0
])
def test_parse_matrix_cases(oracle, text, n, edges):
    got_n, adj = oracle.parse_matrix(text)
    assert got_n == n
    assert {tuple(x) for x in np.argwhere(adj)} == edges


@pytest.mark.parametrize("graph", MATRIX_NAMES)
def test_reference_matrices_parse(oracle, graph):
    n, adj = load(oracle, graph)
    assert 0 < n <= 255
    assert adj.shape == (n, n)
    assert not adj.diagonal().any()


# -------------------------------------------------- configurations / ranks
@pytest.mark.parametrize("n,k", [(1, 1), (3, 2), (5, 3), (10, 3), (11, 4), (7, 5), (4, 8)])
def test_config_enumeration_is_lexicographic_multisets(oracle, n, k):
    cfg = oracle.configs(n, k)
    brute = sorted(itertools.combinations_with_replacement(range(n), k))
    assert [tuple(int(x) for x in r) for r in cfg] == brute
    assert oracle.config_count(n, k) == len(brute)


@pytest.mark.parametrize("n,k", [(2, 1), (6, 2), (9, 3), (11, 4), (5, 6)])
def test_rank_is_enumeration_index(oracle, n, k):
    for i, row in enumerate(oracle.configs(n, k)):
        assert oracle.rank(row, n) == i


def test_config_count_edge_cases(oracle):
    assert oracle.config_count(10, 0) == 0
    assert oracle.config_count(10, 257) == 0
    assert oracle.config_count(1, 5) == 1
    assert oracle.config_count(200, 3) == 1353400
    assert oracle.config_count(200, 4) == 68685050


@pytest.mark.parametrize("name", V1_INLINE[:12])
def test_configs_match_reference_blob(oracle, name):
    e = INDEX[name]
    n, _ = load(oracle, e["graph"])
    g = golden(name)
    assert int(g["header"][3]) == oracle.config_count(n, e["k"])
    assert np.array_equal(oracle.configs(n, e["k"]).reshape(-1), g["configs"])


# ------------------------------------------------------------ team moves
@pytest.mark.parametrize("name", V1_INLINE)
def test_transitions_match_reference_blob(oracle, name):
    e = INDEX[name]
    _, adj = load(oracle, e["graph"])
    heads, trans = oracle.transitions(adj, e["k"])
    g = golden(name)
    assert np.array_equal(heads, g["heads"])
    assert np.array_equal(trans, g["transitions"])
    assert len(trans) == e["edge_pointers"]


def test_transitions_brute_force(oracle):
    rng = np.random.default_rng(7)
    for _ in range(5):
        n = int(rng.integers(4, 9))
        adj = (rng.random((n, n)) < 0.45).astype(np.uint8)
        np.fill_diagonal(adj, 0)
        adj |= np.eye(n, k=1, dtype=np.uint8) | np.eye(n, k=-1, dtype=np.uint8)  # no isolated vertex
        for k in (1, 2, 3):
            cfg = [tuple(r) for r in oracle.configs(n, k)]
            heads, trans = oracle.transitions(adj, k)
            for cid, c in enumerate(cfg):
                opts = [np.flatnonzero(adj[v]) for v in c]
                want = sorted({cfg.index(tuple(sorted(t))) * n for t in itertools.product(*opts)})
                assert list(trans[heads[cid]:heads[cid + 1]]) == want


# ----------------------------------------------------- v1 (column solver)
@pytest.mark.parametrize("name", V1_INLINE)
def test_v1_tables_match_reference(oracle, name):
    e = INDEX[name]
    _, adj = load(oracle, e["graph"])
    s = oracle.v1_solve(adj, e["k"], e["p"])
    g = golden(name)
    assert np.array_equal(s.states, g["states"])
    assert s.pass_counts == e["passes"]
    assert s.base_caps == e["base_caps"]
    best, plys = s.verdict()
    if e["result"] == "WIN":
        cfg = oracle.configs(adj.shape[0], e["k"])[best]
        assert ", ".join(str(int(x)) for x in cfg) == e["positions"]
        assert plys == e["plys"]
        assert plys // (2 * e["p"]) == e["cycles"]
    else:
        assert best == -1


@pytest.mark.parametrize("name", [n for n in V1_INLINE if INDEX[n]["p"] == 1])
def test_levelsync_equals_pass_sweep(oracle, name):
    e = INDEX[name]
    _, adj = load(oracle, e["graph"])
    states, levels, e_cop, e_rob = oracle.v1_levelsync(adj, e["k"], nthreads=3)
    assert np.array_equal(states, golden(name)["states"])
    assert oracle.pass_counts_from_depths(adj.shape[0], e["k"], states) == e["passes"]
    assert oracle.retro_edges(adj, e["k"], states)[:2] == (e_cop, e_rob)
    hist = np.zeros((127, 2), np.int64)
    marked = states & 1
    depth = states >> 1
    for t in (0, 1):
        d = depth[t::2][marked[t::2] == 1]
        hist[:, t] = np.bincount(d, minlength=127)[:127]
    assert np.array_equal(hist, levels.astype(np.int64))


def passes_from_column_depths(states, n, k, cols, C):
    """Per-pass counts of the 2p-column sweep from the depth table alone: a
    state of column c and depth t > 0 is marked in pass
    t - (c + t) // cols + [(c + t) % cols == 0] (the device column solver's
    rule). Also returns the (column, depth) pairs whose successor chain would
    not end on a capture column (0 or 1); the rule needs there to be none."""
    st = states.reshape(C, n, cols)
    per_pass, odd = {}, []
    for c in range(cols):
        col = st[:, :, c]
        d = (col >> 1)[(col & 1) == 1]
        for t, cnt in enumerate(np.bincount(d, minlength=128)):
            if t == 0 or cnt == 0:
                continue
            if (c + t) % cols not in (0, 1):
                odd.append((c, t))
            p = t - (c + t) // cols + (1 if (c + t) % cols == 0 else 0)
            per_pass[p] = per_pass.get(p, 0) + int(cnt)
    last = max(per_pass, default=0)
    return [per_pass.get(i, 0) for i in range(1, last + 1)] + [0], odd


@pytest.mark.parametrize("name", sorted(n for n, v in INDEX.items() if v["solver"] == "v1" and v["p"] >= 2))
def test_column_pass_rule_reproduces_reference_passes(oracle, name):
    """The pass counts the device column solver derives from its depth
    histogram equal the reference's printed passes on every golden p >= 2 case."""
    e = INDEX[name]
    n, adj = load(oracle, e["graph"])
    s = oracle.v1_solve(adj, e["k"], e["p"])
    passes, odd = passes_from_column_depths(s.states, n, e["k"], 2 * e["p"], oracle.config_count(n, e["k"]))
    assert odd == []
    assert passes == e["passes"] == s.pass_counts


def test_column_pass_rule_detects_a_planted_depth(oracle):
    n, adj = load(oracle, "grid9")
    s = oracle.v1_solve(adj, 2, 2)
    bad = s.states.copy()
    i = int(np.flatnonzero((bad & 1) & ((bad >> 1) > 3))[0])
    bad[i] = np.uint8(bad[i] + 2)  # one state one level deeper
    C = oracle.config_count(n, 2)
    passes, odd = passes_from_column_depths(bad, n, 2, 4, C)
    assert odd or passes != INDEX["v1__grid9__k2__p2"]["passes"]


@pytest.mark.parametrize("name", [n for n, v in INDEX.items() if v["solver"] == "v1" and not v["arrays"]])
def test_large_v1_tables_match_reference_hash(oracle, name):
    e = INDEX[name]
    _, adj = load(oracle, e["graph"])
    states, _, _, _ = oracle.v1_levelsync(adj, e["k"], nthreads=8)
    assert hashlib.sha256(states.tobytes()).hexdigest() == e["states_sha256"]
    assert oracle.pass_counts_from_depths(adj.shape[0], e["k"], states) == e["passes"]


def test_retro_edges_brute_force(oracle):
    _, adj = load(oracle, "grid9")
    k = 2
    n = adj.shape[0]
    s = oracle.v1_solve(adj, k, 1)
    cfg = [tuple(r) for r in oracle.configs(n, k)]
    pred = {i: set() for i in range(len(cfg))}
    for i, c in enumerate(cfg):
        for t in itertools.product(*[np.flatnonzero(adj[v]) for v in c]):
            pred[cfg.index(tuple(sorted(t)))].add(i)
    e_cop = e_rob = won = 0
    for cid in range(len(cfg)):
        for r in range(n):
            b0, b1 = s.states[(cid * n + r) * 2], s.states[(cid * n + r) * 2 + 1]
            if b0 & 1:
                won += 1
                e_rob += int(adj[:, r].sum()) if (b0 >> 1) <= 125 else 0
            if b1 & 1:
                won += 1
                e_cop += len(pred[cid]) if (b1 >> 1) <= 125 else 0
    assert oracle.retro_edges(adj, k, s.states) == (e_cop, e_rob, won)


# ------------------------------------------------------------- closed forms
def test_closed_forms(oracle):
    # a single cop on a path wins from the centre in (n-2) plies
    for graph, centre, plys in (("line10", 4, 8), ("line5", 2, 2)):
        _, adj = load(oracle, graph)
        s = oracle.v1_solve(adj, 1, 1)
        best, got = s.verdict()
        assert (best, got) == (centre, plys)
    # one cop loses on cycles of length >= 4
    for graph in ("cycle5", "cycle10"):
        _, adj = load(oracle, graph)
        assert oracle.v1_solve(adj, 1, 1).verdict()[0] == -1
    # a dominating set of size 3 (Petersen {0, 2, 6}) captures in one ply
    _, adj = load(oracle, "peterson")
    s = oracle.v1_solve(adj, 3, 1)
    best, plys = s.verdict()
    assert plys == 1 and tuple(oracle.configs(10, 3)[best]) == (0, 2, 6)
    # base captures = sum over configurations of distinct cop vertices
    cfg = oracle.configs(10, 3)
    assert s.base_caps == sum(len(set(r)) for r in cfg)


def test_depth_cap_is_active_on_line100(oracle):
    _, adj = load(oracle, "line100")
    s = oracle.v1_solve(adj, 1, 1)
    depths = s.states[s.states & 1 == 1] >> 1
    assert depths.max() == 126


def test_planted_mistake_is_detected(oracle):
    """A table with one robber depth off by one must fail the comparison."""
    _, adj = load(oracle, "grid36")
    s = oracle.v1_solve(adj, 2, 1)
    bad = s.states.copy()
    i = int(np.flatnonzero((bad & 1) & ((bad >> 1) > 3))[0])
    bad[i] = (bad[i] & 1) | (((bad[i] >> 1) + 1) << 1)
    assert not np.array_equal(bad, golden("v1__grid36__k2__p1")["states"])
    assert oracle.pass_counts_from_depths(36, 2, bad) != INDEX["v1__grid36__k2__p1"]["passes"] or \
        not np.array_equal(bad, s.states)


# ------------------------------------------------------------- v2 (macro)
@pytest.mark.parametrize("name", V2_INLINE)
def test_v2_tables_match_reference(oracle, name):
    e = INDEX[name]
    _, adj = load(oracle, e["graph"])
    s = oracle.v2_solve(adj, e["k"], e["p"])
    assert np.array_equal(s.states, golden(name)["states"])
    assert s.pass_counts == e["passes"]
    assert s.base_caps == e["base_caps"]
    best, plys = s.verdict()
    if e["result"] == "WIN":
        cfg = oracle.configs(adj.shape[0], e["k"])[best]
        assert ", ".join(str(int(x)) for x in cfg) == e["positions"]
        assert plys == e["plys"]
    else:
        assert best == -1


def test_v2_equals_v1_at_p1(oracle):
    for graph, k in (("peterson", 3), ("grid36", 2), ("planar32", 2)):
        _, adj = load(oracle, graph)
        a = oracle.v1_solve(adj, k, 1)
        b = oracle.v2_solve(adj, k, 1)
        assert np.array_equal(a.states[0::2], b.states)


# ------------------------------------------------- recorded regression CSVs
def csv_rows(fn):
    with open(os.path.join(GOLDEN, "csv", fn)) as f:
        return [r for r in csv.DictReader(f) if r["status"] == "ok"]


def check_row(oracle, row, solver):
    n, adj = load(oracle, row["graph"])
    k, p = int(row["k"]), int(row["p"])
    s = solver(adj, k, p)
    assert str(s.base_caps) == row["base_caps"]
    assert str(len(s.pass_counts)) == row["passes"]
    best, plys = s.verdict()
    if row["result"] == "WIN":
        assert best >= 0
        assert ", ".join(str(int(x)) for x in oracle.configs(n, k)[best]) == row["positions"]
        assert str(plys) == row["capture_plys"]
    else:
        assert row["result"] == "LOSS" and best == -1


V1_ROWS = csv_rows("test_results.csv")
V2_ROWS = [r for r in csv_rows("test_results_v2.csv") if float(r["mainloop_s"] or 0) <= 30]


@pytest.mark.parametrize("row", V1_ROWS, ids=[f"{r['graph']}-k{r['k']}-p{r['p']}" for r in V1_ROWS])
def test_v1_csv_rows(oracle, row):
    check_row(oracle, row, lambda adj, k, p: oracle.v1_solve(adj, k, p))


@pytest.mark.parametrize("row", V2_ROWS, ids=[f"{r['graph']}-k{r['k']}-p{r['p']}" for r in V2_ROWS])
def test_v2_csv_rows(oracle, row):
    check_row(oracle, row, lambda adj, k, p: oracle.v2_solve(adj, k, p))


def test_affected_csv_equals_main_csv():
    main = {(r["graph"], r["k"], r["p"]): r for r in csv_rows("test_results.csv")}
    for r in csv_rows("test_results_affected.csv"):
        m = main[(r["graph"], r["k"], r["p"])]
        for col in ("result", "capture_plys", "passes", "base_caps", "positions"):
            assert r[col] == m[col]


# ------------------------------------- checks for tables the sweep cannot reach
P1_INLINE = [n for n in V1_INLINE if INDEX[n]["p"] == 1]


@pytest.mark.parametrize("n,k", [(1, 1), (5, 3), (7, 4), (11, 2), (4, 6)])
def test_unrank_inverts_rank(oracle, n, k):
    for i, c in enumerate(oracle.configs(n, k)):
        assert oracle.unrank(i, n, k) == tuple(int(x) for x in c)


@pytest.mark.parametrize("name", P1_INLINE)
def test_bellman_check_accepts_reference_tables(oracle, name):
    """Every state of every golden p = 1 table satisfies the local equations."""
    e = INDEX[name]
    _, adj = load(oracle, e["graph"])
    states = golden(name)["states"]
    ids = np.arange(len(states), dtype=np.uint64)
    assert oracle.check_states(adj, e["k"], states, ids, nthreads=3) == (0, None)


def test_bellman_check_finds_planted_mistakes(oracle):
    """One wrong state -- a depth off by one either way, a won state cleared,
    an unwon state marked, a capture moved -- is found when the check covers
    that state, and checks that skip it still pass."""
    rng = np.random.default_rng(3)
    for name, k in (("grid36", 2), ("robertson", 2)):
        _, adj = load(oracle, name)
        states = golden(f"v1__{name}__k{k}__p1")["states"]
        won = np.flatnonzero((states & 1) == 1)
        deep = won[(states[won] >> 1) > 2]
        unwon = np.flatnonzero((states & 1) == 0)
        caps = np.flatnonzero(states == 1)
        plants = [(int(caps[5]), 1 | (1 << 1))]           # capture given round 1
        for i in rng.choice(deep, min(4, len(deep)), replace=False):
            plants.append((int(i), int(states[i]) + 2))   # one round deeper
            plants.append((int(i), int(states[i]) - 2))   # one round shallower
            plants.append((int(i), 254))                  # cleared
        for i in rng.choice(unwon, min(4, len(unwon)), replace=False):
            plants.append((int(i), 1 | (7 << 1)))         # marked without a witness
        for i in rng.choice(won[(states[won] >> 1) > 0], 2, replace=False):
            plants.append((int(i), 254))                  # shallow won state cleared
        assert len(plants) >= 7
        for i, b in plants:
            bad = states.copy()
            bad[i] = b
            count, first = oracle.check_states(adj, k, bad, np.array([i], np.uint64))
            assert count == 1 and first == i, (name, i, b)
            others = np.setdiff1d(np.arange(len(states), dtype=np.uint64), [i])
            sample = others[rng.choice(len(others), 200, replace=False)]
            assert oracle.check_states(adj, k, states, sample)[0] == 0


@pytest.mark.parametrize("name", P1_INLINE)
def test_table_summary_equals_sweep_verdict_and_levels(oracle, name):
    e = INDEX[name]
    n, adj = load(oracle, e["graph"])
    states = golden(name)["states"]
    lv, best, plys = oracle.table_summary(n, e["k"], states, nthreads=4)
    _, levels, _, _ = oracle.v1_levelsync(adj, e["k"], nthreads=2)
    assert np.array_equal(lv, levels)
    s = oracle.Solved(states, 0, [], 2, n, e["k"])
    assert (best, plys) == s.verdict()
    assert oracle.passes_from_levels(lv) == e["passes"]
    # base captures in closed form: n * (multisets containing a given vertex)
    assert int(lv[0, 0]) == e["base_caps"] == n * (oracle.config_count(n, e["k"]) -
                                                   oracle.config_count(n - 1, e["k"]) if n > 1 else 1)


@pytest.mark.parametrize("name", [n for n in P1_INLINE if INDEX[n]["graph"] != "test"])
def test_capped_port_uncapped_equals_levelsync(oracle, name):
    e = INDEX[name]
    _, adj = load(oracle, e["graph"])
    if not np.array_equal(adj, adj.T):
        pytest.skip("directed graph: the capped port takes undirected graphs only")
    states, info = oracle.v1_levelsync_capped(adj, e["k"], nthreads=3)
    want, _, e_cop, e_rob = oracle.v1_levelsync(adj, e["k"], nthreads=2)
    assert np.array_equal(states, want)
    assert (info["e_cop"], info["e_rob"]) == (e_cop, e_rob)
    assert info["level_fraction"] == 0.0


def test_capped_port_stops_at_its_cap(oracle):
    _, adj = load(oracle, "grid36")
    states, info = oracle.v1_levelsync_capped(adj, 3, nthreads=2, time_cap_s=0.0)
    assert info["levels_done"] <= 1 and info["elapsed_s"] < 5.0
