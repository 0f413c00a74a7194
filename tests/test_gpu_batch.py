"""Batched solving and the continuous-batching server against the reference's
tables and the six-step single-graph solve.

Expected values: the compiled reference's p = 1 cache blobs (tests/golden/ref)
for every golden graph that fits a batch (n <= 64, k <= 4), and the
single-graph device solve (itself checked against the reference and the
oracle in test_gpu_parity.py) for seeded random graphs. Integer work: every
comparison is exact.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np
import pytest

from conftest import GOLDEN, matrix

import copsandrobbers_amd as cra
from copsandrobbers_amd import batch as crb
from copsandrobbers_amd import workloads

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "ref_index.json")) as f:
    INDEX = json.load(f)


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def mix64(x):
    x = np.asarray(x, np.uint64)
    with np.errstate(over="ignore"):
        x = x ^ (x >> np.uint64(30))
        x = x * np.uint64(0xbf58476d1ce4e5b9)
        x = x ^ (x >> np.uint64(27))
        x = x * np.uint64(0x94d049bb133111eb)
        x = x ^ (x >> np.uint64(31))
    return x


def table_hash(states):
    """The documented table hash: sum over i of (byte_i + 1) * mix64(i) mod 2^64."""
    i = np.arange(states.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int(np.sum((states.astype(np.uint64) + np.uint64(1)) * mix64(i), dtype=np.uint64))


def single(solver, adj, k):
    solver.load_graph_matrix(adj, k, 1)
    solver.build_aux_graph()
    base = solver.initialize_captures()
    passes = solver.main_loop()
    verdict = solver.find_final_result()
    return base, passes, verdict, solver.states(), solver.stats().won_states


@pytest.fixture(scope="module")
def batch():
    b = crb.Batch(0)
    yield b
    b.close()


@pytest.fixture(scope="module")
def solver():
    s = cra.Solver(0)
    yield s
    s.close()


def golden_cases():
    out = []
    for name, e in sorted(INDEX.items()):
        if e["solver"] != "v1" or e["p"] != 1 or e["k"] > crb.MAX_BATCH_K:
            continue
        n, adj = cra.parse_matrix(open(matrix(e["graph"]), "rb").read())
        if n <= crb.MAX_N and adj.sum(1).min() > 0:
            out.append((name, adj, e["k"]))
    return out


def test_batch_matches_reference_tables(batch):
    cases = golden_cases()
    say("batch golden", len(cases))
    res, tabs = batch.solve([(adj, k) for _, adj, k in cases], tables=True)
    for i, (name, adj, k) in enumerate(cases):
        e = INDEX[name]
        r = res[i]
        assert r["status"] == 0, name
        if e["arrays"]:
            g = np.load(os.path.join(GOLDEN, "ref", e["arrays"]))
            assert np.array_equal(tabs[i], g["states"]), name
        assert hashlib.sha256(tabs[i].tobytes()).hexdigest() == e["states_sha256"], name
        assert r["table_hash"] == table_hash(tabs[i]), name
        assert batch.pass_counts(i) == e["passes"], name
        assert r["n_passes"] == len(e["passes"]), name
        assert r["base_captures"] == e["base_caps"], name
        if e["result"] == "WIN":
            assert r["win"] and r["plys"] == e["plys"], name
            assert ", ".join(str(x) for x in r["positions"]) == e["positions"], name
        else:
            assert not r["win"], name


def random_graphs(count, seed_base):
    graphs = workloads.batch_graphs(count, seed_base=seed_base)
    # a directed graph (every vertex keeps an out-edge) and one with stay moves
    rng = workloads.SplitMix64(77)
    n = 13
    di = np.zeros((n, n), np.uint8)
    for u in range(n):
        di[u, (u + 3) % n] = 1
        v = rng.below(n)
        if v != u:
            di[u, v] = 1
    graphs.append((di, 2))
    loops, _ = workloads.er_graph(11, 0.25, 5)
    loops = loops.copy()
    loops[np.arange(11), np.arange(11)] = 1
    graphs.append((loops, 3))
    path64 = np.zeros((64, 64), np.uint8)
    for u in range(63):
        path64[u, u + 1] = path64[u + 1, u] = 1
    graphs.append((path64, 1))
    k4, _ = workloads.er_graph(12, 0.3, 2)
    graphs.append((k4, 4))
    return graphs


def test_batch_matches_single_graph_solves(batch, solver, oracle):
    graphs = random_graphs(48, seed_base=5000)
    say("batch random", len(graphs))
    res, tabs = batch.solve(graphs, tables=True)
    for i, (adj, k) in enumerate(graphs[-4:], start=len(graphs) - 4):  # directed, stay moves, n = 64, k = 4
        want = oracle.v1_solve(adj, k, 1)
        assert np.array_equal(tabs[i], want.states), i
        assert batch.pass_counts(i) == want.pass_counts, i
    for i, (adj, k) in enumerate(graphs):
        base, passes, verdict, states, won = single(solver, adj, k)
        r = res[i]
        assert np.array_equal(tabs[i], states), i
        assert batch.pass_counts(i) == passes, i
        assert r["base_captures"] == base and r["won_states"] == won, i
        assert r["win"] == (verdict is not None), i
        if verdict is not None:
            assert r["plys"] == verdict["plys"] and r["positions"] == verdict["positions"], i


@pytest.mark.parametrize("push_pct", [0, 10**9], ids=["always_pull", "always_push"])
def test_push_and_pull_levels_give_the_same_tables(batch, oracle, push_pct):
    """Every level pulled, or every level's cop turn pushed from its robber
    frontier (predecessors enumerated from in-neighbour lists), gives the
    tables of the default settings (which push only the sparse levels), and
    the oracle's on the directed, stay-move, n = 64 and k = 4 graphs."""
    graphs = random_graphs(24, seed_base=7100)
    say("batch push", push_pct, len(graphs))
    _, ref = batch.solve(graphs, tables=True)
    try:
        batch.set_tuning(push_pct=push_pct)
        res, tabs = batch.solve(graphs, tables=True)
        counts = [batch.pass_counts(i) for i in range(len(graphs))]
    finally:
        batch.set_tuning(push_pct=100)
    for i, (adj, k) in enumerate(graphs):
        assert res[i]["status"] == 0, i
        assert np.array_equal(tabs[i], ref[i]), (push_pct, i)
    for i, (adj, k) in enumerate(graphs[-4:], start=len(graphs) - 4):  # directed, stay moves, n = 64, k = 4
        want = oracle.v1_solve(adj, k, 1)
        assert np.array_equal(tabs[i], want.states), (push_pct, i)
        assert counts[i] == want.pass_counts, (push_pct, i)


def test_benchmarked_batch_equals_single_solves_and_oracle(batch, solver, oracle):
    """The 4096 graphs the batch benchmark solves (workloads.batch_graphs(4096),
    one packed launch): every table's hash, pass count, verdict and capture
    count equals the single-graph solve's, and every 64th graph's table equals
    the oracle's byte for byte."""
    graphs = workloads.batch_graphs(4096)
    say("benchmarked batch", len(graphs))
    arr = batch.solve(crb.pack(graphs), results="array")
    for i, (adj, k) in enumerate(graphs):
        base, passes, verdict, states, won = single(solver, adj, k)
        assert int(arr["table_hash"][i]) == table_hash(states), i
        assert batch.pass_counts(i) == passes, i
        assert int(arr["base_captures"][i]) == base and int(arr["won_states"][i]) == won, i
        assert bool(arr["win"][i]) == (verdict is not None), i
        if i % 64 == 0:
            assert np.array_equal(states, oracle.v1_solve(adj, k, 1).states), i


def test_packed_inputs_and_array_results_equal_the_dicts(batch):
    """Graphs packed once and results as a record array (the form the bench
    times) give the same results and tables as the list-of-dicts form."""
    graphs = random_graphs(24, seed_base=7000)
    say("batch packed", len(graphs))
    res, tabs = batch.solve(graphs, tables=True)
    packed = crb.pack(graphs)
    for _ in range(2):  # the packed arrays are reused
        arr, tabs2 = batch.solve(packed, tables=True, results="array")
        assert len(arr) == len(graphs)
        for i, r in enumerate(res):
            k = graphs[i][1]
            for f in ("status", "plys", "n_passes", "best_cid", "base_captures", "won_states", "table_hash",
                      "table_bytes"):
                assert int(arr[i][f]) == r[f], (i, f)
            assert bool(arr[i]["win"]) == r["win"], i
            assert tuple(int(x) for x in arr[i]["positions"][:k]) == r["positions"], i
            assert np.array_equal(tabs2[i], tabs[i]), i
    assert len(batch.solve(crb.pack([]), results="array")) == 0


def test_batch_reports_invalid_graphs_and_solves_the_rest(batch):
    good, _ = workloads.er_graph(9, 0.4, 1)
    iso = good.copy()
    iso[4, :] = 0
    big = np.ones((65, 65), np.uint8)
    graphs = [(good, 2), (iso, 2), (good, 5), (big, 1), (good, 1)]
    say("batch errors")
    res = batch.solve(graphs, raise_on_error=False)
    assert [r["status"] for r in res] == [0, -5, -6, -4, 0]
    alone = batch.solve([(good, 2), (good, 1)])
    assert res[0]["table_hash"] == alone[0]["table_hash"] and res[4]["table_hash"] == alone[1]["table_hash"]
    with pytest.raises(cra.SolverError) as exc:
        batch.solve(graphs)
    assert exc.value.code == -1
    assert batch.solve([]) == []


def test_server_results_equal_alone_and_in_the_mix(batch):
    """Requests arrive in bursts while earlier ones are on the device; each
    result equals the same graph solved alone."""
    graphs = workloads.batch_graphs(400, seed_base=9000)
    alone = [batch.solve([g])[0] for g in graphs[:40]]
    mix = batch.solve(graphs)
    for a, m in zip(alone, mix[:40]):
        for key in ("table_hash", "win", "plys", "positions", "base_captures", "won_states", "n_passes"):
            assert a[key] == m[key]
    say("server")
    got = {}
    with crb.Server(0, hbm_budget=8 << 30, max_batch=64, streams=4) as srv:
        tickets = []
        for lo in range(0, len(graphs), 50):
            for adj, k in graphs[lo:lo + 50]:
                tickets.append(srv.submit(adj, k))
            for r in srv.poll(wait_us=200):
                got[r["ticket"]] = r
        deadline = time.time() + 60
        while len(got) < len(graphs) and time.time() < deadline:
            for r in srv.poll(wait_us=2000):
                got[r["ticket"]] = r
        stats = srv.stats()
    assert len(got) == len(graphs)
    assert stats["submitted"] == stats["finished"] == len(graphs) and stats["launches"] >= 2
    for t, m in zip(tickets, mix):
        r = got[t]
        assert r["status"] == 0
        for key in ("table_hash", "win", "plys", "positions", "base_captures", "won_states", "n_passes"):
            assert r[key] == (m[key] if key != "positions" else m[key] + (0,) * (len(r[key]) - len(m[key])))
        assert r["latency_ms"] > 0


def request_bytes(n, k):
    """A request's admitted footprint: its state table plus its work planes
    (won, frontier, next x 2 columns, factorised-pull intermediates, depth bytes)."""
    from math import comb
    C = comb(n + k - 1, k)
    hmax = max([comb(n + i - 1, i) * comb(n + k - i - 1, k - i) for i in range(1, k)], default=0)
    return 2 * C * n + (5 * C + 2 * hmax + (C + 7) // 8) * 8


def test_server_budget_admits_one_at_a_time_and_rejects_oversize(batch):
    adj, _ = workloads.er_graph(30, 0.15, 4)
    graphs = [(adj, 3)] * 12
    ref = batch.solve(graphs[:1])[0]
    need = request_bytes(30, 3)
    say("server budget", need)
    got = {}
    # room for one request and not two: one graph per launch, one launch at a time
    with crb.Server(0, hbm_budget=int(need * 1.5), max_batch=64, streams=2) as srv:
        tickets = [srv.submit(a, k) for a, k in graphs]
        big, _ = workloads.er_graph(60, 0.2, 3)
        with pytest.raises(cra.SolverError) as exc:
            srv.submit(big, 4)
        assert exc.value.code == -8
        deadline = time.time() + 60
        while len(got) < len(graphs) and time.time() < deadline:
            for r in srv.poll(wait_us=2000):
                got[r["ticket"]] = r
        stats = srv.stats()
    assert stats["launches"] == len(graphs)
    assert need <= stats["peak_bytes"] <= need * 1.5
    for t in tickets:
        assert got[t]["table_hash"] == ref["table_hash"] and got[t]["plys"] == ref["plys"]


def test_server_keeps_tables_resident_and_readable(batch):
    """With retention on, each served request's table stays in HBM and reads
    back byte-equal to the batched solve of the same graph (whose tables equal
    the single-graph solver's, above); releasing frees them."""
    graphs = workloads.batch_graphs(300, seed_base=4000)
    want, tabs = batch.solve(graphs, tables=True)
    say("server resident tables")
    got = {}
    with crb.Server(0, hbm_budget=8 << 30, max_batch=64, streams=2) as srv:
        srv.retain_tables(True)
        tickets, status = srv.submit_many(graphs)
        assert all(t > 0 for t in tickets) and not any(status)
        deadline = time.time() + 60
        while len(got) < len(graphs) and time.time() < deadline:
            for r in srv.poll(wait_us=2000):
                got[r["ticket"]] = r
        assert len(got) == len(graphs)
        st = srv.table_stats()
        assert st["resident_tables"] == len(graphs) and st["evicted_launches"] == 0
        for i, t in enumerate(tickets):
            assert got[t]["table_hash"] == want[i]["table_hash"]
            assert np.array_equal(srv.table(t), tabs[i]), i
        for t in tickets[::2]:
            srv.release(t)
        with pytest.raises(cra.SolverError):
            srv.table(tickets[0])
        assert np.array_equal(srv.table(tickets[1]), tabs[1])
        for t in tickets[1::2]:
            srv.release(t)
        st = srv.table_stats()
        assert st["resident_tables"] == 0 and st["resident_bytes"] == 0


def test_server_evicts_oldest_tables_within_the_budget(batch):
    adj, _ = workloads.er_graph(30, 0.15, 4)
    graphs = [(adj, 3)] * 10
    ref, tabs = batch.solve(graphs[:1], tables=True)
    need = request_bytes(30, 3)
    say("server table eviction", need)
    got = {}
    with crb.Server(0, hbm_budget=int(need * 2.5), max_batch=1, streams=1) as srv:
        srv.retain_tables(True)
        tickets = []
        for a, k in graphs:  # one at a time, so each launch retains one table
            tickets.append(srv.submit(a, k))
            deadline = time.time() + 30
            while tickets[-1] not in got and time.time() < deadline:
                for r in srv.poll(wait_us=2000):
                    got[r["ticket"]] = r
        st = srv.table_stats()
        assert len(got) == len(graphs)
        assert st["evicted_launches"] >= 1
        assert np.array_equal(srv.table(tickets[-1]), tabs[0])
        with pytest.raises(cra.SolverError):
            srv.table(tickets[0])
        assert st["resident_bytes"] <= need * 2.5
    for t in tickets:
        assert got[t]["table_hash"] == ref[0]["table_hash"]
