"""Randomised parity of the plain solve (v1, p = 1) with the oracle: random small
graphs (6..40 vertices, 1..4 cops), and graphs of 65..130 vertices with 1..2
cops whose robber rows take two or three 64-bit words, under random speed
settings, none of which may change a table; and 66- and 80-vertex graphs
(two-word rows) with 3 cops under the pull's kernel variants. Fixed seeds; every case prints its
graph size and settings before it runs."""
import numpy as np
import pytest

import copsandrobbers_amd as cra

pytestmark = pytest.mark.gpu

CHOICES = {
    "pull_scale": [-1, 0, 10**9, 150],
    "depth_log": [40, 2, 0],
    "depth_log_min": [0, 1 << 24],
    "fuse_finalize": [0, 1],
    "pull_tiles": [0, 1, 1],
    "pull_tile1": [0, 3, 9],
    "pull_tile_rows": [2, 4],
    "pull_tile_chunk": [1, 16],
    "pull_gate": [0, 1, 4, 16],
    "robber_open": [0, 2, 1000],
    "levels_per_sync": [1, 2, 3, 5, 8],
    "scan_tile": [4, 7, 8192],
    "push_dirty": [0, 1],
    "pred_csr": [0, 0, 1],
    "push_filter": [0, 1, 1],
    "push_blocks": [8, 2048],
    "pull_pair": [0, 1, 2, 2],
    "pair_rows": [1, 3, 4, 8],
    "pair_blocks": [8, 1024],
    "end_pass": [0, 1, 1],
}


def random_graph(rng, lo=6, hi=40):
    """A random undirected graph in which every vertex has a move."""
    while True:
        n = int(rng.integers(lo, hi + 1))
        adj = (rng.random((n, n)) < rng.uniform(0.06, 0.35)).astype(np.uint8)
        adj = np.triu(adj, 1)
        adj = adj | adj.T
        if adj.sum(1).min() > 0:
            return n, adj


@pytest.mark.parametrize("seed", [11, 12])
def test_random_graphs_and_settings_match_oracle(oracle, seed):
    rng = np.random.default_rng(seed)
    s = cra.Solver(0)
    try:
        for case in range(30):
            n, adj = random_graph(rng)
            k = int(rng.integers(1, 5 if n <= 24 else 4))
            tuning = {name: int(rng.choice(values)) for name, values in CHOICES.items()}
            print(f"random case {seed}/{case}: n={n} k={k} {tuning}", flush=True)
            s.set_tuning(**tuning)
            s.load_graph_matrix(adj, k, 1)
            s.build_aux_graph()
            s.initialize_captures()
            s.main_loop()
            want = oracle.v1_solve(adj, k, 1).states
            assert np.array_equal(s.states(), want), (seed, case, n, k, tuning)
    finally:
        s.close()


@pytest.mark.parametrize("seed", [21])
def test_random_wide_rows_and_settings_match_oracle(oracle, seed):
    rng = np.random.default_rng(seed)
    s = cra.Solver(0)
    try:
        for case in range(12):
            n, adj = random_graph(rng, 65, 130)
            k = int(rng.integers(1, 3))
            tuning = {name: int(rng.choice(values)) for name, values in CHOICES.items()}
            print(f"random wide case {seed}/{case}: n={n} k={k} {tuning}", flush=True)
            s.set_tuning(**tuning)
            s.load_graph_matrix(adj, k, 1)
            s.build_aux_graph()
            s.initialize_captures()
            s.main_loop()
            want = oracle.v1_levelsync(adj, k, nthreads=8)[0]
            assert np.array_equal(s.states(), want), (seed, case, n, k, tuning)
    finally:
        s.close()


@pytest.mark.parametrize("seed", [31])
def test_wide_rows_with_three_cops_match_oracle(oracle, seed):
    """Two-word robber rows at k = 3 (66 and 80 vertices): the default settings,
    every level pulled (stage 1 by tiles of consecutive multisets, the last two
    stages in one k_pull_pair launch, or stage by stage) and the depth log with
    the fused last stage; each table equals the oracle's."""
    rng = np.random.default_rng(seed)
    cases = [({}, 66), ({"pull_scale": 0}, 66), ({"pull_scale": 0, "pull_pair": 0}, 80),
             ({"pull_scale": 0, "depth_log_min": 0}, 80)]
    s = cra.Solver(0)
    try:
        for case, (tuning, n) in enumerate(cases):
            while True:
                adj = np.triu((rng.random((n, n)) < 0.06).astype(np.uint8), 1)
                adj = adj | adj.T
                if adj.sum(1).min() > 0:
                    break
            print(f"wide k=3 case {seed}/{case}: n={n} {tuning}", flush=True)
            s.set_tuning(pull_scale=-1, pull_pair=1, depth_log_min=1 << 24)
            s.set_tuning(**tuning)
            s.load_graph_matrix(adj, 3, 1)
            s.build_aux_graph()
            s.initialize_captures()
            s.main_loop()
            want = oracle.v1_levelsync(adj, 3, nthreads=8)[0]
            assert np.array_equal(s.states(), want), (seed, case, n, tuning)
    finally:
        s.close()
