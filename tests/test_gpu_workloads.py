"""Parity at the benchmarked sizes (p = 1, plain rules).

* ER G(200, 0.03), k = 3 (541 M states): the device table is byte-equal to the
  oracle's multithreaded level-synchronous solve, with equal pass counts,
  level histogram, retro-edge counts and verdict -- under the default speed
  settings and under settings that force the depth log, the two-level level
  scan, pull-only and push-only levels and the predecessor-CSR push at this size.
* Planar Delaunay graph, n = 200, 400 edges, k = 4 (27.5 G states): no CPU
  solve fits, so the table is checked through properties that fix it. Each
  sampled state must hold the byte its successors in the same table imply
  (oracle.check_states: captures, cop-turn 1 + min over team moves, robber-turn
  1 + max over moves, the 127 cap; a table satisfying this at every state is
  the reference's), on 10^6 random states and every state of 1,000 random
  configuration rows; the base captures equal their closed form; the pass
  counts and the verdict are re-derived from a full scan of the copied table.
  Solves through other code paths (separate finalize, a small depth log, a
  tiled level scan; other tile shapes, robber-step and push variants) give the
  same table byte for byte.
"""
import sys

import numpy as np
import pytest

from conftest import matrix

import copsandrobbers_amd as cra

pytestmark = pytest.mark.gpu

THREADS = 16


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def solve(s, path, k):
    s.load_graph_file(path, k, 1)
    s.build_aux_graph()
    base = s.initialize_captures()
    passes = s.main_loop()
    return base, passes, s.find_final_result()


@pytest.fixture(scope="module")
def er_k3(oracle):
    n, adj = oracle.load_matrix(matrix("er200_p0.03_s1"))
    say("oracle: er200 k=3 level-synchronous port")
    states, levels, e_cop, e_rob = oracle.v1_levelsync(adj, 3, nthreads=THREADS)
    return n, adj, states, levels, e_cop, e_rob


ER_SETTINGS = [{}, {"depth_log_min": 0, "scan_tile": 4, "pull_pair": 0}, {"pull_scale": 0}, {"pull_scale": 0, "pull_pair": 0},
               {"pull_scale": 10**9, "pred_csr": 1}]
ER_IDS = ["default", "depth_log_two_level_scan", "pull_every_level", "pull_stage_by_stage", "push_csr_every_level"]


@pytest.mark.parametrize("setting", ER_SETTINGS, ids=ER_IDS)
def test_er200_k3_table_equals_oracle_port(oracle, er_k3, setting):
    n, adj, want, levels, e_cop, e_rob = er_k3
    say("er200 k=3 on the device", setting)
    with cra.Solver(0) as s:
        s.set_tuning(**setting)
        base, passes, verdict = solve(s, matrix("er200_p0.03_s1"), 3)
        got = s.states()
        assert np.array_equal(got, want)
        assert np.array_equal(s.level_counts(), levels)
        assert passes == oracle.passes_from_levels(levels)
        assert base == int(levels[0, 0])
        assert s.retro_edges() == (e_cop, e_rob, int((want & 1).sum()))
    best, plys = oracle.Solved(want, base, passes, 2, n, 3).verdict()
    assert (verdict is None) == (best == -1)
    if verdict is not None:
        assert verdict["best_cid"] == best and verdict["plys"] == plys


@pytest.fixture(scope="module")
def planar_k4(oracle):
    """The default-setting device solve of planar200 k = 4, its table copied to the host."""
    path = matrix("planar200_e400_s1")
    k = 4
    say("planar200 k=4 on the device")
    with cra.Solver(0) as s:
        base, passes, verdict = solve(s, path, k)
        dev_levels = s.level_counts()
        say("copy table to host:", s.info().num_states, "bytes")
        states = s.states()
    return path, k, base, passes, verdict, dev_levels, states


def test_planar200_k4_table_satisfies_the_reference_equations(oracle, planar_k4):
    path, k, base, passes, verdict, dev_levels, states = planar_k4
    n, adj = oracle.load_matrix(path)
    C = oracle.config_count(n, k)
    S = C * n * 2
    assert states.size == S
    rng = np.random.default_rng(2026)
    ids = [rng.integers(0, S, 1_000_000, dtype=np.uint64)]
    rows = rng.choice(C, 1000, replace=False).astype(np.uint64)
    per_row = np.arange(2 * n, dtype=np.uint64)
    ids.append((rows[:, None] * np.uint64(2 * n) + per_row[None, :]).ravel())
    if verdict is not None:
        ids.append(np.uint64(verdict["best_cid"]) * np.uint64(2 * n) + per_row)
    ids = np.concatenate(ids)
    say("bellman check of", ids.size, "states")
    bad, first = oracle.check_states(adj, k, states, ids, nthreads=THREADS)
    assert bad == 0, f"{bad} states differ; first at index {first}"
    say("full scan for levels and verdict")
    levels, best, plys = oracle.table_summary(n, k, states, nthreads=THREADS)
    assert np.array_equal(levels, dev_levels)
    assert base == int(levels[0, 0]) == n * (C - oracle.config_count(n - 1, k))
    assert int(levels[0, 1]) == base
    assert passes == oracle.passes_from_levels(levels)
    assert (verdict is None) == (best == -1)
    if verdict is not None:
        assert verdict["best_cid"] == best and verdict["plys"] == plys
        assert verdict["positions"] == oracle.unrank(best, n, k)
    # the table's depths must fit under the cap and every depth level be contiguous
    depths = np.flatnonzero(levels.sum(axis=1))
    assert depths.max() <= 126 and np.array_equal(depths, np.arange(depths.max() + 1))


# Non-default paths at this size: (1) the separate finalize after a cop step
# that runs first, a depth log too small for a chunk of levels (applied at every
# synchronisation, and depths that do not fit write their bytes directly), the
# two-level level scan with 512-group tiles; (2) the first pull stage through
# k_pull_tile, 2-row tiles, the robber step from the new cop-turn states, a
# push that marks no groups, and a higher pull threshold (more push levels).
PLANAR_OTHER_PATHS = [{"fuse_finalize": 0, "depth_log": 4, "scan_tile": 512},
                      {"pull_tile1": 0, "pull_tile_rows": 2, "robber_open": 0, "push_dirty": 0, "pull_scale": 400}]
PLANAR_OTHER_IDS = ["finalize_small_log_tiled_scan", "tile_shapes_robber_by_new_more_pushes"]


@pytest.mark.parametrize("setting", PLANAR_OTHER_PATHS, ids=PLANAR_OTHER_IDS)
def test_planar200_k4_table_is_the_same_on_other_paths(planar_k4, setting):
    path, k, base, passes, verdict, dev_levels, want = planar_k4
    say("planar200 k=4 on the device", setting)
    with cra.Solver(0) as s:
        s.set_tuning(**setting)
        got_base, got_passes, got_verdict = solve(s, path, k)
        assert (got_base, got_passes, got_verdict) == (base, passes, verdict)
        assert np.array_equal(s.level_counts(), dev_levels)
        step = 1 << 30
        say("compare the table in slices of", step, "bytes")
        for off in range(0, want.size, step):
            cnt = min(step, want.size - off)
            got = s.states(off, cnt)
            if not np.array_equal(got, want[off:off + cnt]):
                bad = np.flatnonzero(got != want[off:off + cnt])
                pytest.fail(f"{bad.size} states differ in [{off}, {off + cnt}), the first at {off + int(bad[0])}")
