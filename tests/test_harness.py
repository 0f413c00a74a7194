"""The sweep runner (copsandrobbers_amd/harness.py) on the CPU: its stdout
parser on known lines, its case plan, and, with the reference binary built
from the reference's sources (oracle/build_ref.sh), a sweep of small recorded
cases whose parsed values equal the reference's own CSV rows
(tests/golden/csv/test_results.csv, written by util/test_k_cops_visibility.py)."""
import csv
import os

import pytest

from copsandrobbers_amd import harness

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
MATRICES = os.path.join(GOLDEN, "matrices")
REF_EXE = os.path.join(REPO, "oracle", "_ref", "k_cops_visibility")


def test_parse_stdout_reads_every_column():
    text = ("Initialized 123 base captures\nPass 1: 10\nPass 2: 4\nPass 3: 0\nRESULT: WIN\n"
            "Optimal Cop Start Positions: ( 0, 7 )\nCapture Time: 5 plys\n"
            "Main Loop ---=> 0.25 s\nTotal Footprint ---=> 2500000 B\nLoaded AuxGraph from cache\n")
    got = harness.parse_stdout(text)
    assert got == {"result": "WIN", "positions": "0, 7", "capture_plys": 5, "mainloop_s": 0.25, "mem_mb": 2.5,
                   "base_caps": 123, "cache_hit": True, "passes": 3, "solver_s": ""}


def test_parse_stdout_leaves_missing_columns_empty():
    got = harness.parse_stdout("RESULT: LOSS\n")
    assert got["result"] == "LOSS" and got["cache_hit"] is False
    assert all(got[c] == "" for c in ("positions", "capture_plys", "mainloop_s", "mem_mb", "base_caps", "passes"))
    assert harness.parse_stdout("")["result"] == "?"


def test_plan_is_cheapest_first_within_budget():
    cases = harness.plan_sweep(MATRICES, 3, 2, 1e6)
    assert cases
    assert [c.cost for c in cases] == sorted(c.cost for c in cases)
    for c in cases:
        assert c.cost == c.n ** c.k * c.n * 2 * c.p <= 1e6
        assert 1 <= c.k < c.n and 1 <= c.p <= 2
        assert c.graph not in harness.SLOW_GRAPHS
        assert c.n == harness.vertex_count(c.path)
    picked = harness.plan_sweep(MATRICES, 1, 1, 1e12, name_filters=("cycle",), include_slow=True)
    assert {c.graph for c in picked} >= {"cycle100"} and all("cycle" in c.graph for c in picked)


@pytest.mark.skipif(not os.path.exists(REF_EXE), reason="reference binary not built (oracle/build_ref.sh)")
def test_sweep_of_the_reference_binary_matches_its_recorded_rows(tmp_path):
    with open(os.path.join(GOLDEN, "csv", "test_results.csv")) as f:
        rows = [r for r in csv.DictReader(f)
                if r["status"] == "ok" and r["p"] == "1" and int(r["n"]) <= 12 and int(r["k"]) <= 2]
    assert len(rows) >= 5
    for row in rows[:12]:
        got = harness.run_solver(REF_EXE, os.path.join(MATRICES, row["graph"] + ".txt"), int(row["k"]), 1,
                                 timeout=60, cwd=str(tmp_path))
        assert got["status"] == "ok", got
        for col in ("result", "capture_plys", "passes", "base_caps", "positions"):
            assert str(got[col]) == row[col], (row["graph"], row["k"], col)
