"""The autotuner on the device: whatever it picks, the tables stay the
reference's (the settings it searches never change a result), its choice is
stored and found again, and tuned batches equal untuned ones."""
import hashlib
import json
import os

import pytest

from conftest import GOLDEN, matrix

import copsandrobbers_amd as cra
from copsandrobbers_amd import autotune, workloads
from copsandrobbers_amd.batch import Batch

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "ref_index.json")) as f:
    INDEX = json.load(f)


def test_tuned_plain_settings_keep_the_reference_table(tmp_path):
    e = INDEX["v1__grid36__k3__p1"]
    print("autotune grid36 k=3", flush=True)
    db = str(tmp_path / "db.json")
    with cra.Solver(0) as s:
        s.load_graph_file(matrix("grid36"), 3, 1)
        s.build_aux_graph()
        space = [("pull_scale", [0, 150, 10**9]), ("pull_pair", [0, 2]), ("pull_tile_rows", [2, 8]),
                 ("levels_per_sync", [1, 16])]
        best, t, base, evals = autotune.tune_plain(s, reps=1, space=space, rounds=1)
        assert len(evals) == 1 + 3 + 2 + 2 + 2
        autotune.save_entry(autotune.plain_entry_key(s, 2 * 60), autotune.entry(best, t, base, evals), db)
        assert autotune.lookup_plain(s, 2 * 60, db) == best
        # every evaluated setting solves the same table; check the last applied and a few explicitly
        for settings in [best] + [ev[0] for ev in evals[1::3]]:
            autotune._apply_plain(s, settings)
            s.initialize_captures()
            passes = s.main_loop()
            assert hashlib.sha256(s.states().tobytes()).hexdigest() == e["states_sha256"], settings
            assert passes == e["passes"]


def test_tuned_batch_equals_default_batch(tmp_path):
    graphs = workloads.batch_graphs(128, seed_base=900)
    print("autotune batch of 128", flush=True)
    with Batch(0) as b:
        want = b.solve(graphs)
        best, t, base, evals = autotune.tune_batch(b, graphs, reps=1, rounds=1)
        got = b.solve(graphs)
    assert [r["table_hash"] for r in got] == [r["table_hash"] for r in want]
    assert [r["plys"] for r in got] == [r["plys"] for r in want]
    assert set(best) <= {"threads", "blocks_per_cu"}
