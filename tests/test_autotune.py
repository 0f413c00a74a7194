"""The autotuner's search and database (CPU only; the device runs are in
test_gpu_autotune.py)."""
import json
import os
import re

import numpy as np
import pytest

from conftest import REPO

from copsandrobbers_amd import TUNING, autotune
from copsandrobbers_amd.batch import TUNING as BATCH_TUNING


def header_defaults():
    text = open(os.path.join(REPO, "include", "crsolve.h")).read()
    block = text[text.index("enum cr_tuning"):]
    block = block[:block.index("};")]
    out = {}
    for m in re.finditer(r"CR_TUNE_(\w+)\s*=\s*\d+,?\s*/\*(.*?)\*/", block, re.S):
        d = re.search(r"\(default (-?\d+)", m.group(2)) or re.search(r"(-?\d+) \(default\)", m.group(2))
        if d:
            out[m.group(1).lower()] = int(d.group(1))
    return out


def test_defaults_match_the_header():
    hd = header_defaults()
    for name, value in autotune.PLAIN_DEFAULTS.items():
        assert name in TUNING
        if name == "pull_scale":
            assert value == -1  # the size-dependent default
        else:
            assert hd[name] == value, name
    assert set(autotune.BATCH_DEFAULTS) == set(BATCH_TUNING)


def test_space_names_are_knobs():
    assert all(name in TUNING for name, _ in autotune.PLAIN_SPACE)
    assert all(name in BATCH_TUNING for name, _ in autotune.BATCH_SPACE)


def test_coordinate_descent_finds_the_separable_minimum():
    space = [("a", [1, 2, 3, 4]), ("b", [10, 20, 30]), ("c", [0, 1])]
    target = {"a": 3, "b": 10, "c": 1}
    defaults = {"a": 1, "b": 20, "c": 0}

    def cost(s):
        v = {**defaults, **s}
        return 1.0 + sum((v[k] - target[k]) ** 2 for k in target) * 0.1

    best, t, base, evals = autotune.coordinate_descent(cost, space)
    assert {**defaults, **best} == target
    assert t == 1.0 and base > t
    assert len(evals) >= 1 + 3 + 2 + 1


def test_coordinate_descent_ignores_gains_below_the_threshold():
    space = [("a", [1, 2])]
    best, t, base, _ = autotune.coordinate_descent(lambda s: 1.0 - 0.01 * (s.get("a", 0) == 2), space,
                                                   min_gain=0.02)
    assert best == {} and t == base == 1.0


def test_database_round_trip(tmp_path):
    db = str(tmp_path / "t.json")
    key = autotune.plain_key(200, 4, 1, 68685050, 4, 800)
    autotune.save_entry(key, autotune.entry({"pull_scale": 190}, 0.6, 0.7, [1, 2]), db)
    autotune.save_entry("batch/x", autotune.entry({"threads": 1024}, 0.02, 0.03, [1]), db)
    data = autotune.load_db(db)
    assert data[key]["settings"] == {"pull_scale": 190}
    assert abs(data[key]["best_ms"] - 600.0) < 1e-9
    assert set(data) == {key, "batch/x"}
    assert autotune.load_db(str(tmp_path / "missing.json")) == {}


def test_batch_key_is_the_shape():
    g = [(np.zeros((8, 8), np.uint8), 2), (np.zeros((12, 12), np.uint8), 3)]
    assert autotune.batch_key(g) == f"batch/g2/nmax12/kmax3/S{2 * 36 * 8 + 2 * 364 * 12}"


def test_shipped_database_entries_are_valid():
    data = autotune.load_db(autotune.DEFAULT_DB)
    cands = {**dict(autotune.PLAIN_SPACE), **dict(autotune.BATCH_SPACE)}
    for key, e in data.items():
        assert key.startswith(("plain/", "batch/"))
        for name, value in e["settings"].items():
            assert value in cands[name], (key, name)
        assert e["best_ms"] <= e["default_ms"]
