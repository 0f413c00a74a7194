"""Cache blob header reader (CacheManager.h:49-58, no device).

The inputs are the reference's own cache blobs, rebuilt byte for byte from
their recorded sections (tests/golden/ref/*.npz; the rebuilt bytes are
checked against the recorded sha256 of each blob file), and damaged copies
of them.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

import copsandrobbers_amd as cra

with open(os.path.join(GOLDEN, "ref_index.json")) as f:
    INDEX = json.load(f)

NAMES = sorted(k for k, v in INDEX.items() if v.get("arrays"))
SECTIONS = ("header", "configs", "heads", "transitions", "states")


def reference_blob(name):
    z = np.load(os.path.join(GOLDEN, "ref", INDEX[name]["arrays"]))
    return b"".join(z[s].tobytes() for s in SECTIONS), z


@pytest.mark.parametrize("name", NAMES)
def test_header_of_reference_blobs(name, tmp_path):
    data, z = reference_blob(name)
    assert hashlib.sha256(data).hexdigest() == INDEX[name]["blob_sha256"]
    path = tmp_path / "blob.bin"
    path.write_bytes(data)
    h = cra.read_aux_header(path)
    assert np.array_equal(h, z["header"])
    # the sections sit where the offsets say
    assert int(h[8]) == 96
    assert data[int(h[8]):int(h[9])] == z["configs"].tobytes()
    assert data[int(h[11]):] == z["states"].tobytes()


def code_of(path):
    with pytest.raises(cra.SolverError) as exc:
        cra.read_aux_header(path)
    return exc.value.code


def test_header_errors(tmp_path):
    data, z = reference_blob("v1__peterson__k3__p1")
    h = z["header"].copy()
    assert code_of(tmp_path / "missing.bin") == -2  # CR_ERR_IO
    short = tmp_path / "short.bin"
    short.write_bytes(data[:50])
    assert code_of(short) == -2
    cut = tmp_path / "cut.bin"
    cut.write_bytes(data[:int(h[11]) - 1])  # ends inside the transitions
    assert code_of(cut) == -13  # CR_ERR_FORMAT
    for field, value in ((8, 95), (9, 90), (10, int(h[11]) + 1), (11, len(data) + 1), (5, 0), (6, 2), (7, 3)):
        bad = h.copy()
        bad[field] = value
        p = tmp_path / f"bad{field}.bin"
        p.write_bytes(bad.tobytes() + data[96:])
        assert code_of(p) == -13, field
    # heads and transitions must be whole 8-byte words
    bad = h.copy()
    bad[10] = int(h[10]) + 4
    p = tmp_path / "odd.bin"
    p.write_bytes(bad.tobytes() + data[96:])
    assert code_of(p) == -13
