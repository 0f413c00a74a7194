"""cr_batch_solve_noalloc (the batch launch with every work buffer in the
library's file-scope device array) against cr_batch_solve on the same graphs:
equal results and equal table bytes, on two calls in a row, the second with a
smaller input, plus its size limits. The caller owns the output tables (a
torch device tensor here)."""
import sys

import numpy as np
import pytest
import torch

from copsandrobbers_amd import SolverError
from copsandrobbers_amd import batch as crb
from copsandrobbers_amd import workloads

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def strip(r):
    return {k: v for k, v in r.items() if k != "latency_ms"}


def check_call(ref_batch, graphs, d_tables):
    ref, tabs = ref_batch.solve(graphs, tables=True)
    _, offs = crb.solve_noalloc(graphs)  # sizing only
    need = int(offs[-1])
    assert need <= d_tables.numel()
    d_tables.fill_(0xAB)
    torch.cuda.synchronize()
    got, offs2 = crb.solve_noalloc(graphs, d_tables.data_ptr(), d_tables.numel())
    assert np.array_equal(offs, offs2)
    host = d_tables.cpu().numpy()
    for i, (a, b) in enumerate(zip(got, ref)):
        assert strip(a) == strip(b), i
        t = host[int(offs[i]):int(offs[i]) + a["table_bytes"]]
        assert np.array_equal(t, tabs[i]), i
    return need


def test_noalloc_equals_batch_solve_on_two_calls():
    say("no-alloc: call 1 (640 graphs), call 2 (96 graphs)")
    g1 = workloads.batch_graphs(640)
    g2 = workloads.batch_graphs(96, seed_base=7000)
    d_tables = torch.empty(int(crb.solve_noalloc(g1)[1][-1]), dtype=torch.uint8, device="cuda")
    with crb.Batch(0) as ref:
        n1 = check_call(ref, g1, d_tables)
        n2 = check_call(ref, g2, d_tables)
    assert n2 < n1


def test_noalloc_limits_and_invalid_graphs():
    say("no-alloc: limits")
    g = workloads.batch_graphs(8)
    need = int(crb.solve_noalloc(g)[1][-1])
    d_tables = torch.empty(need, dtype=torch.uint8, device="cuda")
    with pytest.raises(SolverError) as exc:
        crb.solve_noalloc(g, d_tables.data_ptr(), need - 1)
    assert exc.value.code == -8
    with pytest.raises(SolverError) as exc:
        crb.solve_noalloc(g * (crb.NOALLOC_MAX_GRAPHS // 8 + 1))
    assert exc.value.code == -8
    bad = list(g)
    iso = np.array(bad[3][0], np.uint8)
    iso[0, :] = 0
    iso[:, 0] = 0
    bad[3] = (iso, bad[3][1])
    got, offs = crb.solve_noalloc(bad, d_tables.data_ptr(), need, raise_on_error=False)
    assert got[3]["status"] == -5
    with crb.Batch(0) as ref:
        want = ref.solve(g)
    for i in range(8):
        if i != 3:
            assert strip(got[i]) == strip(want[i])
