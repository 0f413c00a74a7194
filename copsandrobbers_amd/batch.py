"""Many small plain games in one launch, and the continuous-batching server.

Thin ctypes layer over cr_batch_* / cr_server_* (include/crsolve.h). Every
result equals the six-step solve of that graph: the same table (compared by
``table_hash`` or by the table itself), passes and verdict.
"""
import ctypes

import numpy as np

from . import MAX_K, SolverError, _check, _u8, lib

MAX_N = 64
MAX_BATCH_K = 4
TUNING = {"threads": 1, "blocks_per_cu": 2, "push_pct": 3}


class GraphResult(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("status", ctypes.c_int32), ("win", ctypes.c_int32),
                ("plys", ctypes.c_int32), ("n_passes", ctypes.c_int32), ("best_cid", ctypes.c_uint64),
                ("positions", ctypes.c_uint8 * MAX_K), ("base_captures", ctypes.c_uint64),
                ("won_states", ctypes.c_uint64), ("table_hash", ctypes.c_uint64),
                ("table_bytes", ctypes.c_uint64), ("latency_ms", ctypes.c_double)]


_ready = False


def _lib():
    global _ready
    L = lib()
    if not _ready:
        vp = ctypes.c_void_p
        i32, u64 = ctypes.c_int32, ctypes.c_uint64
        u64p = ctypes.POINTER(ctypes.c_uint64)
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.cr_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.cr_batch_destroy.argtypes = [vp]
        L.cr_batch_destroy.restype = None
        L.cr_batch_solve.argtypes = [vp, i32, i32p, i32p, ctypes.POINTER(ctypes.c_uint8), u64p,
                                     ctypes.POINTER(GraphResult), ctypes.POINTER(ctypes.c_uint8), u64p]
        L.cr_batch_pass_counts.argtypes = [vp, i32, u64p, i32, i32p]
        L.cr_batch_set_tuning.argtypes = [vp, i32, ctypes.c_int64]
        L.cr_server_set_tuning.argtypes = [vp, i32, ctypes.c_int64]
        L.cr_batch_last_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_double), i32p]
        L.cr_server_create.argtypes = [ctypes.c_int, u64, i32, i32, ctypes.POINTER(vp)]
        L.cr_server_submit.argtypes = [vp, i32, i32, ctypes.POINTER(ctypes.c_uint8), u64, u64p]
        L.cr_server_submit_many.argtypes = [vp, i32, i32p, i32p, ctypes.POINTER(ctypes.c_uint8), u64p, u64p, u64p,
                                            i32p]
        L.cr_server_poll.argtypes = [vp, ctypes.POINTER(GraphResult), i32, i32, i32p]
        L.cr_server_stats.argtypes = [vp, u64p, u64p, u64p, u64p]
        L.cr_server_destroy.argtypes = [vp]
        L.cr_server_destroy.restype = None
        L.cr_clock_ns.restype = u64
        L.cr_server_retain_tables.argtypes = [vp, i32]
        L.cr_server_copy_table.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_uint8), u64, u64p]
        L.cr_server_release_table.argtypes = [vp, u64]
        L.cr_server_table_stats.argtypes = [vp, u64p, u64p, u64p]
        L.cr_batch_solve_noalloc.argtypes = [ctypes.c_int, i32, i32p, i32p, ctypes.POINTER(ctypes.c_uint8), u64p,
                                             ctypes.POINTER(GraphResult), ctypes.c_void_p, u64, u64p]
        _ready = True
    return L


def to_dict(r):
    k = MAX_K
    return {"ticket": int(r.ticket), "status": int(r.status), "win": bool(r.win), "plys": int(r.plys),
            "n_passes": int(r.n_passes), "best_cid": int(r.best_cid), "positions": tuple(r.positions[:k]),
            "base_captures": int(r.base_captures), "won_states": int(r.won_states),
            "table_hash": int(r.table_hash), "table_bytes": int(r.table_bytes), "latency_ms": float(r.latency_ms)}


def clock_ns():
    return int(_lib().cr_clock_ns())


class Packed:
    """Graphs packed once for repeated launches: n int32[], k int32[], the
    concatenated adjacency matrices and their offsets (what cr_batch_solve
    reads)."""
    __slots__ = ("ns", "ks", "adj", "offs")

    def __init__(self, graphs):
        self.ns, self.ks, self.adj, self.offs = _pack(graphs)

    def __len__(self):
        return len(self.ns)


def pack(graphs):
    """graphs: list of (adj uint8[n, n], k) -> Packed."""
    return Packed(graphs)


# GraphResult as a numpy record (results="array" returns the launch's results
# in this layout without a Python object per graph)
RESULT_DTYPE = np.dtype({"names": [f for f, _ in GraphResult._fields_],
                         "formats": [np.uint64, np.int32, np.int32, np.int32, np.int32, np.uint64, (np.uint8, MAX_K),
                                     np.uint64, np.uint64, np.uint64, np.uint64, np.float64],
                         "offsets": [getattr(GraphResult, f).offset for f, _ in GraphResult._fields_],
                         "itemsize": ctypes.sizeof(GraphResult)})


def _pack(graphs):
    """(n int32[], k int32[], concatenated matrices, offsets uint64[])."""
    ns = np.array([np.asarray(a).shape[0] for a, _ in graphs], np.int32)
    ks = np.array([int(k) for _, k in graphs], np.int32)
    offs = np.zeros(len(graphs), np.uint64)
    if len(graphs):
        offs[1:] = np.cumsum(ns.astype(np.uint64) ** 2)[:-1]
    adj = np.concatenate([np.ascontiguousarray(a, np.uint8).ravel() for a, _ in graphs]) if graphs else \
        np.zeros(1, np.uint8)
    if adj.size == 0:
        adj = np.zeros(1, np.uint8)
    return ns, ks, adj, offs


NOALLOC_SCRATCH_BYTES = 256 << 20
NOALLOC_MAX_GRAPHS = 4096


def solve_noalloc(graphs, d_tables=None, tables_cap=0, device=0, raise_on_error=True):
    """cr_batch_solve_noalloc: one launch over `graphs` with no allocation by
    the library. d_tables is a device address (int) of tables_cap bytes owned
    by the caller (e.g. a torch uint8 tensor's data_ptr()); graph i's table
    lands at offsets[i]. With d_tables None only the offsets are computed
    (offsets[-1] = bytes needed). Returns (results as dicts, offsets)."""
    L = _lib()
    ns, ks, adj, offs = _pack(graphs)
    count = len(graphs)
    res = (GraphResult * max(count, 1))()
    toff = np.zeros(count + 1, np.uint64)
    i32p = ctypes.POINTER(ctypes.c_int32)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    rc = L.cr_batch_solve_noalloc(device, count, ns.ctypes.data_as(i32p), ks.ctypes.data_as(i32p), _u8(adj),
                                  offs.ctypes.data_as(u64p), res, d_tables, tables_cap, toff.ctypes.data_as(u64p))
    if rc != 0 and (raise_on_error or rc != -1):
        _check(rc)
    out = [to_dict(res[i]) for i in range(count)]
    for i, r in enumerate(out):
        r["positions"] = r["positions"][:int(ks[i])]
    return out, toff


class Batch:
    """One launch over many graphs: ``solve([(adj, k), ...])``."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(_lib().cr_batch_create(device, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib().cr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def solve(self, graphs, tables=False, raise_on_error=True, results="dicts"):
        """graphs: list of (adj uint8[n, n], k), or a Packed. Returns the
        results as a list of dicts, or with results="array" as a RESULT_DTYPE
        record array (positions hold MAX_K entries, the first k used); plus
        the list of state tables when tables=True."""
        count = len(graphs)
        ns, ks, adj, offs = (graphs.ns, graphs.ks, graphs.adj, graphs.offs) if isinstance(graphs, Packed) else _pack(graphs)
        res = (GraphResult * max(count, 1))()
        tab = None
        toffs = None
        if tables:
            from math import comb
            sizes = [2 * comb(int(n) + int(k) - 1, int(k)) * int(n) if 1 <= k <= MAX_BATCH_K and 1 <= n <= MAX_N else 0
                     for n, k in zip(ns, ks)]
            toffs = np.zeros(count, np.uint64)
            if count:
                toffs[1:] = np.cumsum(np.array(sizes, np.uint64))[:-1]
            tab = np.zeros(max(1, int(sum(sizes))), np.uint8)
        L = _lib()
        i32p = ctypes.POINTER(ctypes.c_int32)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        rc = L.cr_batch_solve(self._h, count, ns.ctypes.data_as(i32p), ks.ctypes.data_as(i32p), _u8(adj),
                              offs.ctypes.data_as(u64p), res, _u8(tab) if tab is not None else None,
                              toffs.ctypes.data_as(u64p) if toffs is not None else None)
        if rc != 0 and (raise_on_error or rc != -1):
            raise SolverError(rc, L.cr_last_error().decode(errors="replace"))
        if results == "array":
            arr = np.frombuffer(res, dtype=RESULT_DTYPE, count=count).copy()
            return (arr, [tab[int(toffs[i]):int(toffs[i]) + int(arr[i]["table_bytes"])] for i in range(count)]) \
                if tables else arr
        out = [to_dict(res[i]) for i in range(count)]
        for i, r in enumerate(out):
            r["positions"] = r["positions"][:int(ks[i])]
        if tables:
            tlist = [tab[int(toffs[i]):int(toffs[i]) + out[i]["table_bytes"]] for i in range(count)]
            return out, tlist
        return out

    def set_tuning(self, **knobs):
        """Speed settings (results never change): threads (0, 256, 512, 1024), blocks_per_cu,
        push_pct (0 = every level pulls)."""
        for name, value in knobs.items():
            _check(_lib().cr_batch_set_tuning(self._h, TUNING[name], int(value)))

    def pass_counts(self, index):
        buf = np.zeros(4096, np.uint64)
        n = ctypes.c_int32()
        _check(_lib().cr_batch_pass_counts(self._h, index, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                           4096, ctypes.byref(n)))
        return [int(x) for x in buf[:n.value]]

    def last_timing(self):
        ms = ctypes.c_double()
        g = ctypes.c_int32()
        _check(_lib().cr_batch_last_timing(self._h, ctypes.byref(ms), ctypes.byref(g)))
        return float(ms.value), int(g.value)


class Server:
    """Continuous batching: submit requests as they arrive, poll results."""

    def __init__(self, device=0, hbm_budget=64 << 30, max_batch=1024, streams=4):
        h = ctypes.c_void_p()
        _check(_lib().cr_server_create(device, int(hbm_budget), int(max_batch), int(streams), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib().cr_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_tuning(self, **knobs):
        for name, value in knobs.items():
            _check(_lib().cr_server_set_tuning(self._h, TUNING[name], int(value)))

    def retain_tables(self, on=True):
        """Keep each finished request's state table resident (within the HBM budget)."""
        _check(_lib().cr_server_retain_tables(self._h, 1 if on else 0))

    def table(self, ticket):
        """The resident table of a finished request (uint8, the reference's layout)."""
        n = ctypes.c_uint64()
        _check(_lib().cr_server_copy_table(self._h, int(ticket), None, 0, ctypes.byref(n)))
        out = np.empty(max(int(n.value), 1), np.uint8)
        _check(_lib().cr_server_copy_table(self._h, int(ticket), _u8(out), out.size, ctypes.byref(n)))
        return out[:int(n.value)]

    def release(self, ticket):
        _check(_lib().cr_server_release_table(self._h, int(ticket)))

    def table_stats(self):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(_lib().cr_server_table_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"resident_tables": int(a.value), "resident_bytes": int(b.value), "evicted_launches": int(c.value)}

    def submit(self, adj, k, arrival_ns=0):
        adj = np.ascontiguousarray(adj, np.uint8)
        t = ctypes.c_uint64()
        _check(_lib().cr_server_submit(self._h, adj.shape[0], int(k), _u8(adj), int(arrival_ns), ctypes.byref(t)))
        return int(t.value)

    def submit_many(self, graphs, arrival_ns=None):
        """graphs: list of (adj, k); arrival_ns: per-request arrival times (or
        None = now). Returns (tickets, statuses); a rejected request has
        ticket 0 and its status code."""
        count = len(graphs)
        if count == 0:
            return [], []
        ns, ks, adj, offs = _pack(graphs)
        arr = None if arrival_ns is None else np.ascontiguousarray(arrival_ns, np.uint64)
        tickets = np.zeros(count, np.uint64)
        status = np.zeros(count, np.int32)
        i32p = ctypes.POINTER(ctypes.c_int32)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        rc = _lib().cr_server_submit_many(self._h, count, ns.ctypes.data_as(i32p), ks.ctypes.data_as(i32p),
                                          _u8(adj), offs.ctypes.data_as(u64p),
                                          arr.ctypes.data_as(u64p) if arr is not None else None,
                                          tickets.ctypes.data_as(u64p), status.ctypes.data_as(i32p))
        if rc not in (0, -1):
            _check(rc)
        return [int(t) for t in tickets], [int(x) for x in status]

    def poll(self, cap=4096, wait_us=0):
        res = (GraphResult * cap)()
        n = ctypes.c_int32()
        _check(_lib().cr_server_poll(self._h, res, cap, int(wait_us), ctypes.byref(n)))
        return [to_dict(res[i]) for i in range(n.value)]

    def stats(self):
        v = [ctypes.c_uint64() for _ in range(4)]
        _check(_lib().cr_server_stats(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("submitted", "finished", "launches", "peak_bytes"), (int(x.value) for x in v)))
