"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, always as the checker of the HIP path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MAX_ROUND = 127

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        i32, i64, u64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
        L.or_parse_matrix.argtypes = [u8p, u64, u8p, i32]
        L.or_parse_matrix.restype = i32
        L.or_config_count.argtypes = [i32, i32]
        L.or_config_count.restype = u64
        L.or_enum_configs.argtypes = [i32, i32, u8p, u64]
        L.or_enum_configs.restype = i32
        L.or_rank.argtypes = [u8p, i32, i32]
        L.or_rank.restype = u64
        L.or_transitions.argtypes = [i32, u8p, i32, i32, u64p, u64p, u64]
        L.or_transitions.restype = i64
        L.or_v1_solve.argtypes = [i32, u8p, i32, i32, i32, i32, u8p, u64p, u64p, i32,
                                  ctypes.POINTER(i32)]
        L.or_v1_solve.restype = i32
        L.or_verdict.argtypes = [i32, i32, i32, u8p, ctypes.POINTER(i64), ctypes.POINTER(i32)]
        L.or_verdict.restype = i32
        L.or_v2_solve.argtypes = [i32, u8p, i32, i32, u8p, u64p, u64p, i32, ctypes.POINTER(i32)]
        L.or_v2_solve.restype = i32
        L.or_v1_levelsync.argtypes = [i32, u8p, i32, i32, u8p, u64p, u64p, u64p]
        L.or_v1_levelsync.restype = i32
        L.or_retro_edges.argtypes = [i32, u8p, i32, u8p, u64p, u64p, u64p]
        L.or_retro_edges.restype = i32
        L.or_pass_counts_from_depths.argtypes = [i32, i32, u8p, u64p, i32]
        L.or_pass_counts_from_depths.restype = i32
        L.or_unrank.argtypes = [u64, i32, i32, u8p]
        L.or_unrank.restype = i32
        L.or_check_states.argtypes = [i32, u8p, i32, u8p, u64p, u64, i32, u64p, u64p]
        L.or_check_states.restype = i32
        L.or_table_summary.argtypes = [i32, i32, u8p, i32, u64p, ctypes.POINTER(i64), ctypes.POINTER(i32)]
        L.or_table_summary.restype = i32
        dp = ctypes.POINTER(ctypes.c_double)
        L.or_v1_levelsync_capped.argtypes = [i32, u8p, i32, i32, ctypes.c_double, u8p, u64p, u64p,
                                             ctypes.POINTER(i32), dp, dp]
        L.or_v1_levelsync_capped.restype = i32
        _lib = L
    return _lib


def _p8(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _p64(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def parse_matrix(text: bytes):
    """Graph.cpp:15-79. Returns (n, adj uint8[n, n])."""
    buf = np.frombuffer(text, dtype=np.uint8).copy() if text else np.zeros(1, np.uint8)
    n = lib().or_parse_matrix(_p8(buf), len(text), None, 0)
    if n == 0:
        return 0, np.zeros((0, 0), np.uint8)
    adj = np.zeros((n, n), np.uint8)
    lib().or_parse_matrix(_p8(buf), len(text), _p8(adj), n)
    return n, adj


def load_matrix(path):
    with open(path, "rb") as f:
        return parse_matrix(f.read())


def config_count(n, k):
    return int(lib().or_config_count(n, k))


def configs(n, k):
    C = config_count(n, k)
    out = np.zeros((C, k), np.uint8)
    assert lib().or_enum_configs(n, k, _p8(out), C) == 0
    return out


def rank(tup, n):
    t = np.ascontiguousarray(np.asarray(tup, np.uint8))
    return int(lib().or_rank(_p8(t), n, len(t)))


def transitions(adj, k, cop_stay=0):
    """AuxGraph.h:179-275: (heads uint64[C+1], transitions uint64[...] = nextId*N)."""
    adj = np.ascontiguousarray(adj, np.uint8)
    n = adj.shape[0]
    C = config_count(n, k)
    heads = np.zeros(C + 1, np.uint64)
    cap = 1 << 16
    while True:
        trans = np.zeros(cap, np.uint64)
        tot = lib().or_transitions(n, _p8(adj), k, cop_stay, _p64(heads), _p64(trans), cap)
        if tot >= 0:
            return heads, trans[:tot].copy()
        need = int(heads[-1])
        if need == 0:
            raise RuntimeError("or_transitions failed")
        cap = need


class Solved:
    def __init__(self, states, base_caps, pass_counts, columns, n, k):
        self.states = states
        self.base_caps = base_caps
        self.pass_counts = pass_counts
        self.columns = columns
        self.n = n
        self.k = k

    def verdict(self):
        best = ctypes.c_int64()
        plys = ctypes.c_int32()
        lib().or_verdict(self.n, self.k, self.columns, _p8(self.states), ctypes.byref(best),
                         ctypes.byref(plys))
        return int(best.value), int(plys.value)


def v1_solve(adj, k, p=1, cop_stay=0, robber_stay=0, max_passes=8192):
    adj = np.ascontiguousarray(adj, np.uint8)
    n = adj.shape[0]
    C = config_count(n, k)
    states = np.zeros(C * n * 2 * p, np.uint8)
    base = ctypes.c_uint64()
    pc = np.zeros(max_passes, np.uint64)
    npass = ctypes.c_int32()
    rc = lib().or_v1_solve(n, _p8(adj), k, p, cop_stay, robber_stay, _p8(states),
                           ctypes.byref(base), _p64(pc), max_passes, ctypes.byref(npass))
    if rc != 0:
        raise RuntimeError("or_v1_solve failed")
    return Solved(states, int(base.value), [int(x) for x in pc[:npass.value]], 2 * p, n, k)


def v2_solve(adj, k, p, max_passes=8192):
    adj = np.ascontiguousarray(adj, np.uint8)
    n = adj.shape[0]
    C = config_count(n, k)
    states = np.zeros(C * n, np.uint8)
    base = ctypes.c_uint64()
    pc = np.zeros(max_passes, np.uint64)
    npass = ctypes.c_int32()
    rc = lib().or_v2_solve(n, _p8(adj), k, p, _p8(states), ctypes.byref(base), _p64(pc),
                           max_passes, ctypes.byref(npass))
    if rc != 0:
        raise RuntimeError("or_v2_solve failed")
    return Solved(states, int(base.value), [int(x) for x in pc[:npass.value]], 1, n, k)


def v1_levelsync(adj, k, nthreads=1):
    """Returns (states, level_counts[127, 2], e_cop, e_rob)."""
    adj = np.ascontiguousarray(adj, np.uint8)
    n = adj.shape[0]
    C = config_count(n, k)
    states = np.zeros(C * n * 2, np.uint8)
    lv = np.zeros(2 * MAX_ROUND, np.uint64)
    ec = ctypes.c_uint64()
    er = ctypes.c_uint64()
    rc = lib().or_v1_levelsync(n, _p8(adj), k, nthreads, _p8(states), _p64(lv), ctypes.byref(ec),
                               ctypes.byref(er))
    if rc != 0:
        raise RuntimeError("or_v1_levelsync failed")
    return states, lv.reshape(MAX_ROUND, 2), int(ec.value), int(er.value)


def retro_edges(adj, k, states):
    adj = np.ascontiguousarray(adj, np.uint8)
    ec, er, w = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    assert lib().or_retro_edges(adj.shape[0], _p8(adj), k, _p8(states), ctypes.byref(ec),
                                ctypes.byref(er), ctypes.byref(w)) == 0
    return int(ec.value), int(er.value), int(w.value)


def unrank(rank_, n, k):
    out = np.zeros(k, np.uint8)
    if lib().or_unrank(rank_, n, k, _p8(out)) != 0:
        raise ValueError("rank out of range")
    return tuple(int(x) for x in out)


def check_states(adj, k, states, ids, nthreads=1):
    """Bellman check of a solved p = 1 table at the state indices `ids`.
    Returns (number of states that differ, index of one of them or None)."""
    adj = np.ascontiguousarray(adj, np.uint8)
    ids = np.ascontiguousarray(ids, np.uint64)
    bad, first = ctypes.c_uint64(), ctypes.c_uint64()
    rc = lib().or_check_states(adj.shape[0], _p8(adj), k, _p8(states), _p64(ids), len(ids), nthreads,
                               ctypes.byref(bad), ctypes.byref(first))
    if rc != 0:
        raise RuntimeError("or_check_states failed")
    return int(bad.value), (int(first.value) if bad.value else None)


def table_summary(n, k, states, nthreads=1):
    """(level_counts[127, 2], best_cid or -1, plys or -1) of a solved p = 1 table."""
    lv = np.zeros(2 * MAX_ROUND, np.uint64)
    best, plys = ctypes.c_int64(), ctypes.c_int32()
    if lib().or_table_summary(n, k, _p8(states), nthreads, _p64(lv), ctypes.byref(best),
                              ctypes.byref(plys)) != 0:
        raise RuntimeError("or_table_summary failed")
    return lv.reshape(MAX_ROUND, 2), int(best.value), int(plys.value)


def v1_levelsync_capped(adj, k, nthreads=1, time_cap_s=1e9, states=None):
    """Time-capped level-synchronous solve with on-the-fly moves. Returns
    (states, dict(e_cop, e_rob, levels_done, level_fraction, elapsed_s))."""
    adj = np.ascontiguousarray(adj, np.uint8)
    n = adj.shape[0]
    if states is None:
        states = np.empty(config_count(n, k) * n * 2, np.uint8)
    ec, er = ctypes.c_uint64(), ctypes.c_uint64()
    done, frac, el = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
    rc = lib().or_v1_levelsync_capped(n, _p8(adj), k, nthreads, float(time_cap_s), _p8(states),
                                      ctypes.byref(ec), ctypes.byref(er), ctypes.byref(done),
                                      ctypes.byref(frac), ctypes.byref(el))
    if rc != 0:
        raise RuntimeError("or_v1_levelsync_capped failed")
    return states, {"e_cop": int(ec.value), "e_rob": int(er.value), "levels_done": int(done.value),
                    "level_fraction": float(frac.value), "elapsed_s": float(el.value)}


def passes_from_levels(levels):
    """The per-pass counts of the p = 1 sweep from a (round, turn) histogram of
    marked states: pass i prints the cop-turn states of rounds 2i-2 (above 0)
    and 2i-1 plus the robber-turn states of rounds 2i-1 and 2i, and the sweep
    stops after the first pass that finds nothing (the rule of
    or_pass_counts_from_depths, applied to counts instead of a table)."""
    lv = np.asarray(levels, np.uint64).reshape(MAX_ROUND, 2)
    get = lambda d, t: int(lv[d, t]) if 0 <= d < MAX_ROUND else 0
    out = []
    i = 1
    while True:
        c0 = 2 * i - 2
        cnt = (get(c0, 0) if c0 > 0 else 0) + get(2 * i - 1, 0) + get(2 * i - 1, 1) + get(2 * i, 1)
        out.append(cnt)
        if cnt == 0:
            return out
        i += 1


def pass_counts_from_depths(n, k, states, max_passes=8192):
    pc = np.zeros(max_passes, np.uint64)
    np_ = lib().or_pass_counts_from_depths(n, k, _p8(states), _p64(pc), max_passes)
    return [int(x) for x in pc[:np_]]


# ---------------------------------------------------------------------------
# Round-solver outputs (legacy/src/k_cops_rounds.cpp, legacy/python/k_cops.py).
# Pure-Python restatements for small cases: the solve itself is v1_solve with
# cop_stay = robber_stay = 1 (k_cops.py:33-36 gives every vertex a self-loop;
# its synchronous passes, :86-131, are the level-synchronous depths), and a
# cop-turn state won at depth t has (t + 1) // 2 rounds left (k_cops.py:124).
# ---------------------------------------------------------------------------
def dp_table_rows(n, k, states, columns=2):
    """The dp_tables JSON rows: one dict per (cId, r) in cId order."""
    cfg = configs(n, k)
    st = np.asarray(states).reshape(len(cfg), n, columns)[:, :, 0].astype(np.int64)
    rows = []
    for c in range(len(cfg)):
        cops = [int(x) for x in cfg[c]]
        for r in range(n):
            b = int(st[c, r])
            win = bool(b & 1)
            rows.append({"cops": cops, "robber": r, "cop_wins": win,
                         "steps_to_win": ((b >> 1) + 1) // 2 if win else -1})
    return rows


def perfect_game(adj, k, states, cop_stay=1, robber_stay=1, columns=2, order="py"):
    """Perfect-game extraction over a solved table. The start is the universal
    configuration with the least worst case, lowest cId on ties, and the robber
    start with the most rounds (first in r order).
    order="py": legacy/python/k_cops.py:174-241 -- cop moves in
      itertools.product order over [neighbours ascending, stay], the robber's
      replies and moves over the same options (that producer iterates a set for
      the start, :142, so its own tie order there is arbitrary);
    order="cpp": legacy/src/k_cops_rounds.cpp:307-397 -- cop moves in ascending
      order of the next configuration (its CSR transition list), the robber's
      replies and moves over its neighbours only (no stay), no loop limit."""
    adj = np.asarray(adj)
    n = adj.shape[0]
    cfg = [tuple(int(x) for x in c) for c in configs(n, k)]
    index = {c: i for i, c in enumerate(cfg)}
    st = np.asarray(states).reshape(len(cfg), n, columns)[:, :, 0]

    def options(u, stay):
        o = [v for v in range(n) if adj[u, v]]
        if stay and not adj[u, u]:
            o.append(u)
        return o

    cop_opt = [options(u, cop_stay) for u in range(n)]
    rob_opt = [options(u, robber_stay and order == "py") for u in range(n)]
    won = lambda c, r: bool(st[index[c], r] & 1)
    rounds = lambda c, r: ((int(st[index[c], r]) >> 1) + 1) // 2
    best, best_worst = None, None
    for c in cfg:
        if all(won(c, r) for r in range(n)):
            w = max(rounds(c, r) for r in range(n))
            if best_worst is None or w < best_worst:
                best, best_worst = c, w
    if best is None:
        return []
    r0, most = None, -1
    for r in range(n):
        if won(best, r) and rounds(best, r) > most:
            most, r0 = rounds(best, r), r
    path, cur, rob = [], best, r0
    import itertools
    loops = 0
    while True:
        if rob in cur:  # the robber walked onto a cop: the C++ producer records it (:325-327)
            if order == "cpp":
                path.append({"cops": list(cur), "robber": rob, "turn": "Game Over - Captured!"})
            break
        loops += 1
        if order == "py" and loops > 5 * n:
            break
        path.append({"cops": list(cur), "robber": rob, "turn": "Cop's Turn"})
        nxt, least = cur, None
        moves = [tuple(sorted(m)) for m in itertools.product(*[cop_opt[c] for c in cur])]
        if order == "cpp":
            moves = sorted(set(moves), key=lambda c: index[c])
        for nc in moves:
            if rob in nc:
                score, ok = 0, True
            else:
                score, ok = -1, True
                for v in rob_opt[rob]:
                    if not won(nc, v):
                        ok = False
                        break
                    score = max(score, rounds(nc, v))
            if ok and (least is None or score < least):
                least, nxt = score, nc
        cur = nxt
        if rob in cur:
            path.append({"cops": list(cur), "robber": rob, "turn": "Game Over - Captured!"})
            break
        path.append({"cops": list(cur), "robber": rob, "turn": "Robber's Turn"})
        go, most = rob, -1
        for v in rob_opt[rob]:
            if won(cur, v) and rounds(cur, v) > most:
                most, go = rounds(cur, v), v
        rob = go
    return path
