"""MI355X-native k-cop retrograde solver.

Python mirror of the reference's solver plug-in boundary: the six steps of
app/main.cpp:11-16 (loadGraphFile, buildAuxGraph, initializeCaptures,
mainLoop, findFinalResult, outputData) as methods of :class:`Solver`, each a
thin ctypes call into libcrsolve.so (include/crsolve.h). The HIP library is
the only compute path: importing this package on a machine without the built
library raises, and every step raises :class:`SolverError` with the library's
message when the C ABI reports a failure.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcrsolve.so")
MAX_K = 8
MAX_N = 255
MAX_ROUND = 127
VARIANT_COLUMNS = 1
VARIANT_MACRO = 2
TUNING = {"push_filter": 1, "push_blocks": 2, "row_blocks": 3, "levels_per_sync": 4, "pred_csr": 5,
          "pull_scale": 6, "pull_tiles": 7, "pull_tile_rows": 8,
          "pull_tile_blocks": 9, "depth_log": 10, "depth_log_min": 11,
          "pull_tile_chunk": 12, "fuse_finalize": 13, "pull_tile1": 14, "pull_gate": 15, "robber_open": 16, "log_blocks": 17, "scan_tile": 18, "push_dirty": 19,
          "pull_pair": 20, "pair_rows": 21, "pair_blocks": 22,
          "end_pass": 23}

STATUS = {
    0: "CR_OK", -1: "CR_ERR_ARG", -2: "CR_ERR_IO", -3: "CR_ERR_EMPTY_GRAPH",
    -4: "CR_ERR_GRAPH_LIMIT", -5: "CR_ERR_ISOLATED_VERTEX", -6: "CR_ERR_K_RANGE",
    -7: "CR_ERR_P_RANGE", -8: "CR_ERR_TOO_LARGE", -9: "CR_ERR_DEVICE", -10: "CR_ERR_ORDER",
    -11: "CR_ERR_UNSUPPORTED", -12: "CR_ERR_NO_DEVICE", -13: "CR_ERR_FORMAT", -14: "CR_ERR_CACHE_RULES",
}

# Names every build of include/crsolve.h exports.
EXPORTS = (
    "cr_solver_create", "cr_solver_destroy", "cr_last_error", "cr_parse_matrix",
    "cr_load_graph_file", "cr_load_graph_matrix", "cr_set_variant", "cr_set_rules", "cr_parse_edge_list",
    "cr_load_edge_file", "cr_build_aux_graph",
    "cr_initialize_captures", "cr_main_loop", "cr_find_final_result", "cr_save_aux_graph",
    "cr_get_info", "cr_get_stats", "cr_copy_states", "cr_copy_configs", "cr_level_counts",
    "cr_count_retro_edges", "cr_work_profile", "cr_level_rows", "cr_get_macro_stats", "cr_write_dp_table", "cr_extract_game",
    "cr_write_perfect_game", "cr_cop_number", "cr_count_transitions",
    "cr_set_tuning", "cr_set_kernel_timing", "cr_set_kernel_timing_filter", "cr_kernel_times", "cr_cache_key",
    "cr_batch_create", "cr_batch_destroy", "cr_batch_solve", "cr_batch_set_tuning", "cr_batch_pass_counts",
    "cr_batch_last_timing", "cr_server_set_tuning",
    "cr_server_create", "cr_server_submit", "cr_server_submit_many", "cr_server_poll", "cr_server_stats", "cr_server_destroy",
    "cr_clock_ns", "cr_batch_solve_noalloc", "cr_server_retain_tables", "cr_server_copy_table",
    "cr_server_release_table", "cr_server_table_stats", "cr_load_aux_graph", "cr_read_aux_header",
)


class SolverError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{STATUS.get(code, code)}: {message}")
        self.code = code


class Info(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("k", ctypes.c_int32), ("p", ctypes.c_int32),
                ("columns", ctypes.c_int32), ("variant", ctypes.c_int32),
                ("max_degree", ctypes.c_int32), ("config_count", ctypes.c_uint64),
                ("num_states", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64)]


class Verdict(ctypes.Structure):
    _fields_ = [("win", ctypes.c_int32), ("plys", ctypes.c_int32), ("best_cid", ctypes.c_uint64),
                ("positions", ctypes.c_uint8 * MAX_K)]


class GameStep(ctypes.Structure):
    _fields_ = [("cops", ctypes.c_uint8 * MAX_K), ("robber", ctypes.c_int32), ("turn", ctypes.c_int32)]


TURNS = ("Cop's Turn", "Robber's Turn", "Game Over - Captured!")
GAME_ORDER = {"py": 0, "cpp": 1}


class Stats(ctypes.Structure):
    _fields_ = [("base_captures", ctypes.c_uint64), ("won_states", ctypes.c_uint64),
                ("n_passes", ctypes.c_int32), ("n_levels", ctypes.c_int32),
                ("init_ms", ctypes.c_double), ("loop_ms", ctypes.c_double),
                ("pred_entries", ctypes.c_uint64), ("build_ms", ctypes.c_double)]


class MacroStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int32), ("dag_nodes", ctypes.c_uint64), ("dag_edges", ctypes.c_uint64),
                ("build_moves", ctypes.c_uint64), ("leaf_moves", ctypes.c_uint64),
                ("up_edges", ctypes.c_uint64), ("build_ms", ctypes.c_double), ("passes_ms", ctypes.c_double)]


_lib = None


def library_path():
    return LIB_PATH


def lib():
    """Loads libcrsolve.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                              "(the HIP library is the only compute path)")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        u8p = ctypes.POINTER(ctypes.c_uint8)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        i32 = ctypes.c_int32
        L.cr_solver_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.cr_solver_destroy.argtypes = [vp]
        L.cr_solver_destroy.restype = None
        L.cr_last_error.restype = ctypes.c_char_p
        L.cr_parse_matrix.argtypes = [u8p, ctypes.c_uint64, u8p, i32]
        L.cr_parse_matrix.restype = i32
        L.cr_load_graph_file.argtypes = [vp, ctypes.c_char_p, i32, i32]
        L.cr_load_graph_matrix.argtypes = [vp, i32, u8p, i32, i32]
        L.cr_set_variant.argtypes = [vp, i32]
        L.cr_set_rules.argtypes = [vp, i32, i32]
        L.cr_parse_edge_list.argtypes = [u8p, ctypes.c_uint64, u8p, i32]
        L.cr_parse_edge_list.restype = i32
        L.cr_load_edge_file.argtypes = [vp, ctypes.c_char_p, i32, i32]
        L.cr_write_dp_table.argtypes = [vp, ctypes.c_char_p]
        L.cr_extract_game.argtypes = [vp, i32, ctypes.POINTER(GameStep), i32, ctypes.POINTER(i32)]
        L.cr_write_perfect_game.argtypes = [vp, i32, ctypes.c_char_p]
        L.cr_cop_number.argtypes = [vp, i32, ctypes.POINTER(i32), ctypes.POINTER(Verdict)]
        L.cr_build_aux_graph.argtypes = [vp]
        L.cr_initialize_captures.argtypes = [vp, u64p]
        L.cr_main_loop.argtypes = [vp, u64p, i32, ctypes.POINTER(i32)]
        L.cr_find_final_result.argtypes = [vp, ctypes.POINTER(Verdict)]
        L.cr_save_aux_graph.argtypes = [vp, ctypes.c_char_p]
        L.cr_load_aux_graph.argtypes = [vp, ctypes.c_char_p]
        L.cr_read_aux_header.argtypes = [ctypes.c_char_p, u64p]
        L.cr_get_info.argtypes = [vp, ctypes.POINTER(Info)]
        L.cr_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.cr_copy_states.argtypes = [vp, u8p, ctypes.c_uint64, ctypes.c_uint64]
        L.cr_copy_configs.argtypes = [vp, u8p, ctypes.c_uint64, ctypes.c_uint64]
        L.cr_level_counts.argtypes = [vp, u64p]
        L.cr_count_retro_edges.argtypes = [vp, u64p, u64p, u64p]
        L.cr_count_transitions.argtypes = [vp, u64p]
        L.cr_work_profile.argtypes = [vp, u64p, u64p, u8p]
        L.cr_level_rows.argtypes = [vp, u64p]
        L.cr_get_macro_stats.argtypes = [vp, ctypes.POINTER(MacroStats), u64p, u64p, i32]
        L.cr_set_tuning.argtypes = [vp, i32, ctypes.c_int64]
        L.cr_set_kernel_timing.argtypes = [vp, i32]
        L.cr_set_kernel_timing_filter.argtypes = [vp, ctypes.c_char_p]
        L.cr_kernel_times.argtypes = [vp, ctypes.c_char_p, i32, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_int64), i32, ctypes.POINTER(i32)]
        L.cr_cache_key.argtypes = [ctypes.c_char_p, i32, i32, i32, ctypes.c_char_p, i32]
        _lib = L
    return _lib


def _u8(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _u64(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def _check(rc):
    if rc != 0:
        raise SolverError(rc, lib().cr_last_error().decode(errors="replace"))
    return rc


def parse_matrix(text: bytes):
    """Graph::Graph parsing (lib/src/Graph.cpp:15-79) -> (n, adj uint8[n, n])."""
    buf = np.frombuffer(bytes(text) or b"\0", np.uint8).copy()
    n = lib().cr_parse_matrix(_u8(buf), len(text), None, 0)
    adj = np.zeros((max(n, 0), max(n, 0)), np.uint8)
    if n > 0:
        lib().cr_parse_matrix(_u8(buf), len(text), _u8(adj), n)
    return n, adj


def parse_edge_list(text: bytes):
    """legacy/util/edge_parser.py: 1-based undirected edge list -> (n, adj uint8[n, n])."""
    buf = np.frombuffer(bytes(text) or b"\0", np.uint8).copy()
    n = lib().cr_parse_edge_list(_u8(buf), len(text), None, 0)
    adj = np.zeros((max(n, 0), max(n, 0)), np.uint8)
    if n > 0:
        lib().cr_parse_edge_list(_u8(buf), len(text), _u8(adj), n)
    return n, adj


def cache_key(graph_path, k, p, variant=VARIANT_COLUMNS):
    buf = ctypes.create_string_buffer(4096)
    n = lib().cr_cache_key(str(graph_path).encode(), k, p, variant, buf, len(buf))
    if n < 0:
        _check(n)
    return buf.value.decode()


class Solver:
    """One device solver; the reference's six plug-in steps as methods."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().cr_solver_create(device, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().cr_solver_destroy(self._h)
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

    # ---- the six steps (app/main.cpp:11-16) ----
    def load_graph_file(self, filename, k, p=1, variant=VARIANT_COLUMNS):
        _check(lib().cr_load_graph_file(self._h, str(filename).encode(), k, p))
        _check(lib().cr_set_variant(self._h, variant))

    def load_graph_matrix(self, adj, k, p=1, variant=VARIANT_COLUMNS):
        adj = np.ascontiguousarray(adj, np.uint8)
        _check(lib().cr_load_graph_matrix(self._h, adj.shape[0], _u8(adj), k, p))
        _check(lib().cr_set_variant(self._h, variant))

    def set_rules(self, cop_stay=False, robber_stay=False):
        """Stay moves for cops / robber (AuxGraph::setSelfEdges); next build."""
        _check(lib().cr_set_rules(self._h, int(bool(cop_stay)), int(bool(robber_stay))))

    def load_edge_file(self, filename, k, p=1, variant=VARIANT_COLUMNS):
        _check(lib().cr_load_edge_file(self._h, str(filename).encode(), k, p))
        _check(lib().cr_set_variant(self._h, variant))

    def build_aux_graph(self):
        _check(lib().cr_build_aux_graph(self._h))

    def initialize_captures(self):
        base = ctypes.c_uint64()
        _check(lib().cr_initialize_captures(self._h, ctypes.byref(base)))
        return int(base.value)

    def main_loop(self, max_passes=16384):
        pc = np.zeros(max_passes, np.uint64)
        n = ctypes.c_int32()
        _check(lib().cr_main_loop(self._h, _u64(pc), max_passes, ctypes.byref(n)))
        return [int(x) for x in pc[:min(n.value, max_passes)]]

    def find_final_result(self):
        v = Verdict()
        _check(lib().cr_find_final_result(self._h, ctypes.byref(v)))
        k = self.info().k
        if not v.win:
            return None
        return {"plys": int(v.plys), "best_cid": int(v.best_cid),
                "positions": tuple(int(x) for x in v.positions[:k])}

    def output_data(self, path):
        _check(lib().cr_save_aux_graph(self._h, str(path).encode()))

    def load_aux_graph(self, path):
        """CacheManager::loadAuxGraph (CacheManager.h:77-141) in place of
        build_aux_graph: the blob's states become the table (readable and
        verdict-ready; initialize_captures resets it)."""
        _check(lib().cr_load_aux_graph(self._h, str(path).encode()))

    # ---- accessors ----
    def info(self):
        i = Info()
        _check(lib().cr_get_info(self._h, ctypes.byref(i)))
        return i

    def stats(self):
        s = Stats()
        _check(lib().cr_get_stats(self._h, ctypes.byref(s)))
        return s

    def states(self, offset=0, count=None):
        total = self.info().num_states
        if count is None:
            count = total - offset
        out = np.empty(count, np.uint8)
        _check(lib().cr_copy_states(self._h, _u8(out), offset, count))
        return out

    def configs(self, first=0, count=None):
        info = self.info()
        if count is None:
            count = info.config_count - first
        out = np.empty((count, info.k), np.uint8)
        _check(lib().cr_copy_configs(self._h, _u8(out), first, count))
        return out

    def level_rows(self):
        """Per depth: configurations with cop-turn / robber-turn states of that
        depth (the level lists), shape (127, 2)."""
        out = np.zeros(2 * (MAX_ROUND), np.uint64)
        _check(lib().cr_level_rows(self._h, _u64(out)))
        return out.reshape(MAX_ROUND, 2)

    def level_counts(self):
        out = np.zeros(2 * MAX_ROUND, np.uint64)
        _check(lib().cr_level_counts(self._h, _u64(out)))
        return out.reshape(MAX_ROUND, 2)

    def retro_edges(self):
        ec, er, w = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().cr_count_retro_edges(self._h, ctypes.byref(ec), ctypes.byref(er),
                                          ctypes.byref(w)))
        return int(ec.value), int(er.value), int(w.value)

    def work_profile(self):
        """(e_cop[127], e_rob[127], pulled[127]) by depth of the solved table."""
        ec = np.zeros(MAX_ROUND, np.uint64)
        er = np.zeros(MAX_ROUND, np.uint64)
        pl = np.zeros(MAX_ROUND, np.uint8)
        _check(lib().cr_work_profile(self._h, _u64(ec), _u64(er), _u8(pl)))
        return ec, er, pl

    def macro_stats(self):
        """Trajectory DAG and work of the last limited-visibility main loop:
        dict with steps, dag_nodes, dag_edges, build_moves, leaf_moves,
        up_edges, build_ms, passes_ms, nodes_per_step, edges_per_step."""
        m = MacroStats()
        nodes = np.zeros(128, np.uint64)
        edges = np.zeros(128, np.uint64)
        _check(lib().cr_get_macro_stats(self._h, ctypes.byref(m), _u64(nodes), _u64(edges), 128))
        out = {name: getattr(m, name) for name, _ in MacroStats._fields_}
        out["nodes_per_step"] = [int(x) for x in nodes[:m.steps]]
        out["edges_per_step"] = [int(x) for x in edges[:max(m.steps - 1, 0)]]
        return out

    def write_dp_table(self, path):
        _check(lib().cr_write_dp_table(self._h, str(path).encode()))

    def extract_game(self, order="py", cap=8192):
        """The perfect game as [{"cops": [...], "robber": r, "turn": "..."}, ...];
        order "py" (legacy/python/k_cops.py) or "cpp" (legacy/src/k_cops_rounds.cpp)."""
        steps = (GameStep * cap)()
        n = ctypes.c_int32()
        _check(lib().cr_extract_game(self._h, GAME_ORDER[order], steps, cap, ctypes.byref(n)))
        k = self.info().k
        return [{"cops": [int(x) for x in steps[i].cops[:k]], "robber": int(steps[i].robber),
                 "turn": TURNS[steps[i].turn]} for i in range(min(n.value, cap))]

    def write_perfect_game(self, path, order="py"):
        _check(lib().cr_write_perfect_game(self._h, GAME_ORDER[order], str(path).encode()))

    def cop_number(self, k_max=4):
        """Least k <= k_max with a WIN (0 if none) and that solve's verdict."""
        kk = ctypes.c_int32()
        v = Verdict()
        _check(lib().cr_cop_number(self._h, k_max, ctypes.byref(kk), ctypes.byref(v)))
        verdict = None
        if v.win:
            verdict = {"plys": int(v.plys), "best_cid": int(v.best_cid),
                       "positions": tuple(int(x) for x in v.positions[:max(kk.value, 1)])}
        return int(kk.value), verdict

    def count_transitions(self):
        t = ctypes.c_uint64()
        _check(lib().cr_count_transitions(self._h, ctypes.byref(t)))
        return int(t.value)

    def set_tuning(self, **knobs):
        """Speed knobs (results never change): push_filter, push_blocks,
        row_blocks, levels_per_sync."""
        for name, value in knobs.items():
            _check(lib().cr_set_tuning(self._h, TUNING[name], int(value)))

    def set_kernel_timing(self, enabled=True, only=None):
        """Per-launch HIP-event timing of the main loop's kernels; `only` names
        the one kernel to time (None: all)."""
        _check(lib().cr_set_kernel_timing_filter(self._h, (only or "").encode()))
        _check(lib().cr_set_kernel_timing(self._h, 1 if enabled else 0))

    def kernel_times(self):
        names = ctypes.create_string_buffer(4096)
        ms = (ctypes.c_double * 32)()
        la = (ctypes.c_int64 * 32)()
        cnt = ctypes.c_int32()
        _check(lib().cr_kernel_times(self._h, names, 4096, ms, la, 32, ctypes.byref(cnt)))
        labels = names.raw.split(b"\0")[:cnt.value]
        return {labels[i].decode(): (float(ms[i]), int(la[i])) for i in range(cnt.value)}


def read_aux_header(path):
    """The 12 header words of a cache blob (CacheManager.h:49-58): k, N,
    columns, configCount, numStates, item size, cop and robber self-edge
    flags, and the four section offsets."""
    h = np.zeros(12, np.uint64)
    _check(lib().cr_read_aux_header(str(path).encode(), _u64(h)))
    return h


def solve(adj, k, p=1, variant=VARIANT_COLUMNS, device=0):
    """Runs the six steps on an in-memory matrix; returns a dict of results."""
    with Solver(device) as s:
        s.load_graph_matrix(adj, k, p, variant)
        s.build_aux_graph()
        base = s.initialize_captures()
        passes = s.main_loop()
        verdict = s.find_final_result()
        return {"base_caps": base, "passes": passes, "verdict": verdict,
                "states": s.states(), "stats": s.stats()}
