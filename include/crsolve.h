/*
 * crsolve.h -- C ABI of libcrsolve.so, the MI355X (gfx950) k-cop retrograde
 * solver.
 *
 * The reference has no FFI: its solver boundary is six C++ free functions that
 * app/main.cpp:11-16 declares and each solver translation unit defines
 * (solvers/k_cops_visibility.cpp, solvers/k_cops_visibility_v2.cpp). Every
 * entry point below replaces one of those steps, or one accessor that the
 * steps reach through AuxGraph/CacheManager; the comment on each cites the
 * reference interface it stands in for. Paths are relative to the reference
 * repository root.
 *
 * Conventions
 *   - Status: 0 = success, negative CR_ERR_* on failure (the reference's steps
 *     return bool true on failure, app/main.cpp:29). The message of the last
 *     failure on the calling thread is available from cr_last_error().
 *   - Ownership: the caller owns every pointer it passes in. The library owns
 *     the device memory of a cr_solver until cr_solver_destroy.
 *   - Threading: one cr_solver is not reentrant (like the reference's TU
 *     globals, k_cops_visibility.cpp:28-35); distinct solvers may be used
 *     from distinct threads. Each solver owns one HIP stream.
 *   - State tables use the reference's layout and byte encoding: index
 *     (cId*N + r)*columns + col, byte = marked | markedRound << 1
 *     (DataItem, k_cops_visibility.cpp:20-23 with GCC bit-field order).
 */
#ifndef CRSOLVE_H
#define CRSOLVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CR_MAX_K 8       /* cops per configuration supported on the device */
#define CR_MAX_N 255     /* vertex ids are uint8 with 255 as terminator (AdjacencyList.h) */
#define CR_MAX_ROUND 127 /* MAX_ROUND_COUNT, k_cops_visibility.cpp:26 */

enum cr_status {
    CR_OK = 0,
    CR_ERR_ARG = -1,             /* null pointer / bad argument */
    CR_ERR_IO = -2,              /* file unreadable or unwritable */
    CR_ERR_EMPTY_GRAPH = -3,     /* nodeCount == 0 (k_cops_visibility.cpp:47-50) */
    CR_ERR_GRAPH_LIMIT = -4,     /* N > 255 (k_cops_visibility_v2.cpp:348-351) */
    CR_ERR_ISOLATED_VERTEX = -5, /* a vertex with no move (AuxGraph.h:217-221 reads an
                                    uninitialised option there) */
    CR_ERR_K_RANGE = -6,         /* k <= 0, k > CR_MAX_K, or no configuration (AuxGraph.h:118-122) */
    CR_ERR_P_RANGE = -7,         /* p < 1 */
    CR_ERR_TOO_LARGE = -8,       /* state space exceeds device memory or 32-bit config ids */
    CR_ERR_DEVICE = -9,          /* HIP runtime error */
    CR_ERR_ORDER = -10,          /* step called out of order */
    CR_ERR_UNSUPPORTED = -11,    /* rule combination not implemented on the device */
    CR_ERR_NO_DEVICE = -12,      /* no HIP device visible */
    CR_ERR_FORMAT = -13,         /* a cache blob whose layout does not fit the loaded graph */
    CR_ERR_CACHE_RULES = -14     /* a cache blob solved under other self-edge rules (CacheManager.h:89-95) */
};

/* Solver variants: which reference solver translation unit is replaced. */
enum cr_variant {
    CR_VARIANT_COLUMNS = 1, /* solvers/k_cops_visibility.cpp (2p columns; p = 1 = plain game) */
    CR_VARIANT_MACRO = 2    /* solvers/k_cops_visibility_v2.cpp (one column, p-step macro turns) */
};

typedef struct cr_solver cr_solver;

typedef struct {
    int32_t n;             /* AuxGraph::N */
    int32_t k;             /* AuxGraph::k */
    int32_t p;
    int32_t columns;       /* AuxGraph::columns (2p, or 1 for the macro variant) */
    int32_t variant;
    int32_t max_degree;
    uint64_t config_count; /* AuxGraph::configCount */
    uint64_t num_states;   /* AuxGraph::numStates */
    uint64_t device_bytes; /* device memory held by this solver */
} cr_info;

typedef struct {
    int32_t win;              /* 1 = WIN, 0 = LOSS */
    int32_t plys;             /* overallMinWorstCase (valid when win) */
    uint64_t best_cid;        /* lowest cId achieving it (valid when win) */
    uint8_t positions[CR_MAX_K]; /* configs[best_cid*k .. +k) */
} cr_verdict;

typedef struct {
    uint64_t base_captures; /* "Initialized n base captures." */
    uint64_t won_states;    /* marked states of every column */
    int32_t n_passes;       /* reference passes (last one reports 0) */
    int32_t n_levels;       /* device levels / passes executed */
    double init_ms;         /* device time of initializeCaptures */
    double loop_ms;         /* device time of mainLoop */
    uint64_t pred_entries;  /* predecessor-list entries built by cr_build_aux_graph (0: none) */
    double build_ms;        /* device time to build them */
} cr_stats;

/* ---- lifetime -------------------------------------------------------- */
int cr_solver_create(int device, cr_solver** out);
void cr_solver_destroy(cr_solver* s);
/* Message of the last failure on this thread ("" if none). */
const char* cr_last_error(void);

/* ---- graph input ----------------------------------------------------- */
/* Graph::Graph(const char*) parsing, lib/src/Graph.cpp:15-79: node count from
 * the first line's width, '0'/'1' cells, '-' ends the matrix. Returns N (0 for
 * an empty or malformed text) and fills adj[N*N] when N <= cap. */
int32_t cr_parse_matrix(const uint8_t* text, uint64_t len, uint8_t* adj, int32_t cap);

/* loadGraphFile(filename, k, p), solvers/k_cops_visibility.cpp:38-57
 * (Graph + AdjacencyList::constructFrom, AdjacencyList.cpp:33-72). */
int cr_load_graph_file(cr_solver* s, const char* filename, int32_t k, int32_t p);
/* Same step from an in-memory N*N 0/1 matrix (row u = moves out of u). */
int cr_load_graph_matrix(cr_solver* s, int32_t n, const uint8_t* adj, int32_t k, int32_t p);
/* Selects the solver TU being replaced (default CR_VARIANT_COLUMNS). */
int cr_set_variant(cr_solver* s, int32_t variant);
/* Movement rules, AuxGraph::setSelfEdges (AuxGraph.h:81-91): cops may stay on
 * their vertex (an extra option per cop, AuxGraph.h:203-205) and/or the robber
 * may stay (k_cops_visibility.cpp:253-255). The reference solvers fix both to
 * FALSE (k_cops_visibility.cpp:63-64, the default here); the legacy round
 * solver lets both stay (legacy/src/k_cops_rounds.cpp:112-118, robber :224-233).
 * Takes effect at the next cr_build_aux_graph; the column variant only. */
int cr_set_rules(cr_solver* s, int32_t cop_stay, int32_t robber_stay);
/* 1-based edge-list input, legacy/util/edge_parser.py:4-63: one edge "u v"
 * (or "u, v") per line, '#' comments, undirected, N = the largest id. Returns N
 * (0 when there is no edge) and fills adj[N*N] when N <= cap. */
int32_t cr_parse_edge_list(const uint8_t* text, uint64_t len, uint8_t* adj, int32_t cap);
/* loadGraphFile from an edge list (the matrix edge_parser.py would write). */
int cr_load_edge_file(cr_solver* s, const char* filename, int32_t k, int32_t p);

/* ---- the six solver steps -------------------------------------------- */
/* buildAuxGraph(), k_cops_visibility.cpp:59-83 -> AuxGraph::constructFrom
 * (AuxGraph.h:59-79): configuration count, device state space. Configurations
 * and team moves are not materialised; the device ranks and unranks sorted
 * cop tuples in closed form (AuxGraph.h:117-177 order). */
int cr_build_aux_graph(cr_solver* s);
/* initializeCaptures(), k_cops_visibility.cpp:85-127 (v2 :200-224). */
int cr_initialize_captures(cr_solver* s, uint64_t* base_captures);
/* mainLoop(), k_cops_visibility.cpp:129-299 (v2 :343-565). pass_counts
 * receives the per-pass "Found n new winning states." values (at most
 * max_passes of them); n_passes the number of passes the reference prints. */
int cr_main_loop(cr_solver* s, uint64_t* pass_counts, int32_t max_passes, int32_t* n_passes);
/* findFinalResult(), k_cops_visibility.cpp:301-354 (v2 :567-619). */
int cr_find_final_result(cr_solver* s, cr_verdict* out);
/* outputData() -> CacheManager::saveAuxGraph, CacheManager.h:28-75: writes the
 * 96-byte header, configs, transition heads, transitions (nextId*N) and
 * states to path. */
int cr_save_aux_graph(cr_solver* s, const char* path);
/* CacheManager::loadAuxGraph, CacheManager.h:77-141, in place of buildAuxGraph
 * (k_cops_visibility.cpp:66-71): reads the blob at path for the graph, k, p,
 * variant and rules loaded in s. The blob's self-edge flags must equal the
 * solver's rules (CR_ERR_CACHE_RULES, with the reference's "Cache rules
 * mismatch" text in cr_last_error). Every other header field, the section
 * offsets, the file size, the configs section and the ends of the transition
 * heads must be what cr_save_aux_graph writes for this graph (CR_ERR_FORMAT;
 * the reference takes them from the header, here the device layout is fixed
 * by the graph). The states section becomes the device table: cr_copy_states,
 * cr_find_final_result, cr_write_dp_table, cr_extract_game and
 * cr_save_aux_graph read it as after a solve; the level and retro-edge
 * accessors and cr_main_loop need cr_initialize_captures first, which resets
 * the table as the reference's does on a cache hit (k_cops_visibility.cpp:90-104). */
int cr_load_aux_graph(cr_solver* s, const char* path);
/* The 96-byte header of a cache blob (CacheManager.h:49-58), after checking
 * the section offsets against each other and the file size (no device). */
int cr_read_aux_header(const char* path, uint64_t* header12);

/* ---- accessors ------------------------------------------------------- */
int cr_get_info(const cr_solver* s, cr_info* out);
int cr_get_stats(const cr_solver* s, cr_stats* out);
/* AuxGraph::states bytes [offset, offset+count) (AuxGraph.h:32-33, 96-98). */
int cr_copy_states(cr_solver* s, uint8_t* dst, uint64_t offset, uint64_t count);
/* AuxGraph::configs rows [first, first+count), k bytes each (AuxGraph.h:28). */
int cr_copy_configs(cr_solver* s, uint8_t* dst, uint64_t first, uint64_t count);
/* Depth-level histogram of a solved column-0/1 table: out[2*d + t] = states of
 * column t won at depth d (d < 127). Only for the p = 1 plain game. */
int cr_level_counts(cr_solver* s, uint64_t* out);
/* Retro-edge count of the solved p = 1 table (the benchmark's unit of work):
 * e_cop = sum over won robber-turn states (depth <= 125) of |pred(C')|,
 * e_rob = sum over won cop-turn states (depth <= 125) of in-degree(v). */
int cr_count_retro_edges(cr_solver* s, uint64_t* e_cop, uint64_t* e_rob, uint64_t* won);
/* Retro-edges of the solved p = 1 table by depth (arrays of 127): e_cop[d]
 * from won robber-turn states of depth d, e_rob[d] from won cop-turn states
 * of depth d; pulled[d] = 1 when the cop turn of level d gathered instead of
 * pushing (depth-0 relaxations are resolved in closed form by
 * cr_initialize_captures). */
int cr_work_profile(cr_solver* s, uint64_t* e_cop, uint64_t* e_rob, uint8_t* pulled);
/* Per depth d (127 entries of two): the configurations with cop-turn
 * ([2d]) and robber-turn ([2d + 1]) states of depth d -- the lengths of the
 * level lists. No counterpart in the reference (it has no level lists). */
int cr_level_rows(cr_solver* s, uint64_t* rows);
/* Limited-visibility variant (CR_VARIANT_MACRO): the trajectory DAG of the last
 * mainLoop and the work it did. No counterpart in the reference, whose
 * trajectory search (k_cops_visibility_v2.cpp:243-341) keeps no such graph. */
typedef struct {
    int32_t steps;        /* p: cop steps per macro turn */
    uint64_t dag_nodes;   /* (configuration, cloud) nodes over all steps; step 1 = the states */
    uint64_t dag_edges;   /* edges between consecutive steps */
    uint64_t build_moves; /* team moves enumerated while building the DAG */
    uint64_t leaf_moves;  /* (node, team move) pairs evaluated at step p, summed over passes */
    uint64_t up_edges;    /* DAG edges read on the way up, summed over passes */
    double build_ms;      /* device time of the DAG build */
    double passes_ms;     /* device time of the passes */
} cr_macro_stats;
int cr_get_macro_stats(cr_solver* s, cr_macro_stats* out, uint64_t* nodes_per_step, uint64_t* edges_per_step,
                       int32_t cap);
/* DP-table export (the assets/dp_tables/<graph>_<k>cops_dp_table.json layout of
 * the legacy round solver, legacy/src/k_cops_rounds.cpp:211-281 and
 * legacy/python/k_cops.py:86-131): a JSON list, Python json.dumps spacing, one
 * object per (cId, r) in cId order: {"cops": [...], "robber": r, "cop_wins":
 * true|false, "steps_to_win": s}, s = (depth + 1) / 2 for a won cop-turn state
 * (a round is two plies), -1 otherwise. Solved column variant with p = 1. */
int cr_write_dp_table(cr_solver* s, const char* path);
/* Perfect game (assets/cached_solutions/<graph>_<k>cops_perfect_game.json).
 * From the universal configuration with the least worst-case rounds (lowest
 * cId on ties) and the robber start with the most rounds (lowest r on ties),
 * the cops pick the move minimising the robber's best reply and the robber
 * the reply maximising the rounds left, first best in the order of the
 * reference producer chosen by `order`:
 *   CR_GAME_ORDER_PY  legacy/python/k_cops.py:174-241: cop moves in
 *                     itertools.product order over [neighbours ascending,
 *                     stay], the robber's options likewise;
 *   CR_GAME_ORDER_CPP legacy/src/k_cops_rounds.cpp:307-397: cop moves in
 *                     ascending order of the next configuration, the robber's
 *                     options its neighbours only.
 * Walked on the device. turn: 0 = "Cop's Turn", 1 = "Robber's Turn",
 * 2 = "Game Over - Captured!". length = 0 when no configuration wins. */
enum cr_game_order { CR_GAME_ORDER_PY = 0, CR_GAME_ORDER_CPP = 1 };
typedef struct {
    uint8_t cops[CR_MAX_K];
    int32_t robber;
    int32_t turn;
} cr_game_step;
int cr_extract_game(cr_solver* s, int32_t order, cr_game_step* out, int32_t cap, int32_t* length);
/* The same game as JSON in the reference's layout (json.dump, indent 4). */
int cr_write_perfect_game(cr_solver* s, int32_t order, const char* path);
/* Cop number: the least k in 1..k_max whose solve (same graph, variant, p and
 * rules) is a WIN (the harness sweeps k for this, util/test_k_cops_visibility.py:
 * 121-145). cop_number = 0 when none is; verdict is that solve's. The solver is
 * left solved at the last k tried. */
int cr_cop_number(cr_solver* s, int32_t k_max, int32_t* cop_number, cr_verdict* verdict);
/* Total team moves sum over C of |T(C)|: the "Total edge pointers" the
 * reference prints after AuxGraph::createTransitions (AuxGraph.h:274). */
int cr_count_transitions(cr_solver* s, uint64_t* total);
/* Tuning knobs (no counterpart in the reference; they change speed, never
 * results). */
enum cr_tuning {
    CR_TUNE_PUSH_FILTER = 1,    /* 1: cop push skips targets already won (default 1) */
    CR_TUNE_PUSH_BLOCKS = 2,    /* workgroups of the cop-push kernel (default 2048) */
    CR_TUNE_ROW_BLOCKS = 3,     /* workgroups of the row kernels (default 32768, at most one per group) */
    CR_TUNE_LEVELS_PER_SYNC = 4, /* depth levels enqueued per host synchronisation (default 8) */
    CR_TUNE_PRED_CSR = 5,        /* 1: cr_build_aux_graph precomputes predecessor lists (default 0) */
    CR_TUNE_PULL_SCALE = 6,      /* a level gathers (factorised pull) instead of pushing when its robber
                                    frontier exceeds value/1000 times the estimated break-even row count;
                                    0: always pull, >= 10^9: never; default 150 while the cop-turn
                                    bitsets fit in 256 MiB, else 190; a negative value restores it */
    CR_TUNE_PULL_TILES = 7,      /* 1 (default): the pull's stages load their input rows through LDS
                                    tiles; 0: one row per lane with direct gathers */
    CR_TUNE_PULL_TILE_ROWS = 8,  /* rows of each input block per LDS tile, 1..16 (default 4) */
    CR_TUNE_PULL_TILE_BLOCKS = 9, /* workgroups of a tiled pull stage (default 8192) */
    CR_TUNE_DEPTH_LOG = 10,      /* capacity of the depth log in units of the configuration count
                                    (default 40; 0: no log, every depth writes its bytes directly):
                                    each depth's new states are logged and written to the table in row
                                    chunks. Takes effect at the next cr_build_aux_graph */
    CR_TUNE_DEPTH_LOG_MIN = 11,  /* configurations below which no depth log is kept (default 2^24) */
    CR_TUNE_PULL_TILE_CHUNK = 12, /* consecutive tiles a pull workgroup takes at a time, at most (default 16;
                                    fewer where a stage has under 2048 * 16 tiles) */
    CR_TUNE_FUSE_FINALIZE = 13,  /* 1 (default): where the depth log is kept, a pulled level's last
                                    stage settles the new cop-turn states itself (no finalize pass) */
    CR_TUNE_PULL_TILE1 = 14,     /* consecutive cop multisets per tile of a tiled pull's first stage,
                                    1..16, fewer where the tile does not fit (default 9); 0 = the first
                                    stage runs k_pull_tile */
    CR_TUNE_PULL_GATE = 15,      /* levels enqueued between host synchronisations may pull only if the
                                    first one's robber frontier is above 1/gate of the pull threshold
                                    (else they push and launch no pull kernels); 0 = always (default 4) */
    CR_TUNE_ROBBER_OPEN = 16,    /* the robber step finds a row's new robber-turn states from its open
                                    cop-turn states where open <= value * new cop-turn states, else from
                                    the new ones; 0..1000 (default 2) */
    CR_TUNE_LOG_BLOCKS = 17,     /* workgroups of k_log_lists, which appends the level lists to the
                                    depth log (default 4096) */
    CR_TUNE_SCAN_TILE = 18,      /* list groups per tile of the level scan, 4..8192 (default 8192): with
                                    more groups than one tile, tiles are scanned in parallel first */
    CR_TUNE_PUSH_DIRTY = 19,     /* 1 (default): the push marks the groups of 256 configurations it
                                    writes, and the finalize of a push level reads only those */
    CR_TUNE_PULL_PAIR = 20,      /* a tiled pull runs its last two stages as one kernel that keeps the
                                    next-to-last table in LDS: 0 never, 1 (default) for k = 3 (measured
                                    faster there; no gain at k = 4), 2 for every k >= 3 */
    CR_TUNE_PAIR_ROWS = 21,      /* cop multisets per tile of that kernel, 1..8 (default 4) */
    CR_TUNE_PAIR_BLOCKS = 22,    /* workgroups of that kernel, rounded to a multiple of 8 (default 1024) */
    CR_TUNE_END_PASS = 23        /* 1 (default): where one scan covers a level's group counts (up to 8192
                                    groups), the level ends in one launch (single-pass scan with
                                    decoupled look-back, list writes); 0: a scan launch, then the lists */
};
int cr_set_tuning(cr_solver* s, int32_t key, int64_t value);
/* Per-kernel device time of the last cr_main_loop when timing is enabled:
 * name_out receives a '\0'-separated list, ms/launches one entry per name. */
int cr_set_kernel_timing(cr_solver* s, int32_t enabled);
/* Restricts timing to the launches of one kernel (name as cr_kernel_times
 * reports it); NULL or "" times every launch. Timing events sit between
 * launches, so timing fewer kernels perturbs the loop less. */
int cr_set_kernel_timing_filter(cr_solver* s, const char* name);
int cr_kernel_times(cr_solver* s, char* names_out, int32_t names_cap, double* ms,
                    int64_t* launches, int32_t cap, int32_t* count);
/* Build-time cache key, CacheManager::buildCacheKey (CacheManager.h:16-26),
 * without creating the directory. Returns the length written (excluding
 * '\0') or a negative status. */
int cr_cache_key(const char* graph_path, int32_t k, int32_t p, int32_t variant, char* out,
                 int32_t cap);

/* ---- many small graphs: one launch, and a continuous-batching server ---
 * The plain game (full information, p = 1) of many independent graphs. The
 * reference solves one graph per process (app/main.cpp:92-111, swept by
 * util/test_k_cops_visibility.py:121-145); these entry points have no
 * counterpart there and serve callers that solve graphs in bulk or as a
 * stream of requests. Every result equals the six-step solve of that graph
 * (tables, passes, verdict). */
#define CR_BATCH_MAX_N 64 /* vertices per graph in a batch */
#define CR_BATCH_MAX_K 4  /* cops per graph in a batch */

typedef struct {
    uint64_t ticket;        /* cr_server_submit's ticket (index for cr_batch_solve) */
    int32_t status;         /* CR_OK, or the error that rejected this graph */
    int32_t win;            /* findFinalResult: 1 = WIN */
    int32_t plys;           /* capture time when win, else -1 */
    int32_t n_passes;       /* passes the reference prints (the last reports 0) */
    uint64_t best_cid;      /* optimal start configuration when win */
    uint8_t positions[CR_MAX_K];
    uint64_t base_captures; /* "Initialized n base captures." */
    uint64_t won_states;    /* marked states of both columns */
    uint64_t table_hash;    /* sum over state bytes b_i of (b_i + 1) * mix64(i): equal tables, equal hash */
    uint64_t table_bytes;   /* 2 * C * n */
    double latency_ms;      /* cr_server: submit (or stated arrival) to completion; 0 otherwise */
} cr_graph_result;

typedef struct cr_batch cr_batch;
int cr_batch_create(int device, cr_batch** out);
void cr_batch_destroy(cr_batch* b);
/* Solves count graphs in one persistent launch. Graph i has n[i] <= 64
 * vertices, k[i] <= 4 cops and its n[i]*n[i] 0/1 matrix (Graph.cpp layout,
 * row = moves out of a vertex) at adj + adj_offset[i]. results[i] receives its
 * result. When tables is not null, graph i's state table (the reference's
 * layout, results[i].table_bytes bytes) is copied to tables + table_offset[i].
 * Graphs that fail validation get their status in results[i].status and the
 * call returns CR_ERR_ARG after solving the others. */
int cr_batch_solve(cr_batch* b, int32_t count, const int32_t* n, const int32_t* k, const uint8_t* adj,
                   const uint64_t* adj_offset, cr_graph_result* results, uint8_t* tables,
                   const uint64_t* table_offset);
/* Launch shape of cr_batch_solve and of the server's launches (speed only). */
enum cr_batch_tuning {
    CR_BATCH_TUNE_THREADS = 1,      /* threads per workgroup (one graph at a time): 256, 512, 1024; 0 = 1024
                                       when the batch's largest graph has >= 4096 configurations, else 256 */
    CR_BATCH_TUNE_BLOCKS_PER_CU = 2, /* persistent workgroups per CU; 0 = 2048 threads per CU */
    CR_BATCH_TUNE_PUSH_PCT = 3       /* a level's cop turn pushes from its robber frontier instead of
                                        pulling when the push's bound (frontier states x largest
                                        in-degree^k) is at most this % of the pull's row loads (default
                                        100; 0 = always pull) */
};
int cr_batch_set_tuning(cr_batch* b, int32_t key, int64_t value);
/* Per-pass "Found n" counts of graph `index` of the last cr_batch_solve. */
int cr_batch_pass_counts(cr_batch* b, int32_t index, uint64_t* out, int32_t cap, int32_t* n_passes);
/* Device time of the last cr_batch_solve (launch to last result) and its grid. */
int cr_batch_last_timing(cr_batch* b, double* device_ms, int32_t* workgroups);

/* No counterpart in the reference: a variant of cr_batch_solve for callers
 * that cannot allocate memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc. It makes no device or host
 * allocation: the graph descriptors, robber masks, work counter, per-graph
 * outputs and the workgroups' work planes live in one file-scope device array
 * of CR_NOALLOC_SCRATCH_BYTES, which each device it runs on holds until the
 * process ends. Inputs as in cr_batch_solve (host arrays, at most
 * CR_NOALLOC_MAX_GRAPHS graphs). Graph i's state table is written to the
 * caller's device memory d_tables + table_offsets[i]; the offsets are
 * computed here (each table 256-byte aligned, table_offsets[count] = the bytes
 * needed; table_offsets has count + 1 entries). With d_tables NULL the call
 * only computes the offsets. CR_ERR_TOO_LARGE when count, tables_cap or one
 * graph's work planes exceed these limits. Runs on the device's null stream
 * and returns when the results are in `results`. Calls that overlap on one
 * device (and any concurrent calls: the host staging is shared) race; the
 * array keeps the last call's contents. */
#define CR_NOALLOC_SCRATCH_BYTES (256ull << 20)
#define CR_NOALLOC_MAX_GRAPHS 4096
int cr_batch_solve_noalloc(int device, int32_t count, const int32_t* n, const int32_t* k, const uint8_t* adj,
                           const uint64_t* adj_offset, cr_graph_result* results, uint8_t* d_tables,
                           uint64_t tables_cap, uint64_t* table_offsets);

/* Continuous batching: a dispatcher thread admits pending requests into a new
 * launch whenever one of `streams` launch slots is free and their footprint
 * (state table + per-graph work planes, bytes) fits in what `hbm_budget`
 * leaves after the launches in flight; launches on different slots overlap
 * on the device. Results come back in completion order. */
typedef struct cr_server cr_server;
int cr_server_create(int device, uint64_t hbm_budget, int32_t max_batch, int32_t streams, cr_server** out);
/* Copies the request; arrival_ns (cr_clock_ns time base, 0 = now) is when it
 * arrived, from which its latency is measured. */
int cr_server_submit(cr_server* s, int32_t n, int32_t k, const uint8_t* adj, uint64_t arrival_ns,
                     uint64_t* ticket);
/* Submits count requests in one call (graph i laid out as in cr_batch_solve;
 * arrival_ns may be null = now). tickets[i] gets request i's ticket, or 0
 * when it was rejected, with the reason in status[i] (may be null). Returns
 * CR_OK, or CR_ERR_ARG when some request was rejected. */
int cr_server_submit_many(cr_server* s, int32_t count, const int32_t* n, const int32_t* k, const uint8_t* adj,
                          const uint64_t* adj_offset, const uint64_t* arrival_ns, uint64_t* tickets,
                          int32_t* status);
/* cr_batch_set_tuning for every launch slot (before the first submit). */
int cr_server_set_tuning(cr_server* s, int32_t key, int64_t value);
/* Moves up to cap finished results to out; waits up to wait_us for the first. */
int cr_server_poll(cr_server* s, cr_graph_result* out, int32_t cap, int32_t wait_us, int32_t* count);
/* Requests submitted, finished, launches made, peak bytes admitted at once. */
int cr_server_stats(cr_server* s, uint64_t* submitted, uint64_t* finished, uint64_t* launches,
                    uint64_t* peak_bytes);
/* Finishes the work in flight and stops the dispatcher. */
/* Resident tables. Every reference solve emits its full state table
 * (CacheManager.h:28-75); a server with retention on keeps the table of each
 * request it finishes in HBM, in the reference's layout, until the caller
 * releases it. Retained tables count against hbm_budget: when new work needs
 * the room, the tables of the oldest finished launches are evicted (their
 * tickets then have no table). */
int cr_server_retain_tables(cr_server* s, int32_t on);
/* Copies the table of a finished request to dst (cap bytes; dst NULL: only
 * *bytes is set). CR_ERR_ARG when the ticket has no resident table. */
int cr_server_copy_table(cr_server* s, uint64_t ticket, uint8_t* dst, uint64_t cap, uint64_t* bytes);
/* Frees a request's table (its launch's memory once all its tables are released). */
int cr_server_release_table(cr_server* s, uint64_t ticket);
int cr_server_table_stats(cr_server* s, uint64_t* resident_tables, uint64_t* resident_bytes,
                          uint64_t* evicted_launches);
void cr_server_destroy(cr_server* s);
/* Steady clock in nanoseconds (the arrival_ns time base). */
uint64_t cr_clock_ns(void);

#ifdef __cplusplus
}
#endif
#endif /* CRSOLVE_H */
