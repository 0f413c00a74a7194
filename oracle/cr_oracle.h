/*
 * cr_oracle.h -- CPU restatement of the reference k-cop retrograde solver.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests/, the smoke
 * check and the CPU-baseline leg of bench.py call it to compute the expected
 * tables. The product (libcrsolve.so and its plug-in) never links, loads or
 * calls it.
 *
 * Every function names the reference file:line range it restates (paths are
 * relative to the reference repository root). State tables use the
 * reference's byte layout: index (cId*N + r)*columns + col, byte value
 * marked | markedRound << 1 (GCC bit-field order for `DataItem`,
 * solvers/k_cops_visibility.cpp:20-23).
 *
 * Parity status of each function (how it is pinned) is listed in the header
 * comment of cr_oracle.c.
 */
#ifndef CR_ORACLE_H
#define CR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_K 16
#define OR_MAX_ROUND 127

/* Graph.cpp:15-79 (+ fileio.cpp:6-30 for the whole-file read). Returns the
 * node count N (0 = empty/malformed). Fills adj[N*N] with 0/1 when N <= cap. */
int32_t or_parse_matrix(const uint8_t* buf, uint64_t len, uint8_t* adj, int32_t cap);

/* AuxGraph.h:124-138: C(N+k-1, k), 0 when k > N+k-1. */
uint64_t or_config_count(int32_t n, int32_t k);

/* AuxGraph.h:155-176: lexicographic non-decreasing k-tuples. */
int32_t or_enum_configs(int32_t n, int32_t k, uint8_t* out, uint64_t count);

/* Position of a sorted tuple in the enumeration above (closed form). */
uint64_t or_rank(const uint8_t* sorted_cops, int32_t n, int32_t k);

/* AuxGraph.h:179-275: per-config sorted, de-duplicated team moves stored as
 * nextId*N (the reference's transition value). heads has C+1 entries.
 * Returns the total entry count, or -1 when cap is too small (heads filled). */
int64_t or_transitions(int32_t n, const uint8_t* adj, int32_t k, int32_t cop_stay,
                       uint64_t* heads, uint64_t* trans, uint64_t cap);

/* initializeCaptures + mainLoop of solvers/k_cops_visibility.cpp:85-299,
 * in the reference's own pass-sweep order. states has C*N*2p bytes. */
int32_t or_v1_solve(int32_t n, const uint8_t* adj, int32_t k, int32_t p,
                    int32_t cop_stay, int32_t robber_stay, uint8_t* states,
                    uint64_t* base_caps, uint64_t* pass_counts, int32_t max_passes,
                    int32_t* n_passes);

/* findFinalResult (k_cops_visibility.cpp:301-354; v2 :567-619): best_cid = -1
 * for LOSS. */
int32_t or_verdict(int32_t n, int32_t k, int32_t columns, const uint8_t* states,
                   int64_t* best_cid, int32_t* plys);

/* initializeCaptures + mainLoop of solvers/k_cops_visibility_v2.cpp:200-565
 * (macro-turn value with snapshot passes), evaluated by memoising the
 * (config, cloud, step) sub-games of one pass. states has C*N bytes. */
int32_t or_v2_solve(int32_t n, const uint8_t* adj, int32_t k, int32_t p,
                    uint8_t* states, uint64_t* base_caps, uint64_t* pass_counts,
                    int32_t max_passes, int32_t* n_passes);

/* Level-synchronous counter retrograde for p = 1 on nthreads host threads,
 * structured like legacy/src/k_cops_4.cpp:388-481 but under the v1 rules
 * (no stay moves, depth kept, 127 cap). Produces the same table as
 * or_v1_solve(p=1). level_counts[2*d+t] = states of turn t won at depth d
 * (d < 127). e_cop/e_rob = predecessor relaxations performed. */
int32_t or_v1_levelsync(int32_t n, const uint8_t* adj, int32_t k, int32_t nthreads,
                        uint8_t* states, uint64_t* level_counts, uint64_t* e_cop,
                        uint64_t* e_rob);

/* Retro-edge count of a solved p = 1 table: E_cop = sum over won robber-turn
 * states of depth <= 125 of |pred(C')|, E_rob = sum over won cop-turn states of
 * depth <= 125 of in-degree(v); won = number of won states. */
int32_t or_retro_edges(int32_t n, const uint8_t* adj, int32_t k, const uint8_t* states,
                       uint64_t* e_cop, uint64_t* e_rob, uint64_t* won);

/* Per-pass "Found n new winning states" counts of the p = 1 sweep, derived
 * from a solved table's depths. Returns the number of passes printed. */
int32_t or_pass_counts_from_depths(int32_t n, int32_t k, const uint8_t* states,
                                   uint64_t* pass_counts, int32_t max_passes);

/* Inverse of or_rank: the tuple at lexicographic position rank. */
int32_t or_unrank(uint64_t rank, int32_t n, int32_t k, uint8_t* out);

/* Local (Bellman) check of a solved p = 1 table, plain rules, at the given
 * state indices ((cId*N + r)*2 + turn): each state must hold the byte its
 * successors in the same table imply (captures at round 0; cop turn 1 + min
 * over T(C), robber turn 1 + max over N(r) when all marked; cap 127), per
 * solvers/k_cops_visibility.cpp:85-299. bad = number of states that differ,
 * first_bad = the index of one of them (UINT64_MAX if none). */
int32_t or_check_states(int32_t n, const uint8_t* adj, int32_t k, const uint8_t* states,
                        const uint64_t* ids, uint64_t count, int32_t nthreads, uint64_t* bad,
                        uint64_t* first_bad);

/* One multithreaded pass over a p = 1 table: level_counts[2*d+t] = marked
 * states of turn t and round d, and the findFinalResult verdict
 * (k_cops_visibility.cpp:301-354; best_cid = -1 for LOSS). */
int32_t or_table_summary(int32_t n, int32_t k, const uint8_t* states, int32_t nthreads,
                         uint64_t* level_counts, int64_t* best_cid, int32_t* plys);

/* Level-synchronous p = 1 solver (undirected graphs, plain rules) with team
 * moves generated on the fly and a wall-clock cap: the CPU baseline where a
 * move table does not fit. Same table as or_v1_levelsync when it finishes.
 * levels_done = depth levels completed; level_fraction = share of the
 * configurations the interrupted level had scanned; e_cop/e_rob = the
 * relaxations performed until then; elapsed_s includes the captures. */
int32_t or_v1_levelsync_capped(int32_t n, const uint8_t* adj, int32_t k, int32_t nthreads,
                               double time_cap_s, uint8_t* states, uint64_t* e_cop,
                               uint64_t* e_rob, int32_t* levels_done, double* level_fraction,
                               double* elapsed_s);

#ifdef __cplusplus
}
#endif
#endif
