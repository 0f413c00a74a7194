/* Exercises the oracle (oracle/cr_oracle.c) under sanitizers: the pass-sweep
 * solver, both multithreaded level-synchronous solvers, the Bellman check,
 * the table summary, the limited-visibility solver and the move lists, on
 * one matrix file. Exits non-zero when the solvers disagree. Usage:
 * oracle_driver MATRIX K THREADS */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cr_oracle.h"

static uint8_t* read_file(const char* path, long* len) {
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    *len = ftell(f);
    fseek(f, 0, SEEK_SET);
    uint8_t* buf = malloc((size_t)*len + 1);
    if (fread(buf, 1, (size_t)*len, f) != (size_t)*len) { fclose(f); free(buf); return NULL; }
    fclose(f);
    return buf;
}

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    long len = 0;
    uint8_t* text = read_file(argv[1], &len);
    const int k = atoi(argv[2]), threads = atoi(argv[3]);
    if (!text) return 2;
    const int n = or_parse_matrix(text, (uint64_t)len, NULL, 0);
    uint8_t* adj = malloc((size_t)n * n);
    or_parse_matrix(text, (uint64_t)len, adj, n);
    const uint64_t C = or_config_count(n, k), S = C * (uint64_t)n * 2;
    uint8_t* a = malloc(S);
    uint8_t* b = malloc(S);
    uint8_t* c = malloc(S);
    uint64_t base = 0, passes[4096], levels[2 * OR_MAX_ROUND], ec = 0, er = 0, ec2 = 0, er2 = 0;
    int32_t np = 0, done = 0;
    double frac = 0, el = 0;
    int rc = 0;
    rc |= or_v1_solve(n, adj, k, 1, 0, 0, a, &base, passes, 4096, &np);
    rc |= or_v1_levelsync(n, adj, k, threads, b, levels, &ec, &er);
    rc |= or_v1_levelsync_capped(n, adj, k, threads, 1e9, c, &ec2, &er2, &done, &frac, &el);
    if (rc) { fprintf(stderr, "solver failed\n"); return 1; }
    if (memcmp(a, b, S) || memcmp(a, c, S) || ec != ec2 || er != er2) { fprintf(stderr, "tables differ\n"); return 1; }
    uint64_t* ids = malloc(sizeof(uint64_t) * S);
    for (uint64_t i = 0; i < S; ++i) ids[i] = i;
    uint64_t bad = 0, first = 0;
    if (or_check_states(n, adj, k, a, ids, S, threads, &bad, &first) || bad) { fprintf(stderr, "check %llu\n", (unsigned long long)bad); return 1; }
    uint64_t lv[2 * OR_MAX_ROUND];
    int64_t best = 0, best2 = 0;
    int32_t plys = 0, plys2 = 0;
    or_table_summary(n, k, a, threads, lv, &best, &plys);
    or_verdict(n, k, 2, a, &best2, &plys2);
    if (memcmp(lv, levels, sizeof lv) || best != best2 || plys != plys2) { fprintf(stderr, "summary differs\n"); return 1; }
    /* the limited-visibility solver at p = 2 and the move lists */
    uint8_t* v2 = malloc(C * (uint64_t)n);
    if (or_v2_solve(n, adj, k, 2, v2, &base, passes, 4096, &np)) { fprintf(stderr, "v2 failed\n"); return 1; }
    uint64_t* heads = malloc(sizeof(uint64_t) * (C + 1));
    int64_t total = or_transitions(n, adj, k, 0, heads, NULL, 0);
    uint64_t* trans = malloc(sizeof(uint64_t) * (size_t)(heads[C] ? heads[C] : 1));
    total = or_transitions(n, adj, k, 0, heads, trans, heads[C]);
    if (total < 0) { fprintf(stderr, "transitions failed\n"); return 1; }
    uint8_t t[OR_MAX_K];
    for (uint64_t r = 0; r < C; r += 1 + C / 97) {
        or_unrank(r, n, k, t);
        if (or_rank(t, n, k) != r) { fprintf(stderr, "rank\n"); return 1; }
    }
    free(a); free(b); free(c); free(ids); free(v2); free(heads); free(trans); free(adj); free(text);
    printf("oracle driver ok: n=%d k=%d states=%llu\n", n, k, (unsigned long long)S);
    return 0;
}
