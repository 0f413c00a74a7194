/*
 * cr_oracle.c -- CPU restatement of the reference k-cop retrograde solver.
 *
 * TEST INFRASTRUCTURE ONLY (see cr_oracle.h). Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py load this library, always as the checker.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   or_parse_matrix     golden adjacency of all 28 reference matrices
 *                       (written by the compiled reference) + hand cases.
 *   or_config_count     closed form C(N+k-1,k) vs brute force; golden headers.
 *   or_enum_configs     golden configs sections of reference cache blobs.
 *   or_rank             = enumeration index for every tuple, N<=11, k<=4.
 *   or_transitions      golden heads/transitions sections of reference blobs,
 *                       and brute-force move sets on small graphs.
 *   or_v1_solve         golden states sections (byte-exact) of reference
 *                       blobs for p = 1..3; 180 rows of the reference's
 *                       recorded results CSV (verdict, plys, passes, base
 *                       captures, start positions); closed forms (path
 *                       centre, cycles need 2 cops, dominating sets).
 *   or_verdict          the same CSV rows.
 *   or_v2_solve         golden v2 blobs and the v2 results CSV rows.
 *   or_v1_levelsync     byte-equal to or_v1_solve(p=1) on every golden case.
 *   or_retro_edges      brute-force recount in Python on small cases and
 *                       equality with or_v1_levelsync's own counters.
 *   or_pass_counts_from_depths  equal to or_v1_solve's printed pass counts.
 *   or_unrank           inverse of or_rank over every tuple, N<=11, k<=4.
 *   or_check_states     0 differences on every golden p = 1 table (all states);
 *                       planted one-state mistakes (a depth off by one, a
 *                       marked state cleared, an unmarked state marked) found.
 *   or_table_summary    = or_verdict and or_v1_levelsync level counts on the
 *                       golden cases.
 *   or_v1_levelsync_capped  uncapped: byte-equal to or_v1_levelsync with equal
 *                       relaxation counts on the golden undirected cases.
 */
#define _GNU_SOURCE
#include "cr_oracle.h"

#include <pthread.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>

#define UNMARKED_BYTE ((uint8_t)(OR_MAX_ROUND << 1)) /* marked=0, round=127 */
#define IS_MARKED(b) ((b) & 1u)
#define DEPTH_OF(b) ((b) >> 1)
#define WON_BYTE(d) ((uint8_t)(1u | ((unsigned)(d) << 1)))

/* ------------------------------------------------------------------------ */
/* Graph input: Graph.cpp:26-70                                              */
/* ------------------------------------------------------------------------ */
int32_t or_parse_matrix(const uint8_t* buf, uint64_t len, uint8_t* adj, int32_t cap) {
    if (!buf || len == 0) return 0; /* fileio.cpp:12: empty file -> no buffer */
    uint64_t width = 0;             /* Graph.cpp:27-30: first line width */
    while (width < len && buf[width] != '\n' && buf[width] != '\r' && buf[width] != '-') width++;
    if (width == 0) return 0;
    int32_t n = (int32_t)width;
    if (!adj || n > cap) return n;
    memset(adj, 0, (size_t)n * (size_t)n);
    int32_t row = 0, col = 0;
    for (uint64_t i = 0; i < len; ++i) {
        uint8_t ch = buf[i];
        if (ch == '-') break; /* end marker */
        if (ch == '0' || ch == '1') {
            if (row < n && col < n) adj[(size_t)row * n + col] = (uint8_t)(ch == '1');
            col++;
        } else if (ch == '\n' && col > 0) { /* blank lines do not advance */
            row++;
            col = 0;
        }
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* Configurations: AuxGraph.h:117-177                                        */
/* ------------------------------------------------------------------------ */
uint64_t or_config_count(int32_t n, int32_t k) {
    if (k <= 0 || k > 256) return 0; /* MAX_COPS guard, AuxGraph.h:118-122 */
    int64_t top = (int64_t)n + k - 1;
    int64_t m = k;
    if (m > top) return 0;
    if (m == top) return 1;
    if (m > top / 2) m = top - m;
    uint64_t acc = 1;
    for (int64_t i = 1; i <= m; ++i) acc = acc * (uint64_t)(top - i + 1) / (uint64_t)i;
    return acc;
}

int32_t or_enum_configs(int32_t n, int32_t k, uint8_t* out, uint64_t count) {
    if (k <= 0 || k > OR_MAX_K || n <= 0) return -1;
    uint8_t t[OR_MAX_K];
    memset(t, 0, sizeof t);
    uint64_t idx = 0;
    for (;;) {
        if (idx >= count) return -1;
        memcpy(out + idx * (uint64_t)k, t, (size_t)k);
        idx++;
        int j = k - 1;
        while (j >= 0 && t[j] == n - 1) j--;
        if (j < 0) break;
        uint8_t v = (uint8_t)(t[j] + 1);
        for (int i = j; i < k; ++i) t[i] = v;
    }
    return idx == count ? 0 : -1;
}

static uint64_t binom(int64_t a, int64_t b) {
    if (b < 0 || a < b) return 0;
    if (b > a - b) b = a - b;
    uint64_t acc = 1;
    for (int64_t i = 1; i <= b; ++i) acc = acc * (uint64_t)(a - b + i) / (uint64_t)i;
    return acc;
}

/* Lexicographic rank of a non-decreasing tuple: shift to the strictly
 * increasing subset e_i = c_i + i of [0, M), M = N+k-1, and count the subsets
 * that come after it (complement of the colex sum). */
uint64_t or_rank(const uint8_t* c, int32_t n, int32_t k) {
    int64_t M = (int64_t)n + k - 1;
    uint64_t tail = 0;
    for (int i = 0; i < k; ++i) tail += binom(M - 1 - ((int64_t)c[i] + i), k - i);
    return binom(M, k) - 1 - tail;
}

/* ------------------------------------------------------------------------ */
/* Team moves: AuxGraph.h:179-275                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t n;
    int32_t deg[256];
    uint8_t nb[256][256]; /* ascending out-neighbours (row of the matrix) */
} nbrs_t;

static void build_nbrs(nbrs_t* g, int32_t n, const uint8_t* adj) {
    g->n = n;
    for (int u = 0; u < n; ++u) {
        int d = 0;
        for (int v = 0; v < n; ++v)
            if (adj[(size_t)u * n + v]) g->nb[u][d++] = (uint8_t)v;
        g->deg[u] = d;
    }
}

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

static void sort_small(uint8_t* t, int k) {
    for (int i = 1; i < k; ++i) {
        uint8_t v = t[i];
        int j = i - 1;
        while (j >= 0 && t[j] > v) { t[j + 1] = t[j]; j--; }
        t[j + 1] = v;
    }
}

/* All sorted, de-duplicated successor ranks of one configuration. Returns the
 * count written to out (capacity cap), or -1 if cap is exceeded. */
static int64_t moves_of(const nbrs_t* g, const uint8_t* cops, int32_t k, int cop_stay,
                        uint64_t* out, uint64_t cap) {
    uint8_t opt[OR_MAX_K][257];
    int cnt[OR_MAX_K];
    for (int i = 0; i < k; ++i) {
        int c = 0;
        if (cop_stay) opt[i][c++] = cops[i];
        for (int e = 0; e < g->deg[cops[i]]; ++e) opt[i][c++] = g->nb[cops[i]][e];
        cnt[i] = c;
        if (c == 0) return 0; /* a cop with no move: no team move exists */
    }
    int odo[OR_MAX_K];
    memset(odo, 0, sizeof odo);
    uint64_t m = 0;
    for (;;) {
        uint8_t t[OR_MAX_K];
        for (int i = 0; i < k; ++i) t[i] = opt[i][odo[i]];
        sort_small(t, k);
        if (m >= cap) return -1;
        out[m++] = or_rank(t, g->n, k);
        int j = k - 1;
        while (j >= 0 && ++odo[j] == cnt[j]) odo[j--] = 0;
        if (j < 0) break;
    }
    qsort(out, m, sizeof(uint64_t), cmp_u64);
    uint64_t w = 0;
    for (uint64_t i = 0; i < m; ++i)
        if (w == 0 || out[w - 1] != out[i]) out[w++] = out[i];
    return (int64_t)w;
}

static uint64_t product_bound(const nbrs_t* g, int32_t k, int cop_stay) {
    int maxd = 0;
    for (int u = 0; u < g->n; ++u) if (g->deg[u] > maxd) maxd = g->deg[u];
    uint64_t b = 1;
    for (int i = 0; i < k; ++i) b *= (uint64_t)(maxd + (cop_stay ? 1 : 0));
    return b;
}

/* CSR of successor config ids (not multiplied by N). */
typedef struct {
    uint64_t C;
    uint64_t* head; /* C+1 */
    uint32_t* id;
} csr_t;

static void csr_free(csr_t* c) { free(c->head); free(c->id); c->head = NULL; c->id = NULL; }

static int build_csr(const nbrs_t* g, int32_t k, int cop_stay, const uint8_t* configs,
                     uint64_t C, csr_t* out) {
    uint64_t pb = product_bound(g, k, cop_stay);
    uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * (pb ? pb : 1));
    out->C = C;
    out->head = (uint64_t*)calloc(C + 1, sizeof(uint64_t));
    uint64_t capacity = C * 8 + 16, used = 0;
    out->id = (uint32_t*)malloc(sizeof(uint32_t) * capacity);
    if (!tmp || !out->head || !out->id) { free(tmp); csr_free(out); return -1; }
    for (uint64_t c = 0; c < C; ++c) {
        int64_t m = moves_of(g, configs + c * (uint64_t)k, k, cop_stay, tmp, pb);
        if (m < 0) { free(tmp); csr_free(out); return -1; }
        if (used + (uint64_t)m > capacity) {
            while (used + (uint64_t)m > capacity) capacity *= 2;
            uint32_t* grown = (uint32_t*)realloc(out->id, sizeof(uint32_t) * capacity);
            if (!grown) { free(tmp); csr_free(out); return -1; }
            out->id = grown;
        }
        for (int64_t i = 0; i < m; ++i) out->id[used++] = (uint32_t)tmp[i];
        out->head[c + 1] = used;
    }
    free(tmp);
    return 0;
}

/* Reverse CSR: pred(c2) = { c : c2 in T(c) } (equal to T on symmetric graphs). */
static int transpose_csr(const csr_t* f, csr_t* r) {
    uint64_t C = f->C;
    r->C = C;
    r->head = (uint64_t*)calloc(C + 1, sizeof(uint64_t));
    r->id = (uint32_t*)malloc(sizeof(uint32_t) * (f->head[C] ? f->head[C] : 1));
    if (!r->head || !r->id) { csr_free(r); return -1; }
    for (uint64_t i = 0; i < f->head[C]; ++i) r->head[f->id[i] + 1]++;
    for (uint64_t c = 0; c < C; ++c) r->head[c + 1] += r->head[c];
    uint64_t* fill = (uint64_t*)malloc(sizeof(uint64_t) * (C ? C : 1));
    if (!fill) { csr_free(r); return -1; }
    memcpy(fill, r->head, sizeof(uint64_t) * C);
    for (uint64_t c = 0; c < C; ++c)
        for (uint64_t i = f->head[c]; i < f->head[c + 1]; ++i) r->id[fill[f->id[i]]++] = (uint32_t)c;
    free(fill);
    return 0;
}

static uint8_t* make_configs(int32_t n, int32_t k, uint64_t* C_out) {
    uint64_t C = or_config_count(n, k);
    *C_out = C;
    if (C == 0 || k > OR_MAX_K) return NULL;
    uint8_t* cfg = (uint8_t*)malloc(C * (uint64_t)k);
    if (!cfg) return NULL;
    if (or_enum_configs(n, k, cfg, C) != 0) { free(cfg); return NULL; }
    return cfg;
}

int64_t or_transitions(int32_t n, const uint8_t* adj, int32_t k, int32_t cop_stay,
                       uint64_t* heads, uint64_t* trans, uint64_t cap) {
    if (n <= 0 || n > 255) return -1;
    nbrs_t* g = (nbrs_t*)malloc(sizeof(nbrs_t));
    if (!g) return -1;
    build_nbrs(g, n, adj);
    uint64_t C;
    uint8_t* cfg = make_configs(n, k, &C);
    if (!cfg) { free(g); return -1; }
    csr_t f;
    if (build_csr(g, k, cop_stay, cfg, C, &f) != 0) { free(cfg); free(g); return -1; }
    for (uint64_t c = 0; c <= C; ++c) heads[c] = f.head[c];
    int64_t total = (int64_t)f.head[C];
    if ((uint64_t)total <= cap)
        for (int64_t i = 0; i < total; ++i) trans[i] = (uint64_t)f.id[i] * (uint64_t)n;
    else total = -1;
    csr_free(&f);
    free(cfg);
    free(g);
    return total;
}

static int in_config(const uint8_t* cops, int k, int r) {
    for (int i = 0; i < k; ++i) if (cops[i] == r) return 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* v1 solver: k_cops_visibility.cpp:85-299, reference pass-sweep order       */
/* ------------------------------------------------------------------------ */
int32_t or_v1_solve(int32_t n, const uint8_t* adj, int32_t k, int32_t p,
                    int32_t cop_stay, int32_t robber_stay, uint8_t* states,
                    uint64_t* base_caps, uint64_t* pass_counts, int32_t max_passes,
                    int32_t* n_passes) {
    if (n <= 0 || n > 255 || p < 1) return -1;
    nbrs_t* g = (nbrs_t*)malloc(sizeof(nbrs_t));
    if (!g) return -1;
    build_nbrs(g, n, adj);
    uint64_t C;
    uint8_t* cfg = make_configs(n, k, &C);
    if (!cfg) { free(g); return -1; }
    csr_t f;
    if (build_csr(g, k, cop_stay, cfg, C, &f) != 0) { free(cfg); free(g); return -1; }
    const int cols = 2 * p;
    const uint64_t N = (uint64_t)n;

    /* initializeCaptures, :85-127: reset every column, seed cols 0 and 1 */
    uint64_t base = 0;
    for (uint64_t c = 0; c < C; ++c)
        for (int r = 0; r < n; ++r) {
            uint8_t* s = states + (c * N + (uint64_t)r) * cols;
            memset(s, UNMARKED_BYTE, (size_t)cols);
            if (in_config(cfg + c * k, k, r)) { s[0] = WON_BYTE(0); s[1] = WON_BYTE(0); base++; }
        }
    if (base_caps) *base_caps = base;

    /* Vertices reachable in exactly p robber steps (max-column BFS, :235-288) */
    uint8_t* reach = (uint8_t*)calloc((size_t)n * n, 1);
    uint8_t* cur = (uint8_t*)malloc((size_t)n);
    uint8_t* nxt = (uint8_t*)malloc((size_t)n);
    if (!reach || !cur || !nxt) { free(reach); free(cur); free(nxt); csr_free(&f); free(cfg); free(g); return -1; }
    for (int r = 0; r < n; ++r) {
        memset(cur, 0, (size_t)n);
        cur[r] = 1;
        for (int step = 0; step < p; ++step) {
            memset(nxt, 0, (size_t)n);
            for (int v = 0; v < n; ++v) {
                if (!cur[v]) continue;
                if (robber_stay) nxt[v] = 1;
                for (int e = 0; e < g->deg[v]; ++e) nxt[g->nb[v][e]] = 1;
            }
            memcpy(cur, nxt, (size_t)n);
        }
        memcpy(reach + (size_t)r * n, cur, (size_t)n);
    }

    int passes = 0;
    for (;;) {
        passes++;
        uint64_t found = 0;
        for (int col = 0; col < cols; ++col) {
            const int nc = (col + 1) % cols;
            for (uint64_t c = 0; c < C; ++c) {
                for (int r = 0; r < n; ++r) {
                    uint8_t* s = states + (c * N + (uint64_t)r) * cols + col;
                    if (IS_MARKED(*s)) continue;
                    unsigned best;
                    int ok;
                    if ((col & 1) == 0) { /* cop turn: exists, min (:160-197) */
                        ok = 0;
                        best = OR_MAX_ROUND;
                        for (uint64_t i = f.head[c]; i < f.head[c + 1]; ++i) {
                            uint8_t q = states[((uint64_t)f.id[i] * N + (uint64_t)r) * cols + nc];
                            if (IS_MARKED(q)) { ok = 1; if (DEPTH_OF(q) < best) best = DEPTH_OF(q); }
                        }
                    } else if (col < cols - 1) { /* interior robber turn: copy (:209-228) */
                        uint8_t q = states[(c * N + (uint64_t)r) * cols + nc];
                        ok = IS_MARKED(q);
                        best = DEPTH_OF(q);
                    } else { /* max column: for all reachable v, max (:235-288) */
                        ok = 1;
                        best = 0;
                        const uint8_t* rr = reach + (size_t)r * n;
                        for (int v = 0; v < n && ok; ++v) {
                            if (!rr[v]) continue;
                            uint8_t q = states[(c * N + (uint64_t)v) * cols + 0];
                            if (!IS_MARKED(q)) ok = 0;
                            else if (DEPTH_OF(q) > best) best = DEPTH_OF(q);
                        }
                    }
                    if (ok && best + 1 < OR_MAX_ROUND) { *s = WON_BYTE(best + 1); found++; }
                }
            }
        }
        if (passes <= max_passes && pass_counts) pass_counts[passes - 1] = found;
        if (found == 0) break;
    }
    if (n_passes) *n_passes = passes;
    free(reach); free(cur); free(nxt);
    csr_free(&f);
    free(cfg);
    free(g);
    return 0;
}

/* findFinalResult, k_cops_visibility.cpp:301-354 */
int32_t or_verdict(int32_t n, int32_t k, int32_t columns, const uint8_t* states,
                   int64_t* best_cid, int32_t* plys) {
    uint64_t C = or_config_count(n, k);
    int64_t best = -1;
    unsigned best_worst = OR_MAX_ROUND;
    for (uint64_t c = 0; c < C; ++c) {
        unsigned worst = 0;
        int all = 1;
        for (int r = 0; r < n; ++r) {
            uint8_t b = states[(c * (uint64_t)n + (uint64_t)r) * (uint64_t)columns];
            if (!IS_MARKED(b)) { all = 0; break; }
            if (DEPTH_OF(b) > worst) worst = DEPTH_OF(b);
        }
        if (all && worst < best_worst) { best_worst = worst; best = (int64_t)c; }
    }
    *best_cid = best;
    *plys = best >= 0 ? (int32_t)best_worst : -1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* v2 solver: k_cops_visibility_v2.cpp:200-565                               */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t w[4]; } vmask;

static inline int vm_empty(const vmask* a) { return (a->w[0] | a->w[1] | a->w[2] | a->w[3]) == 0; }

typedef struct {
    uint64_t key_c;  /* config id, or ~0 for an empty slot */
    vmask cloud;
    int32_t d;
    int32_t val;
} memo_slot;

typedef struct {
    int32_t n, k, p;
    uint64_t C;
    csr_t T;
    vmask nbr[256];
    vmask* cop;     /* C */
    vmask* marked;  /* C, snapshot */
    const uint8_t* snap; /* states snapshot C*N */
    memo_slot* slots;
    uint64_t nslots, used;
} v2ctx;

static uint64_t memo_hash(uint64_t c, const vmask* v, int d) {
    uint64_t h = c * 0x9E3779B97F4A7C15ull ^ (uint64_t)d * 0xC2B2AE3D27D4EB4Full;
    for (int i = 0; i < 4; ++i) { h ^= v->w[i] + 0x165667B19E3779F9ull + (h << 6) + (h >> 2); }
    h ^= h >> 31;
    h *= 0xD6E8FEB86659FD93ull;
    h ^= h >> 32;
    return h;
}

static int memo_grow(v2ctx* x);

static memo_slot* memo_find(v2ctx* x, uint64_t c, const vmask* v, int d, int* found) {
    if ((x->used + 1) * 2 > x->nslots && memo_grow(x) != 0) return NULL;
    uint64_t i = memo_hash(c, v, d) & (x->nslots - 1);
    for (;;) {
        memo_slot* s = &x->slots[i];
        if (s->key_c == ~0ull) { *found = 0; return s; }
        if (s->key_c == c && s->d == d && memcmp(&s->cloud, v, sizeof(vmask)) == 0) { *found = 1; return s; }
        i = (i + 1) & (x->nslots - 1);
    }
}

static int memo_grow(v2ctx* x) {
    uint64_t nn = x->nslots ? x->nslots * 2 : 1 << 16;
    memo_slot* ns = (memo_slot*)malloc(sizeof(memo_slot) * nn);
    if (!ns) return -1;
    for (uint64_t i = 0; i < nn; ++i) ns[i].key_c = ~0ull;
    memo_slot* old = x->slots;
    uint64_t on = x->nslots;
    x->slots = ns;
    x->nslots = nn;
    for (uint64_t i = 0; i < on; ++i) {
        if (old[i].key_c == ~0ull) continue;
        uint64_t j = memo_hash(old[i].key_c, &old[i].cloud, old[i].d) & (nn - 1);
        while (ns[j].key_c != ~0ull) j = (j + 1) & (nn - 1);
        ns[j] = old[i];
    }
    free(old);
    return 0;
}

#define V2_INF 0x7fffffff

/* Value of cop step d from config c with the robber cloud V = V_{d-1}:
 * minimum over team moves of the ply cost (:243-341 without its ordering,
 * cut-off and skip filters, which only decide how fast the minimum is found). */
static int32_t v2_value(v2ctx* x, uint64_t c, const vmask* V, int d) {
    int found;
    memo_slot* s = memo_find(x, c, V, d, &found);
    if (!s) return -2;
    if (found) return s->val;
    int32_t best = V2_INF;
    for (uint64_t i = x->T.head[c]; i < x->T.head[c + 1]; ++i) {
        uint64_t c2 = x->T.id[i];
        const vmask* occ = &x->cop[c2];
        vmask mid;
        for (int w = 0; w < 4; ++w) mid.w[w] = V->w[w] & ~occ->w[w];
        if (vm_empty(&mid)) { best = 2 * d - 1; break; } /* capture on the cop ply */
        vmask nx = {{0, 0, 0, 0}};
        for (int w = 0; w < 4; ++w) {
            uint64_t bits = mid.w[w];
            while (bits) {
                int v = w * 64 + __builtin_ctzll(bits);
                bits &= bits - 1;
                for (int q = 0; q < 4; ++q) nx.w[q] |= x->nbr[v].w[q];
            }
        }
        for (int w = 0; w < 4; ++w) nx.w[w] &= ~occ->w[w];
        int32_t cost;
        if (vm_empty(&nx)) cost = 2 * d; /* robber forced onto a cop */
        else if (d == x->p) {            /* boundary: observe robber in V_p */
            int sub = 1;
            for (int w = 0; w < 4; ++w) if (nx.w[w] & ~x->marked[c2].w[w]) sub = 0;
            if (!sub) continue;
            int32_t mx = 0;
            for (int w = 0; w < 4; ++w) {
                uint64_t bits = nx.w[w];
                while (bits) {
                    int v = w * 64 + __builtin_ctzll(bits);
                    bits &= bits - 1;
                    int32_t dd = DEPTH_OF(x->snap[c2 * (uint64_t)x->n + (uint64_t)v]);
                    if (dd > mx) mx = dd;
                }
            }
            cost = 2 * x->p + mx;
        } else {
            cost = v2_value(x, c2, &nx, d + 1);
            if (cost == -2) return -2;
        }
        if (cost < best) best = cost;
    }
    /* s may have moved if the table grew during recursion: look it up again */
    s = memo_find(x, c, V, d, &found);
    if (!s) return -2;
    if (!found) { s->key_c = c; s->cloud = *V; s->d = d; x->used++; }
    s->val = best;
    return best;
}

int32_t or_v2_solve(int32_t n, const uint8_t* adj, int32_t k, int32_t p,
                    uint8_t* states, uint64_t* base_caps, uint64_t* pass_counts,
                    int32_t max_passes, int32_t* n_passes) {
    if (n <= 0 || n > 255 || p < 1) return -1;
    nbrs_t* g = (nbrs_t*)malloc(sizeof(nbrs_t));
    if (!g) return -1;
    build_nbrs(g, n, adj);
    v2ctx x;
    memset(&x, 0, sizeof x);
    x.n = n; x.k = k; x.p = p;
    uint8_t* cfg = make_configs(n, k, &x.C);
    if (!cfg) { free(g); return -1; }
    if (build_csr(g, k, 0, cfg, x.C, &x.T) != 0) { free(cfg); free(g); return -1; }
    for (int v = 0; v < n; ++v) {
        memset(&x.nbr[v], 0, sizeof(vmask));
        for (int e = 0; e < g->deg[v]; ++e) x.nbr[v].w[g->nb[v][e] >> 6] |= 1ull << (g->nb[v][e] & 63);
    }
    const uint64_t C = x.C, N = (uint64_t)n;
    x.cop = (vmask*)calloc(C, sizeof(vmask));
    x.marked = (vmask*)calloc(C, sizeof(vmask));
    uint8_t* snap = (uint8_t*)malloc(C * N);
    int32_t* pend_cost = (int32_t*)malloc(sizeof(int32_t) * C * N);
    if (!x.cop || !x.marked || !snap || !pend_cost) return -1;
    uint64_t base = 0;
    for (uint64_t c = 0; c < C; ++c) {
        for (int i = 0; i < k; ++i) x.cop[c].w[cfg[c * k + i] >> 6] |= 1ull << (cfg[c * k + i] & 63);
        for (int r = 0; r < n; ++r) {
            int cap = in_config(cfg + c * k, k, r);
            states[c * N + r] = cap ? WON_BYTE(0) : UNMARKED_BYTE;
            base += (uint64_t)cap;
        }
    }
    if (base_caps) *base_caps = base;
    int passes = 0;
    for (;;) {
        passes++;
        memcpy(snap, states, C * N);
        x.snap = snap;
        for (uint64_t c = 0; c < C; ++c) {
            memset(&x.marked[c], 0, sizeof(vmask));
            for (int r = 0; r < n; ++r)
                if (IS_MARKED(snap[c * N + r])) x.marked[c].w[r >> 6] |= 1ull << (r & 63);
        }
        x.used = 0;
        for (uint64_t i = 0; i < x.nslots; ++i) x.slots[i].key_c = ~0ull;
        uint64_t found = 0;
        for (uint64_t c = 0; c < C; ++c)
            for (int r = 0; r < n; ++r) {
                pend_cost[c * N + r] = V2_INF;
                if (IS_MARKED(snap[c * N + r])) continue;
                vmask v0 = {{0, 0, 0, 0}};
                v0.w[r >> 6] = 1ull << (r & 63);
                int32_t cost = v2_value(&x, c, &v0, 1);
                if (cost == -2) return -1;
                if (cost < OR_MAX_ROUND) { pend_cost[c * N + r] = cost; found++; }
            }
        for (uint64_t s = 0; s < C * N; ++s)
            if (pend_cost[s] != V2_INF) states[s] = WON_BYTE(pend_cost[s]);
        if (passes <= max_passes && pass_counts) pass_counts[passes - 1] = found;
        if (found == 0) break;
    }
    if (n_passes) *n_passes = passes;
    free(x.slots); free(x.cop); free(x.marked); free(snap); free(pend_cost);
    csr_free(&x.T);
    free(cfg);
    free(g);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Multithreaded level-synchronous p = 1 solver (legacy/src/k_cops_4.cpp     */
/* :388-481 structure, v1 rules).                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint64_t* v;
    uint64_t n, cap;
} vec64;

static int vec_push(vec64* a, uint64_t x) {
    if (a->n == a->cap) {
        uint64_t nc = a->cap ? a->cap * 2 : 1024;
        uint64_t* g = (uint64_t*)realloc(a->v, nc * sizeof(uint64_t));
        if (!g) return -1;
        a->v = g;
        a->cap = nc;
    }
    a->v[a->n++] = x;
    return 0;
}

typedef struct {
    int32_t n;
    uint64_t C;
    const csr_t* pred;
    const nbrs_t* inn; /* in-neighbours: robber predecessors */
    uint8_t* states;   /* C*N*2 */
    uint8_t* escapes;  /* C*N robber escape counters */
    const vec64* f0;   /* cop-turn frontier (state ids c*N+v) */
    const vec64* f1;   /* robber-turn frontier */
    unsigned depth;    /* frontier depth */
    int nthreads;
} lvl_shared;

typedef struct {
    lvl_shared* sh;
    int tid;
    vec64 n0, n1;
    uint64_t e_cop, e_rob;
    int err;
} lvl_worker;

static void* lvl_run(void* arg) {
    lvl_worker* w = (lvl_worker*)arg;
    lvl_shared* s = w->sh;
    const uint64_t N = (uint64_t)s->n;
    const uint8_t nd = WON_BYTE(s->depth + 1);
    /* robber-turn frontier -> cop predecessors (exists): claim once */
    uint64_t lo = s->f1->n * (uint64_t)w->tid / (uint64_t)s->nthreads;
    uint64_t hi = s->f1->n * (uint64_t)(w->tid + 1) / (uint64_t)s->nthreads;
    for (uint64_t i = lo; i < hi; ++i) {
        uint64_t id = s->f1->v[i], c2 = id / N, r = id % N;
        for (uint64_t j = s->pred->head[c2]; j < s->pred->head[c2 + 1]; ++j) {
            uint64_t st = (uint64_t)s->pred->id[j] * N + r;
            uint8_t* b = &s->states[st * 2 + 0];
            w->e_cop++;
            uint8_t expect = UNMARKED_BYTE;
            if (__atomic_load_n(b, __ATOMIC_RELAXED) != UNMARKED_BYTE) continue;
            if (__atomic_compare_exchange_n(b, &expect, nd, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
                if (vec_push(&w->n0, st)) w->err = 1;
        }
    }
    /* cop-turn frontier -> robber predecessors (for all): count down escapes */
    lo = s->f0->n * (uint64_t)w->tid / (uint64_t)s->nthreads;
    hi = s->f0->n * (uint64_t)(w->tid + 1) / (uint64_t)s->nthreads;
    for (uint64_t i = lo; i < hi; ++i) {
        uint64_t id = s->f0->v[i], c = id / N, v = id % N;
        for (int e = 0; e < s->inn->deg[v]; ++e) {
            uint64_t st = c * N + s->inn->nb[v][e];
            w->e_rob++;
            if (IS_MARKED(__atomic_load_n(&s->states[st * 2 + 1], __ATOMIC_RELAXED))) continue;
            if (__atomic_fetch_sub(&s->escapes[st], 1, __ATOMIC_RELAXED) == 1) {
                __atomic_store_n(&s->states[st * 2 + 1], nd, __ATOMIC_RELAXED);
                if (vec_push(&w->n1, st)) w->err = 1;
            }
        }
    }
    return NULL;
}

int32_t or_v1_levelsync(int32_t n, const uint8_t* adj, int32_t k, int32_t nthreads,
                        uint8_t* states, uint64_t* level_counts, uint64_t* e_cop,
                        uint64_t* e_rob) {
    if (n <= 0 || n > 255 || nthreads < 1 || nthreads > 1024) return -1;
    nbrs_t* g = (nbrs_t*)malloc(sizeof(nbrs_t));
    nbrs_t* inn = (nbrs_t*)malloc(sizeof(nbrs_t));
    if (!g || !inn) return -1;
    build_nbrs(g, n, adj);
    inn->n = n;
    for (int v = 0; v < n; ++v) {
        int d = 0;
        for (int u = 0; u < n; ++u) if (adj[(size_t)u * n + v]) inn->nb[v][d++] = (uint8_t)u;
        inn->deg[v] = d;
    }
    uint64_t C;
    uint8_t* cfg = make_configs(n, k, &C);
    if (!cfg) return -1;
    csr_t f, pr;
    if (build_csr(g, k, 0, cfg, C, &f) != 0) return -1;
    if (transpose_csr(&f, &pr) != 0) return -1;
    csr_free(&f);
    const uint64_t N = (uint64_t)n;
    uint8_t* esc = (uint8_t*)malloc(C * N);
    if (!esc) return -1;
    vec64 f0 = {0}, f1 = {0};
    if (level_counts) memset(level_counts, 0, sizeof(uint64_t) * 2 * OR_MAX_ROUND);
    for (uint64_t c = 0; c < C; ++c)
        for (int r = 0; r < n; ++r) {
            uint64_t st = c * N + (uint64_t)r;
            esc[st] = (uint8_t)g->deg[r];
            if (in_config(cfg + c * k, k, r)) {
                states[st * 2] = WON_BYTE(0);
                states[st * 2 + 1] = WON_BYTE(0);
                vec_push(&f0, st);
                vec_push(&f1, st);
            } else {
                states[st * 2] = UNMARKED_BYTE;
                states[st * 2 + 1] = UNMARKED_BYTE;
            }
        }
    vec64 vac = {0}; /* robber states with no move are won at depth 1 */
    for (uint64_t c = 0; c < C; ++c)
        for (int r = 0; r < n; ++r)
            if (g->deg[r] == 0 && !in_config(cfg + c * k, k, r)) {
                states[(c * N + r) * 2 + 1] = WON_BYTE(1);
                vec_push(&vac, c * N + r);
            }
    if (level_counts) { level_counts[0] = f0.n; level_counts[1] = f1.n; }
    uint64_t ec = 0, er = 0;
    lvl_worker* ws = (lvl_worker*)calloc((size_t)nthreads, sizeof(lvl_worker));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    int err = 0;
    for (unsigned d = 0; d + 1 < OR_MAX_ROUND && (f0.n || f1.n); ++d) {
        lvl_shared sh = {n, C, &pr, inn, states, esc, &f0, &f1, d, nthreads};
        for (int t = 0; t < nthreads; ++t) {
            memset(&ws[t], 0, sizeof(lvl_worker));
            ws[t].sh = &sh;
            ws[t].tid = t;
            pthread_create(&th[t], NULL, lvl_run, &ws[t]);
        }
        vec64 n0 = {0}, n1 = {0};
        for (int t = 0; t < nthreads; ++t) {
            pthread_join(th[t], NULL);
            err |= ws[t].err;
            ec += ws[t].e_cop;
            er += ws[t].e_rob;
            for (uint64_t i = 0; i < ws[t].n0.n; ++i) vec_push(&n0, ws[t].n0.v[i]);
            for (uint64_t i = 0; i < ws[t].n1.n; ++i) vec_push(&n1, ws[t].n1.v[i]);
            free(ws[t].n0.v);
            free(ws[t].n1.v);
        }
        if (d == 0) { for (uint64_t i = 0; i < vac.n; ++i) vec_push(&n1, vac.v[i]); }
        free(f0.v);
        free(f1.v);
        f0 = n0;
        f1 = n1;
        if (level_counts) { level_counts[2 * (d + 1)] = f0.n; level_counts[2 * (d + 1) + 1] = f1.n; }
    }
    free(f0.v); free(f1.v); free(vac.v); free(ws); free(th); free(esc);
    csr_free(&pr);
    free(cfg); free(g); free(inn);
    if (e_cop) *e_cop = ec;
    if (e_rob) *e_rob = er;
    return err ? -1 : 0;
}

int32_t or_retro_edges(int32_t n, const uint8_t* adj, int32_t k, const uint8_t* states,
                       uint64_t* e_cop, uint64_t* e_rob, uint64_t* won) {
    if (n <= 0 || n > 255) return -1;
    nbrs_t* g = (nbrs_t*)malloc(sizeof(nbrs_t));
    if (!g) return -1;
    build_nbrs(g, n, adj);
    int indeg[256] = {0};
    for (int u = 0; u < n; ++u)
        for (int e = 0; e < g->deg[u]; ++e) indeg[g->nb[u][e]]++;
    uint64_t C;
    uint8_t* cfg = make_configs(n, k, &C);
    if (!cfg) { free(g); return -1; }
    csr_t f;
    if (build_csr(g, k, 0, cfg, C, &f) != 0) { free(cfg); free(g); return -1; }
    uint64_t* npred = (uint64_t*)calloc(C, sizeof(uint64_t));
    for (uint64_t i = 0; i < f.head[C]; ++i) npred[f.id[i]]++;
    uint64_t ec = 0, er = 0, w = 0;
    const uint64_t N = (uint64_t)n;
    for (uint64_t c = 0; c < C; ++c)
        for (int r = 0; r < n; ++r) {
            uint8_t b0 = states[(c * N + r) * 2], b1 = states[(c * N + r) * 2 + 1];
            if (IS_MARKED(b0)) { w++; if (DEPTH_OF(b0) <= OR_MAX_ROUND - 2) er += (uint64_t)indeg[r]; }
            if (IS_MARKED(b1)) { w++; if (DEPTH_OF(b1) <= OR_MAX_ROUND - 2) ec += npred[c]; }
        }
    free(npred);
    csr_free(&f);
    free(cfg);
    free(g);
    *e_cop = ec;
    *e_rob = er;
    *won = w;
    return 0;
}

int32_t or_pass_counts_from_depths(int32_t n, int32_t k, const uint8_t* states,
                                   uint64_t* pass_counts, int32_t max_passes) {
    uint64_t C = or_config_count(n, k);
    uint64_t lv[2][OR_MAX_ROUND];
    memset(lv, 0, sizeof lv);
    for (uint64_t s = 0; s < C * (uint64_t)n; ++s)
        for (int t = 0; t < 2; ++t) {
            uint8_t b = states[s * 2 + t];
            if (IS_MARKED(b)) lv[t][DEPTH_OF(b)]++;
        }
    int passes = 0;
    for (int i = 1;; ++i) {
        uint64_t cnt = 0;
        int c0 = 2 * i - 2, c1 = 2 * i - 1, r0 = 2 * i - 1, r1 = 2 * i;
        if (c0 > 0 && c0 < OR_MAX_ROUND) cnt += lv[0][c0];
        if (c1 < OR_MAX_ROUND) cnt += lv[0][c1];
        if (r0 < OR_MAX_ROUND) cnt += lv[1][r0];
        if (r1 < OR_MAX_ROUND) cnt += lv[1][r1];
        if (i <= max_passes && pass_counts) pass_counts[i - 1] = cnt;
        passes = i;
        if (cnt == 0) break;
    }
    return passes;
}

/* ------------------------------------------------------------------------ */
/* Checks of a solved p = 1 table at sizes the sweep cannot reach            */
/* ------------------------------------------------------------------------ */
static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Inverse of or_rank: the non-decreasing tuple at lexicographic position
 * `rank` (AuxGraph.h:155-176 enumeration order). A prefix ending in value v
 * at position i is followed by C(N-v+k-i-2, k-i-1) completions. */
int32_t or_unrank(uint64_t rank, int32_t n, int32_t k, uint8_t* out) {
    if (k <= 0 || k > OR_MAX_K || n <= 0 || rank >= or_config_count(n, k)) return -1;
    int v = 0;
    for (int i = 0; i < k; ++i) {
        for (;; ++v) {
            uint64_t span = binom((int64_t)n - v + k - i - 2, k - i - 1);
            if (rank < span) break;
            rank -= span;
        }
        out[i] = (uint8_t)v;
    }
    return 0;
}

/* The value a state must hold given its successors' values in the same table:
 * the local (Bellman) form of solvers/k_cops_visibility.cpp:85-299 at p = 1.
 *   capture (r in C, both turns, :98-110): marked at round 0;
 *   cop turn (:160-197): 1 + min over team moves C' in T(C) of the marked
 *     robber-turn rounds of (C', r);
 *   robber turn (:235-288): 1 + max over r -> v of the cop-turn rounds of
 *     (C, v) when all of them are marked (vacuous for a robber with no move);
 * and a state is marked only below the cap (round + 1 < 127, :189, :280).
 * Applied at every state, these equations have exactly one solution (by
 * induction on the round: a state holds round d iff its true depth is d), so
 * a table that satisfies them everywhere is the reference's table. */
static uint8_t expected_byte(const nbrs_t* g, int32_t k, const uint8_t* states,
                             const uint8_t* cops, uint64_t c, int r, int t, uint64_t* buf,
                             uint64_t cap) {
    const uint64_t N = (uint64_t)g->n;
    if (in_config(cops, k, r)) return WON_BYTE(0);
    unsigned best;
    if (t == 0) {
        int64_t m = moves_of(g, cops, k, 0, buf, cap);
        best = OR_MAX_ROUND;
        for (int64_t i = 0; i < m; ++i) {
            uint8_t b = states[(buf[i] * N + (uint64_t)r) * 2 + 1];
            if (IS_MARKED(b) && DEPTH_OF(b) < best) best = DEPTH_OF(b);
        }
        if (best == OR_MAX_ROUND) return UNMARKED_BYTE;
    } else {
        best = 0;
        for (int e = 0; e < g->deg[r]; ++e) {
            uint8_t b = states[(c * N + g->nb[r][e]) * 2];
            if (!IS_MARKED(b)) return UNMARKED_BYTE;
            if (DEPTH_OF(b) > best) best = DEPTH_OF(b);
        }
    }
    return best + 1 < OR_MAX_ROUND ? WON_BYTE(best + 1) : UNMARKED_BYTE;
}

typedef struct {
    const nbrs_t* g;
    int32_t k;
    const uint8_t* states;
    const uint64_t* ids;
    uint64_t lo, hi;
    uint64_t bad, first_bad;
    int err;
} chk_job;

static void* chk_run(void* arg) {
    chk_job* j = (chk_job*)arg;
    const uint64_t N = (uint64_t)j->g->n;
    uint64_t cap = product_bound(j->g, j->k, 0);
    uint64_t* buf = (uint64_t*)malloc(sizeof(uint64_t) * (cap ? cap : 1));
    if (!buf) { j->err = 1; return NULL; }
    uint8_t cops[OR_MAX_K];
    uint64_t last_c = UINT64_MAX;
    for (uint64_t i = j->lo; i < j->hi; ++i) {
        uint64_t id = j->ids[i], t = id & 1, c = (id >> 1) / N;
        int r = (int)((id >> 1) % N);
        if (c != last_c) { or_unrank(c, j->g->n, j->k, cops); last_c = c; }
        uint8_t want = expected_byte(j->g, j->k, j->states, cops, c, r, (int)t, buf, cap);
        if (j->states[id] != want) {
            if (j->bad++ == 0) j->first_bad = id;
        }
    }
    free(buf);
    return NULL;
}

int32_t or_check_states(int32_t n, const uint8_t* adj, int32_t k, const uint8_t* states,
                        const uint64_t* ids, uint64_t count, int32_t nthreads, uint64_t* bad,
                        uint64_t* first_bad) {
    if (n <= 0 || n > 255 || k <= 0 || k > OR_MAX_K || nthreads < 1 || nthreads > 1024) return -1;
    uint64_t S = or_config_count(n, k) * (uint64_t)n * 2;
    for (uint64_t i = 0; i < count; ++i) if (ids[i] >= S) return -1;
    nbrs_t* g = (nbrs_t*)malloc(sizeof(nbrs_t));
    chk_job* jobs = (chk_job*)calloc((size_t)nthreads, sizeof(chk_job));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    if (!g || !jobs || !th) { free(g); free(jobs); free(th); return -1; }
    build_nbrs(g, n, adj);
    for (int t = 0; t < nthreads; ++t) {
        jobs[t] = (chk_job){g, k, states, ids, count * (uint64_t)t / (uint64_t)nthreads,
                            count * (uint64_t)(t + 1) / (uint64_t)nthreads, 0, 0, 0};
        pthread_create(&th[t], NULL, chk_run, &jobs[t]);
    }
    uint64_t nb = 0, fb = UINT64_MAX;
    int err = 0;
    for (int t = 0; t < nthreads; ++t) {
        pthread_join(th[t], NULL);
        err |= jobs[t].err;
        if (jobs[t].bad && nb == 0) fb = jobs[t].first_bad;
        nb += jobs[t].bad;
    }
    free(g); free(jobs); free(th);
    *bad = nb;
    *first_bad = fb;
    return err ? -1 : 0;
}

/* One pass over a p = 1 table: histogram of marked states by (round, turn)
 * and the verdict of findFinalResult (k_cops_visibility.cpp:301-354: the
 * configuration whose column-0 states are all marked with the least worst
 * round, lowest cId on ties). Rows are split over threads; each thread keeps
 * its own best, and the lowest cId wins among equal worst rounds. */
typedef struct {
    int32_t n;
    const uint8_t* states;
    uint64_t lo, hi;
    uint64_t lv[2][OR_MAX_ROUND];
    int64_t best;
    unsigned best_worst;
} sum_job;

static void* sum_run(void* arg) {
    sum_job* j = (sum_job*)arg;
    const uint64_t N = (uint64_t)j->n;
    j->best = -1;
    j->best_worst = OR_MAX_ROUND;
    for (uint64_t c = j->lo; c < j->hi; ++c) {
        const uint8_t* row = j->states + c * N * 2;
        unsigned worst = 0;
        int all = 1;
        for (uint64_t r = 0; r < N; ++r) {
            uint8_t b0 = row[2 * r], b1 = row[2 * r + 1];
            if (IS_MARKED(b0)) { j->lv[0][DEPTH_OF(b0)]++; if (DEPTH_OF(b0) > worst) worst = DEPTH_OF(b0); }
            else all = 0;
            if (IS_MARKED(b1)) j->lv[1][DEPTH_OF(b1)]++;
        }
        if (all && worst < j->best_worst) { j->best_worst = worst; j->best = (int64_t)c; }
    }
    return NULL;
}

int32_t or_table_summary(int32_t n, int32_t k, const uint8_t* states, int32_t nthreads,
                         uint64_t* level_counts, int64_t* best_cid, int32_t* plys) {
    if (n <= 0 || n > 255 || nthreads < 1 || nthreads > 1024) return -1;
    uint64_t C = or_config_count(n, k);
    sum_job* jobs = (sum_job*)calloc((size_t)nthreads, sizeof(sum_job));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    if (!jobs || !th) { free(jobs); free(th); return -1; }
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].n = n;
        jobs[t].states = states;
        jobs[t].lo = C * (uint64_t)t / (uint64_t)nthreads;
        jobs[t].hi = C * (uint64_t)(t + 1) / (uint64_t)nthreads;
        pthread_create(&th[t], NULL, sum_run, &jobs[t]);
    }
    memset(level_counts, 0, sizeof(uint64_t) * 2 * OR_MAX_ROUND);
    int64_t best = -1;
    unsigned best_worst = OR_MAX_ROUND;
    for (int t = 0; t < nthreads; ++t) { /* threads in cId order: strict < keeps the lowest */
        pthread_join(th[t], NULL);
        for (int d = 0; d < OR_MAX_ROUND; ++d) {
            level_counts[2 * d] += jobs[t].lv[0][d];
            level_counts[2 * d + 1] += jobs[t].lv[1][d];
        }
        if (jobs[t].best >= 0 && jobs[t].best_worst < best_worst) {
            best_worst = jobs[t].best_worst;
            best = jobs[t].best;
        }
    }
    free(jobs); free(th);
    *best_cid = best;
    *plys = best >= 0 ? (int32_t)best_worst : -1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Level-synchronous p = 1 solver with moves generated on the fly and a time */
/* cap (the CPU baseline at sizes where a move table does not fit).          */
/* Structure of legacy/src/k_cops_4.cpp:388-481 (threads over the frontier,  */
/* atomic claims and escape counters), v1 rules of k_cops_visibility.cpp.    */
/* Undirected graphs only: the predecessors of (C', r, 1) are then           */
/* {(C, r, 0) : C in T(C')}. The frontier of depth d is found by scanning    */
/* the table for that depth instead of being kept in lists.                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    const nbrs_t* g;
    int32_t k;
    uint64_t C;
    uint8_t* states;
    uint8_t* escapes;
    unsigned depth;
    double deadline;
    int* stop;          /* set (atomically) by the first thread past the deadline */
    uint64_t* next_row; /* shared chunk counter */
    int phase;          /* 0 = init, 1 = level */
} cap_shared;

typedef struct {
    cap_shared* sh;
    uint64_t e_cop, e_rob, rows;
    int err;
} cap_worker;

#define CAP_CHUNK 256

static void* cap_run(void* arg) {
    cap_worker* w = (cap_worker*)arg;
    cap_shared* s = w->sh;
    const nbrs_t* g = s->g;
    const uint64_t N = (uint64_t)g->n;
    uint64_t cap = product_bound(g, s->k, 0);
    uint64_t* buf = (uint64_t*)malloc(sizeof(uint64_t) * (cap ? cap : 1));
    if (!buf) { w->err = 1; return NULL; }
    uint8_t cops[OR_MAX_K];
    const uint8_t nd = WON_BYTE(s->depth + 1);
    for (;;) {
        if (__atomic_load_n(s->stop, __ATOMIC_RELAXED)) break;
        uint64_t c0 = __atomic_fetch_add(s->next_row, CAP_CHUNK, __ATOMIC_RELAXED);
        if (c0 >= s->C) break;
        uint64_t c1 = c0 + CAP_CHUNK < s->C ? c0 + CAP_CHUNK : s->C;
        for (uint64_t c = c0; c < c1; ++c) {
            or_unrank(c, g->n, s->k, cops);
            uint8_t* row = s->states + c * N * 2;
            if (s->phase == 0) {
                for (uint64_t r = 0; r < N; ++r) {
                    uint8_t b = in_config(cops, s->k, (int)r) ? WON_BYTE(0) : UNMARKED_BYTE;
                    row[2 * r] = b;
                    row[2 * r + 1] = b;
                    s->escapes[c * N + r] = (uint8_t)g->deg[r];
                }
                continue;
            }
            int moves_known = 0;
            int64_t m = 0;
            for (uint64_t r = 0; r < N; ++r) {
                /* robber-turn state of this depth -> cop-turn predecessors */
                uint8_t b1 = __atomic_load_n(&row[2 * r + 1], __ATOMIC_RELAXED);
                if (b1 == WON_BYTE(s->depth)) {
                    if (!moves_known) { m = moves_of(g, cops, s->k, 0, buf, cap); moves_known = 1; }
                    for (int64_t i = 0; i < m; ++i) {
                        uint8_t* b = &s->states[(buf[i] * N + r) * 2];
                        uint8_t expect = UNMARKED_BYTE;
                        w->e_cop++;
                        if (__atomic_load_n(b, __ATOMIC_RELAXED) != UNMARKED_BYTE) continue;
                        __atomic_compare_exchange_n(b, &expect, nd, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED);
                    }
                }
                /* cop-turn state of this depth -> robber-turn predecessors */
                uint8_t b0 = __atomic_load_n(&row[2 * r], __ATOMIC_RELAXED);
                if (b0 == WON_BYTE(s->depth)) {
                    for (int e = 0; e < g->deg[r]; ++e) {
                        uint64_t u = g->nb[r][e]; /* undirected: in-neighbours = neighbours */
                        w->e_rob++;
                        uint8_t* b = &row[2 * u + 1];
                        if (IS_MARKED(__atomic_load_n(b, __ATOMIC_RELAXED))) continue;
                        if (__atomic_fetch_sub(&s->escapes[c * N + u], 1, __ATOMIC_RELAXED) == 1)
                            __atomic_store_n(b, nd, __ATOMIC_RELAXED);
                    }
                }
            }
        }
        w->rows += c1 - c0;
        if (now_s() > s->deadline) __atomic_store_n(s->stop, 1, __ATOMIC_RELAXED);
    }
    free(buf);
    return NULL;
}

int32_t or_v1_levelsync_capped(int32_t n, const uint8_t* adj, int32_t k, int32_t nthreads,
                               double time_cap_s, uint8_t* states, uint64_t* e_cop,
                               uint64_t* e_rob, int32_t* levels_done, double* level_fraction,
                               double* elapsed_s) {
    if (n <= 0 || n > 255 || k <= 0 || k > OR_MAX_K || nthreads < 1 || nthreads > 1024) return -1;
    for (int u = 0; u < n; ++u)
        for (int v = 0; v < n; ++v)
            if (adj[(size_t)u * n + v] != adj[(size_t)v * n + u]) return -1;
    nbrs_t* g = (nbrs_t*)malloc(sizeof(nbrs_t));
    if (!g) return -1;
    build_nbrs(g, n, adj);
    for (int u = 0; u < n; ++u) if (g->deg[u] == 0) { free(g); return -1; }
    uint64_t C = or_config_count(n, k);
    uint8_t* esc = (uint8_t*)malloc(C * (uint64_t)n);
    cap_worker* ws = (cap_worker*)calloc((size_t)nthreads, sizeof(cap_worker));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    if (!esc || !ws || !th) { free(g); free(esc); free(ws); free(th); return -1; }
    double t0 = now_s();
    int stop = 0;
    uint64_t ec = 0, er = 0;
    int done = 0, err = 0;
    double frac = 0.0;
    /* phase 0 = the captures (initializeCaptures, :85-127), then one phase
     * per depth until a level marks nothing or the cap stops it */
    for (int phase = 0; !__atomic_load_n(&stop, __ATOMIC_RELAXED); ++phase) {
        unsigned depth = phase == 0 ? 0 : (unsigned)(phase - 1);
        if (phase > 0 && depth + 1 >= OR_MAX_ROUND) break;
        uint64_t next = 0;
        cap_shared sh = {g, k, C, states, esc, depth, t0 + time_cap_s, &stop, &next, phase == 0 ? 0 : 1};
        for (int t = 0; t < nthreads; ++t) {
            memset(&ws[t], 0, sizeof(cap_worker));
            ws[t].sh = &sh;
            pthread_create(&th[t], NULL, cap_run, &ws[t]);
        }
        uint64_t rows = 0, lec = 0, ler = 0;
        for (int t = 0; t < nthreads; ++t) {
            pthread_join(th[t], NULL);
            err |= ws[t].err;
            rows += ws[t].rows;
            lec += ws[t].e_cop;
            ler += ws[t].e_rob;
        }
        ec += lec;
        er += ler;
        if (phase == 0) {
            if (rows < C) { frac = 0.0; break; } /* the cap ended the captures */
            continue;
        }
        if (rows < C) { frac = (double)rows / (double)C; break; }
        done++;
        if (lec == 0 && ler == 0) break; /* no state of this depth: fixed point */
    }
    if (elapsed_s) *elapsed_s = now_s() - t0;
    free(g); free(esc); free(ws); free(th);
    *e_cop = ec;
    *e_rob = er;
    *levels_done = done;
    *level_fraction = frac;
    return err ? -1 : 0;
}
