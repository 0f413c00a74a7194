// macro.hip -- the limited-visibility solver (solvers/k_cops_visibility_v2.cpp)
// on the device.
//
// Game (k_cops_visibility_v2.cpp:243-341, 343-565; docs/k_cops_visibility_v2.md):
// one column of states (C, r), cops to move. The cops see the robber every p
// moves, so a macro turn is p cop steps against the robber's "cloud": V_1 = {r},
// and after the cops move C_d -> C_{d+1} in T(C_d) (every cop to a neighbour)
//   * V_d \ C_{d+1} empty           -> caught on the cop ply, cost 2d-1;
//   * V_{d+1} = N(V_d \ C_{d+1}) \ C_{d+1} empty -> no safe robber move, cost 2d;
//   * d = p: the robber is seen again; if every vertex of V_{p+1} is a marked
//     state (C_{p+1}, v) of the snapshot the cost is 2p + max depth, else no win;
//   * otherwise the cost is the value of (C_{d+1}, V_{d+1}) at step d+1.
// The value of (C, {r}) is the minimum over trajectories; an unmarked state is
// marked with it when it is below the 127 cap. Passes use the table as it was
// at the start of the pass and stop after the first pass that marks nothing.
//
// Device form:
//   1. Forward: the trajectory DAG. Step-1 nodes are the states themselves.
//      A node with a terminal move (cost 2d-1 or 2d) has that value (every
//      deeper cost is at least 2d+1), so only the other nodes expand: each
//      (node, team move) pair becomes a candidate (C', V') of step d+1, and
//      equal candidates are merged in a device hash table (open addressing,
//      the slot holds the index of the first candidate with that key; keys
//      are compared in the candidate arrays, so nothing waits on a lock).
//      The DAG does not depend on the table and is built once per solve.
//   2. Per pass, from the snapshot: for every configuration C and robber
//      vertex u, R(C, u) = max depth of the escapes N(u) \ C (127 if one is
//      unmarked, 0 if there is none), stored as 7 bit-planes per C so that
//      max_{u in X} R(C, u) over a cloud X is 7 masked word tests.
//      Step p nodes then take min over moves of 2p + max R (or 2p-1 on a
//      capture), steps p-1..1 take the minimum over their DAG edges, and the
//      step-1 values mark the states.
// All values saturate at 127 ("not won within the cap"), so they fit a byte.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "cr_macro.h"

namespace {

constexpr uint8_t kOpen = 0xFF;              // value comes from the node's DAG edges
constexpr uint32_t kEmptySlot = 0xFFFFFFFFu;
constexpr uint32_t kCap = kMaxRound;         // 127
constexpr int kPlanes = 7;                   // bits of a depth value (<= 127)
constexpr uint32_t kMaxBlocks = 2048;

__device__ __forceinline__ uint32_t sat(uint32_t v) { return v < kCap ? v : kCap; }

// Nodes of one cop step. cfg == nullptr: step 1, node i is the state
// (i / n, {i % n}).
struct Nodes {
    uint64_t n;
    const uint32_t* cfg;
    const uint64_t* cloud;  // [n * W]
};

template <int W>
__device__ __forceinline__ void node_key(const Nodes& nd, uint64_t i, int n, uint32_t& c, uint64_t* V,
                                         int& robber) {
    if (nd.cfg == nullptr) {
        c = static_cast<uint32_t>(i / static_cast<uint64_t>(n));
        robber = static_cast<int>(i - static_cast<uint64_t>(c) * n);
#pragma unroll
        for (int q = 0; q < W; ++q) V[q] = (robber >> 6) == q ? (1ull << (robber & 63)) : 0ull;
    } else {
        c = nd.cfg[i];
        robber = -1;
#pragma unroll
        for (int q = 0; q < W; ++q) V[q] = nd.cloud[i * W + q];
    }
}

template <int K, int W>
__device__ __forceinline__ void tuple_mask(const uint32_t* tup, uint64_t* occ) {
#pragma unroll
    for (int q = 0; q < W; ++q) occ[q] = 0;
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
        for (int q = 0; q < W; ++q)
            if ((tup[j] >> 6) == static_cast<uint32_t>(q)) occ[q] |= 1ull << (tup[j] & 63);
}

// Neighbourhood of a vertex set minus the cops: OR of the robber move masks.
template <int W>
__device__ __forceinline__ bool expand_cloud(const uint64_t* mid, const uint64_t* occ, const uint64_t* lnbr,
                                             uint64_t* nx) {
#pragma unroll
    for (int q = 0; q < W; ++q) nx[q] = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
        uint64_t bits = mid[q];
        while (bits) {
            const int v = q * 64 + __builtin_ctzll(bits);
            bits &= bits - 1;
#pragma unroll
            for (int w = 0; w < W; ++w) nx[w] |= lnbr[v * W + w];
        }
    }
    uint64_t any = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
        nx[q] &= ~occ[q];
        any |= nx[q];
    }
    return any != 0;
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

template <int W>
__device__ __forceinline__ uint64_t key_hash(uint32_t c, const uint64_t* V) {
    uint64_t h = mix64(static_cast<uint64_t>(c) + 0x9e3779b97f4a7c15ull);
#pragma unroll
    for (int q = 0; q < W; ++q) h = mix64(h ^ V[q]);
    return h;
}

// LDS layout of the expand and leaf kernels: graph image, then n*W robber masks.
__host__ __device__ inline uint32_t nbr_offset(uint32_t image_bytes) { return (image_bytes + 15u) & ~15u; }

template <int W>
__device__ __forceinline__ const uint64_t* stage_nbr(uint8_t* lds, const GraphView& gv, const uint64_t* nbr_mask) {
    uint64_t* lnbr = reinterpret_cast<uint64_t*>(lds + nbr_offset(gv.image_bytes));
    for (int i = threadIdx.x; i < gv.n * W; i += blockDim.x) lnbr[i] = nbr_mask[(i / W) * 4 + (i % W)];
    __syncthreads();
    return lnbr;
}

// ---------------------------------------------------------------------------
// captures (k_cops_visibility_v2.cpp:200-223); one wavefront per (C, word q)
// ---------------------------------------------------------------------------
template <int K, int W>
__global__ __launch_bounds__(kBlock) void k_mx_init(GraphView gv, uint8_t* table, uint64_t* occ,
                                                    unsigned long long* base) {
    extern __shared__ __align__(16) uint8_t lds[];
    load_image(lds, gv.image, gv.image_bytes);
    const uint32_t* binom = reinterpret_cast<const uint32_t*>(lds + gv.off_binom);
    const int lane = threadIdx.x & 63;
    const uint64_t nw = static_cast<uint64_t>(gv.C) * W;
    unsigned long long caps = 0, unused = 0;
    for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * kWavesPerBlock + wave_id(); g < nw;
         g += static_cast<uint64_t>(gridDim.x) * kWavesPerBlock) {
        const uint32_t c = static_cast<uint32_t>(g / W);
        const int q = static_cast<int>(g - static_cast<uint64_t>(c) * W);
        uint32_t cops[K];
        unrank<K>(c, gv, binom, cops);
        const int u = 64 * q + lane;
        bool cap = false;
#pragma unroll
        for (int i = 0; i < K; ++i) cap = cap || cops[i] == static_cast<uint32_t>(u);
        if (u < gv.n) table[static_cast<uint64_t>(c) * gv.n + u] = cap ? won_byte(0) : kUnmarked;
        const uint64_t m = __ballot(cap && u < gv.n);
        if (lane == 0) {
            occ[g] = m;
            caps += __builtin_popcountll(m);
        }
    }
    block_sum2(caps, unused);
    if (threadIdx.x == 0 && caps) atomicAdd(base, caps);
}

// ---------------------------------------------------------------------------
// forward: terminal values and successor candidates of one cop step.
// Count mode (off == nullptr): val[i] = terminal value or kOpen, cnt[i] = the
// node's candidates (0 for terminal nodes). Emit mode: writes the candidates
// of open nodes at off[i]...
// ---------------------------------------------------------------------------
template <int K, int W>
__global__ __launch_bounds__(kBlock) void k_mx_expand(GraphView gv, const uint64_t* nbr_mask, Nodes nodes,
                                                      uint32_t d, uint32_t* cnt, uint8_t* val,
                                                      const unsigned long long* off, uint32_t* cand_cfg,
                                                      uint64_t* cand_cloud, unsigned long long* moves) {
    extern __shared__ __align__(16) uint8_t lds[];
    load_image(lds, gv.image, gv.image_bytes);
    const uint64_t* lnbr = stage_nbr<W>(lds, gv, nbr_mask);
    const uint32_t* rank_tab = reinterpret_cast<const uint32_t*>(lds);
    const uint32_t* binom = reinterpret_cast<const uint32_t*>(lds + gv.off_binom);
    const uint16_t* out_off = reinterpret_cast<const uint16_t*>(lds + gv.off_out_off);
    const uint8_t* out_nbr = lds + gv.off_out_nbr;
    const bool emit = off != nullptr;
    const uint32_t t_capture = sat(2 * d - 1), t_trapped = sat(2 * d);
    unsigned long long mv = 0, unused = 0;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nodes.n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        unsigned long long o = 0;
        if (emit) {
            o = off[i];
            if (off[i + 1] == o) continue;
        }
        uint32_t c;
        uint64_t V[W];
        int robber;
        node_key<W>(nodes, i, gv.n, c, V, robber);
        uint32_t cops[K], deg[K], base[K];
        unrank<K>(c, gv, binom, cops);
        bool captured = false;
#pragma unroll
        for (int j = 0; j < K; ++j) captured = captured || static_cast<int>(cops[j]) == robber;
        if (captured) {  // base capture: depth 0, no successors
            val[i] = 0;
            cnt[i] = 0;
            continue;
        }
        uint32_t P = 1;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            base[j] = out_off[cops[j]];
            deg[j] = out_off[cops[j] + 1] - base[j];
            P *= deg[j];
        }
        uint32_t term = kOpen, kept = 0;
        for (uint32_t t = 0; t < P; ++t) {
            uint32_t tup[K];
            if (!product_tuple<K>(t, cops, deg, base, out_nbr, tup)) continue;
            ++mv;
            uint64_t occ[W], mid[W], nx[W];
            tuple_mask<K, W>(tup, occ);
            uint64_t any = 0;
#pragma unroll
            for (int q = 0; q < W; ++q) {
                mid[q] = V[q] & ~occ[q];
                any |= mid[q];
            }
            if (any == 0) {
                term = t_capture;
                break;
            }
            if (!expand_cloud<W>(mid, occ, lnbr, nx)) {
                term = min(term, t_trapped);
                continue;
            }
            if (emit) {
                const unsigned long long e = o + kept;
                cand_cfg[e] = rank_sorted<K>(tup, rank_tab, gv.n, gv.ctot_m1);
#pragma unroll
                for (int q = 0; q < W; ++q) cand_cloud[e * W + q] = nx[q];
            }
            ++kept;
        }
        if (!emit) {
            const bool terminal = term != kOpen;
            val[i] = terminal ? static_cast<uint8_t>(term) : kOpen;
            cnt[i] = terminal ? 0u : kept;
        }
    }
    if (!emit) {
        mv = wave_sum(mv);
        block_sum2(mv, unused);
        if (threadIdx.x == 0 && mv) atomicAdd(moves, mv);
    }
}

// Merges equal candidates: rep[e] = index of the candidate that owns e's key.
template <int W>
__global__ __launch_bounds__(kBlock) void k_mx_insert(const uint32_t* cfg, const uint64_t* cloud, uint32_t ncand,
                                                      uint32_t* slots, uint64_t mask, uint32_t* rep) {
    for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ncand;
         e += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint32_t c = cfg[e];
        uint64_t V[W];
#pragma unroll
        for (int q = 0; q < W; ++q) V[q] = cloud[static_cast<uint64_t>(e) * W + q];
        uint64_t h = key_hash<W>(c, V) & mask;
        // at most ncand of the >= 2*ncand slots are taken, so the probe ends
        for (;;) {
            const uint32_t prev = atomicCAS(&slots[h], kEmptySlot, static_cast<uint32_t>(e));
            if (prev == kEmptySlot) {
                rep[e] = static_cast<uint32_t>(e);
                break;
            }
            bool same = cfg[prev] == c;
#pragma unroll
            for (int q = 0; q < W; ++q) same = same && cloud[static_cast<uint64_t>(prev) * W + q] == V[q];
            if (same) {
                rep[e] = prev;
                break;
            }
            h = (h + 1) & mask;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_mx_flag(const uint32_t* rep, uint32_t ncand, uint32_t* flag) {
    for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ncand;
         e += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        flag[e] = rep[e] == e ? 1u : 0u;
}

// Owners become the next step's nodes (in candidate order); every candidate
// becomes a DAG edge to its owner's node.
template <int W>
__global__ __launch_bounds__(kBlock) void k_mx_compact(const uint32_t* cand_cfg, const uint64_t* cand_cloud,
                                                       const uint32_t* rep, const unsigned long long* pos,
                                                       uint32_t ncand, uint32_t* node_cfg, uint64_t* node_cloud,
                                                       uint32_t* child) {
    for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ncand;
         e += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint32_t r = rep[e];
        const uint32_t id = static_cast<uint32_t>(pos[r]);
        child[e] = id;
        if (r == e) {
            node_cfg[id] = cand_cfg[e];
#pragma unroll
            for (int q = 0; q < W; ++q) node_cloud[static_cast<uint64_t>(id) * W + q] = cand_cloud[static_cast<uint64_t>(e) * W + q];
        }
    }
}

// ---------------------------------------------------------------------------
// per pass: bit-planes of R(C, u) from the snapshot; one wavefront per (C, q)
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(kBlock) void k_mx_planes(int32_t n, uint32_t C, const uint64_t* nbr_mask,
                                                      const uint8_t* table, const uint64_t* occ, uint64_t* planes) {
    __shared__ uint64_t lnbr[CR_MAX_N * W];
    for (int i = threadIdx.x; i < n * W; i += blockDim.x) lnbr[i] = nbr_mask[(i / W) * 4 + (i % W)];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const uint64_t nw = static_cast<uint64_t>(C) * W;
    for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * kWavesPerBlock + wave_id(); g < nw;
         g += static_cast<uint64_t>(gridDim.x) * kWavesPerBlock) {
        const uint32_t c = static_cast<uint32_t>(g / W);
        const int q = static_cast<int>(g - static_cast<uint64_t>(c) * W);
        const uint8_t* row = table + static_cast<uint64_t>(c) * n;
        uint64_t oc[W];
#pragma unroll
        for (int w = 0; w < W; ++w) oc[w] = occ[static_cast<uint64_t>(c) * W + w];
        const int u = 64 * q + lane;
        uint32_t R = 0;
        if (u < n) {
            bool bad = false;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                uint64_t bits = lnbr[u * W + w] & ~oc[w];
                while (bits && !bad) {
                    const int v = w * 64 + __builtin_ctzll(bits);
                    bits &= bits - 1;
                    const uint32_t b = row[v];
                    if (!(b & 1u)) bad = true;
                    else R = max(R, b >> 1);
                }
            }
            if (bad) R = kCap;
        }
        uint64_t* dst = planes + static_cast<uint64_t>(c) * kPlanes * W + q;
#pragma unroll
        for (int b = 0; b < kPlanes; ++b) {
            const uint64_t m = __ballot((R >> b) & 1u);
            if (lane == b) dst[b * W] = m;
        }
    }
}

// Step-p values: min over team moves of 2p-1 (capture) or 2p + max R.
template <int K, int W>
__global__ __launch_bounds__(kBlock) void k_mx_leaf(GraphView gv, Nodes nodes, const uint64_t* planes, uint32_t p,
                                                    uint8_t* val, const uint8_t* table,
                                                    unsigned long long* moves) {
    extern __shared__ __align__(16) uint8_t lds[];
    load_image(lds, gv.image, gv.image_bytes);
    const uint32_t* rank_tab = reinterpret_cast<const uint32_t*>(lds);
    const uint32_t* binom = reinterpret_cast<const uint32_t*>(lds + gv.off_binom);
    const uint16_t* out_off = reinterpret_cast<const uint16_t*>(lds + gv.off_out_off);
    const uint8_t* out_nbr = lds + gv.off_out_nbr;
    const uint32_t t_capture = sat(2 * p - 1);
    unsigned long long mv = 0, unused = 0;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nodes.n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        uint32_t c;
        uint64_t V[W];
        int robber;
        node_key<W>(nodes, i, gv.n, c, V, robber);
        // step 1 is the leaf step (p = 1): states already marked keep their value
        if (robber >= 0 && (table[i] & 1u)) {
            val[i] = 0;
            continue;
        }
        uint32_t cops[K], deg[K], base[K];
        unrank<K>(c, gv, binom, cops);
        uint32_t P = 1;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            base[j] = out_off[cops[j]];
            deg[j] = out_off[cops[j] + 1] - base[j];
            P *= deg[j];
        }
        uint32_t best = kCap;
        for (uint32_t t = 0; t < P; ++t) {
            uint32_t tup[K];
            if (!product_tuple<K>(t, cops, deg, base, out_nbr, tup)) continue;
            ++mv;
            uint64_t occ[W], cand[W];
            tuple_mask<K, W>(tup, occ);
            uint64_t any = 0;
#pragma unroll
            for (int q = 0; q < W; ++q) {
                cand[q] = V[q] & ~occ[q];
                any |= cand[q];
            }
            if (any == 0) {
                best = min(best, t_capture);
                break;
            }
            const uint64_t* pl = planes + static_cast<uint64_t>(rank_sorted<K>(tup, rank_tab, gv.n, gv.ctot_m1)) * kPlanes * W;
            uint32_t mx = 0;
#pragma unroll
            for (int b = kPlanes - 1; b >= 0; --b) {
                uint64_t hit[W], h = 0;
#pragma unroll
                for (int q = 0; q < W; ++q) {
                    hit[q] = cand[q] & pl[b * W + q];
                    h |= hit[q];
                }
                if (h) {
                    mx |= 1u << b;
#pragma unroll
                    for (int q = 0; q < W; ++q) cand[q] = hit[q];
                }
            }
            best = min(best, sat(2 * p + mx));
        }
        val[i] = static_cast<uint8_t>(best);
    }
    mv = wave_sum(mv);
    block_sum2(mv, unused);
    if (threadIdx.x == 0 && mv) atomicAdd(moves, mv);
}

// Steps p-1..1: open nodes take the minimum over their DAG edges.
__global__ __launch_bounds__(kBlock) void k_mx_up(uint64_t n, const unsigned long long* off, const uint32_t* child,
                                                  const uint8_t* next_val, uint8_t* val) {
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const unsigned long long b = off[i], e = off[i + 1];
        if (b == e) continue;  // terminal: value fixed when the DAG was built
        uint32_t m = kCap;
        for (unsigned long long j = b; j < e; ++j) m = min(m, static_cast<uint32_t>(next_val[child[j]]));
        val[i] = static_cast<uint8_t>(m);
    }
}

// Marks the unmarked states whose value is below the cap; counts them.
__global__ __launch_bounds__(kBlock) void k_mx_mark(uint64_t S, const uint8_t* val, uint8_t* table,
                                                    unsigned long long* found) {
    unsigned long long f = 0, unused = 0;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < S;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint32_t v = val[i];
        if (!(table[i] & 1u) && v < kCap) {
            table[i] = won_byte(v);
            ++f;
        }
    }
    f = wave_sum(f);
    block_sum2(f, unused);
    if (threadIdx.x == 0 && f) atomicAdd(found, f);
}

// Marked-state plane for the verdict: bit r of row C = state (C, r) marked.
template <int W>
__global__ __launch_bounds__(kBlock) void k_mx_marked(int32_t n, uint32_t C, const uint8_t* table, uint64_t* marked) {
    const int lane = threadIdx.x & 63;
    const uint64_t nw = static_cast<uint64_t>(C) * W;
    for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * kWavesPerBlock + wave_id(); g < nw;
         g += static_cast<uint64_t>(gridDim.x) * kWavesPerBlock) {
        const uint32_t c = static_cast<uint32_t>(g / W);
        const int u = 64 * static_cast<int>(g - static_cast<uint64_t>(c) * W) + lane;
        const bool m = u < n && (table[static_cast<uint64_t>(c) * n + u] & 1u);
        const uint64_t b = __ballot(m);
        if (lane == 0) marked[g] = b;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <int K, typename F>
int dispatch_w_of(int w, F& f) {
    switch (w) {
        case 1: return f(std::integral_constant<int, K>{}, std::integral_constant<int, 1>{});
        case 2: return f(std::integral_constant<int, K>{}, std::integral_constant<int, 2>{});
        case 3: return f(std::integral_constant<int, K>{}, std::integral_constant<int, 3>{});
        case 4: return f(std::integral_constant<int, K>{}, std::integral_constant<int, 4>{});
        default: return fail(CR_ERR_ARG, "row width %d", w);
    }
}

template <typename F>
int dispatch_w(int w, F&& f) {
    switch (w) {
        case 1: return f(std::integral_constant<int, 1>{});
        case 2: return f(std::integral_constant<int, 2>{});
        case 3: return f(std::integral_constant<int, 3>{});
        case 4: return f(std::integral_constant<int, 4>{});
        default: return fail(CR_ERR_ARG, "row width %d", w);
    }
}

template <typename F>
int dispatch_kw(int k, int w, F&& f) {
    switch (k) {
        case 1: return dispatch_w_of<1>(w, f);
        case 2: return dispatch_w_of<2>(w, f);
        case 3: return dispatch_w_of<3>(w, f);
        case 4: return dispatch_w_of<4>(w, f);
        case 5: return dispatch_w_of<5>(w, f);
        case 6: return dispatch_w_of<6>(w, f);
        case 7: return dispatch_w_of<7>(w, f);
        case 8: return dispatch_w_of<8>(w, f);
        default: return fail(CR_ERR_K_RANGE, "k=%d outside 1..%d", k, CR_MAX_K);
    }
}

uint32_t blocks_for_threads(uint64_t threads, uint32_t cap = kMaxBlocks) {
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((threads + kBlock - 1) / kBlock, cap)));
}

uint32_t blocks_for_waves(uint64_t waves, uint32_t cap = 8192) {
    return static_cast<uint32_t>(
        std::max<uint64_t>(1, std::min<uint64_t>((waves + kWavesPerBlock - 1) / kWavesPerBlock, cap)));
}

// Device buffer that only grows, so repeated solves of one shape allocate once.
struct DevBuf {
    void* p = nullptr;
    uint64_t cap = 0;
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

template <typename T>
int ensure(DevBuf& b, uint64_t count, uint64_t& device_bytes) {
    const uint64_t bytes = std::max<uint64_t>(count, 1) * sizeof(T);
    if (b.cap >= bytes) return 0;
    if (b.p) {
        HIP_TRY(hipFree(b.p));
        device_bytes -= b.cap;
        b.p = nullptr;
        b.cap = 0;
    }
    size_t freeb = 0, totalb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    if (bytes + (256ull << 20) > freeb)
        return fail(CR_ERR_TOO_LARGE, "trajectory DAG needs %llu more bytes; %zu free",
                    static_cast<unsigned long long>(bytes), freeb);
    HIP_TRY(hipMalloc(&b.p, bytes));
    b.cap = bytes;
    device_bytes += bytes;
    return 0;
}

// Event pairs around launches, read after the stream is synchronised.
struct Spans {
    crd::TimeSink sink;
    hipStream_t st;
    std::vector<hipEvent_t>* pool;  // reused across solves
    struct Span {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<Span> open;
    hipEvent_t take() {
        hipEvent_t e = nullptr;
        if (pool->empty()) {
            hipEventCreate(&e);
        } else {
            e = pool->back();
            pool->pop_back();
        }
        return e;
    }
    void begin(const char* name) {
        if (!sink.add) return;
        Span s{name, take(), take()};
        hipEventRecord(s.a, st);
        open.push_back(s);
    }
    void end() {
        if (!sink.add) return;
        hipEventRecord(open.back().b, st);
    }
    void flush() {
        for (auto& s : open) {
            float ms = 0;
            hipEventElapsedTime(&ms, s.a, s.b);
            sink.add(sink.ctx, s.name, ms);
            pool->push_back(s.a);
            pool->push_back(s.b);
        }
        open.clear();
    }
};

}  // namespace

namespace crd {

struct MacroPlan {
    struct Step {
        DevBuf cfg, cloud, val, off, child;
        uint64_t n = 0, edges = 0;
    };
    std::vector<Step> steps;  // steps[d-1] = cop step d
    DevBuf cnt, cand_cfg, cand_cloud, rep, flag, slots, tiles, pos, planes, counters;
    unsigned long long* h_counters = nullptr;  // pinned: found, leaf moves, build moves, scratch
    std::vector<hipEvent_t> events;            // timing events, reused

    void release(uint64_t& device_bytes) {
        auto drop = [&](DevBuf& b) {
            if (b.p) {
                hipFree(b.p);
                device_bytes -= b.cap;
            }
            b = DevBuf{};
        };
        for (auto& s : steps) {
            drop(s.cfg);
            drop(s.cloud);
            drop(s.val);
            drop(s.off);
            drop(s.child);
        }
        steps.clear();
        for (DevBuf* b : {&cnt, &cand_cfg, &cand_cloud, &rep, &flag, &slots, &tiles, &pos, &planes, &counters}) drop(*b);
        if (h_counters) hipHostFree(h_counters);
        h_counters = nullptr;
        for (hipEvent_t e : events) hipEventDestroy(e);
        events.clear();
    }
};

void macro_free(MacroPlan* plan, uint64_t& device_bytes) {
    if (!plan) return;
    plan->release(device_bytes);
    delete plan;
}

int macro_init(const GraphView& gv, int k, hipStream_t st, uint8_t* table, uint64_t* occ, uint64_t* base) {
    unsigned long long* d_base = nullptr;
    HIP_TRY(hipMalloc(&d_base, sizeof(unsigned long long)));
    HIP_TRY(hipMemsetAsync(d_base, 0, sizeof(unsigned long long), st));
    const int rc = dispatch_kw(k, gv.W, [&](auto kc, auto wc) -> int {
        constexpr int K = decltype(kc)::value, W = decltype(wc)::value;
        if (int r = raise_lds_limit(reinterpret_cast<const void*>(k_mx_init<K, W>), gv.image_bytes)) return r;
        hipLaunchKernelGGL((k_mx_init<K, W>), dim3(blocks_for_waves(static_cast<uint64_t>(gv.C) * W, 2048)),
                           dim3(kBlock), gv.image_bytes, st, gv, table, occ, d_base);
        return 0;
    });
    unsigned long long h = 0;
    if (rc == 0) {
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&h, d_base, sizeof h, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    hipFree(d_base);
    if (base) *base = h;
    return rc;
}

namespace {

// exclusive scan of u32 counts into u64 positions; pos[n] = total
int scan_counts(MacroPlan& pl, const uint32_t* cnt, uint64_t n, unsigned long long* pos, hipStream_t st,
                uint64_t& device_bytes) {
    const uint64_t ntiles = (n + kScanTile - 1) / kScanTile;
    if (int rc = ensure<unsigned long long>(pl.tiles, ntiles, device_bytes)) return rc;
    auto* tiles = pl.tiles.as<unsigned long long>();
    if (ntiles > 0) {
        hipLaunchKernelGGL(k_tile_sums, dim3(static_cast<uint32_t>(ntiles)), dim3(kScanThreads), 0, st, cnt, n, tiles);
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, st, tiles, static_cast<uint32_t>(ntiles), pos, n);
        hipLaunchKernelGGL(k_tile_apply, dim3(static_cast<uint32_t>(ntiles)), dim3(kScanThreads), 0, st, cnt, n, tiles, pos);
    } else {
        HIP_TRY(hipMemsetAsync(pos, 0, sizeof(unsigned long long), st));
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

uint64_t read_u64(const unsigned long long* d, unsigned long long* h, hipStream_t st, int& rc) {
    if (hipMemcpyAsync(h, d, sizeof *h, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        rc = fail(CR_ERR_DEVICE, "device read failed");
        return 0;
    }
    return *h;
}

}  // namespace

int macro_solve(MacroPlan** planp, const GraphView& gv, const uint64_t* d_nbr_mask, int k, int p, hipStream_t st,
                uint8_t* table, const uint64_t* occ, uint64_t* marked, uint64_t& device_bytes, TimeSink sink,
                std::vector<uint64_t>& passes, MacroProfile& prof) {
    if (!*planp) *planp = new MacroPlan();
    MacroPlan& pl = **planp;
    const int W = gv.W, n = gv.n;
    const uint64_t C = gv.C, S = C * static_cast<uint64_t>(n);
    const uint32_t lds_expand = nbr_offset(gv.image_bytes) + static_cast<uint32_t>(n * W * 8);
    if (lds_expand > 160 * 1024) return fail(CR_ERR_UNSUPPORTED, "graph image and robber masks exceed LDS");
    if (p > 127) return fail(CR_ERR_P_RANGE, "p=%d: every trajectory longer than 63 cop steps passes the 127 cap", p);
    int rc = 0;
    if (!pl.h_counters)
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&pl.h_counters), 4 * sizeof(unsigned long long), hipHostMallocDefault));
    if ((rc = ensure<unsigned long long>(pl.counters, 4, device_bytes))) return rc;
    auto* ctr = pl.counters.as<unsigned long long>();
    HIP_TRY(hipMemsetAsync(ctr, 0, 4 * sizeof(unsigned long long), st));
    pl.steps.resize(p);
    prof = MacroProfile{};
    prof.nodes.assign(p, 0);
    prof.candidates.assign(p > 1 ? p - 1 : 0, 0);
    prof.edges.assign(p > 1 ? p - 1 : 0, 0);
    passes.clear();
    Spans spans{sink, st, &pl.events, {}};

    hipEvent_t e0, e1, e2;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventCreate(&e2));
    HIP_TRY(hipEventRecord(e0, st));

    // ---- forward: the DAG ---------------------------------------------------
    pl.steps[0].n = S;
    for (int d = 1; d < p; ++d) pl.steps[d].n = 0;
    for (auto& s : pl.steps) s.edges = 0;
    if ((rc = ensure<uint8_t>(pl.steps[0].val, S, device_bytes))) return rc;
    for (int d = 1; d < p; ++d) {
        MacroPlan::Step& cur = pl.steps[d - 1];
        MacroPlan::Step& nxt = pl.steps[d];
        if (cur.n == 0) break;
        const Nodes nodes{cur.n, d == 1 ? nullptr : cur.cfg.as<uint32_t>(), d == 1 ? nullptr : cur.cloud.as<uint64_t>()};
        if ((rc = ensure<uint32_t>(pl.cnt, cur.n, device_bytes))) return rc;
        if ((rc = ensure<unsigned long long>(cur.off, cur.n + 1, device_bytes))) return rc;
        if (d > 1 && (rc = ensure<uint8_t>(cur.val, cur.n, device_bytes))) return rc;
        auto* off = cur.off.as<unsigned long long>();
        spans.begin("mx_expand");
        rc = dispatch_kw(k, W, [&](auto kc, auto wc) -> int {
            constexpr int K = decltype(kc)::value, WW = decltype(wc)::value;
            if (int r = raise_lds_limit(reinterpret_cast<const void*>(k_mx_expand<K, WW>), lds_expand)) return r;
            hipLaunchKernelGGL((k_mx_expand<K, WW>), dim3(blocks_for_threads(cur.n)), dim3(kBlock), lds_expand, st, gv,
                               d_nbr_mask, nodes, static_cast<uint32_t>(d), pl.cnt.as<uint32_t>(),
                               cur.val.as<uint8_t>(), static_cast<const unsigned long long*>(nullptr),
                               static_cast<uint32_t*>(nullptr), static_cast<uint64_t*>(nullptr), ctr + 2);
            return 0;
        });
        spans.end();
        if (rc) return rc;
        if ((rc = scan_counts(pl, pl.cnt.as<uint32_t>(), cur.n, off, st, device_bytes))) return rc;
        const uint64_t E = read_u64(off + cur.n, pl.h_counters, st, rc);
        if (rc) return rc;
        spans.flush();
        prof.candidates[d - 1] = E;
        if (E == 0) break;
        if (E >= kEmptySlot) return fail(CR_ERR_TOO_LARGE, "%llu trajectory candidates at step %d exceed 32-bit ids",
                                         static_cast<unsigned long long>(E), d + 1);
        uint64_t slots = 1;
        while (slots < 2 * E) slots <<= 1;
        if ((rc = ensure<uint32_t>(pl.cand_cfg, E, device_bytes)) || (rc = ensure<uint64_t>(pl.cand_cloud, E * W, device_bytes)) ||
            (rc = ensure<uint32_t>(pl.rep, E, device_bytes)) || (rc = ensure<uint32_t>(pl.flag, E, device_bytes)) ||
            (rc = ensure<unsigned long long>(pl.pos, E + 1, device_bytes)) || (rc = ensure<uint32_t>(pl.slots, slots, device_bytes)) ||
            (rc = ensure<uint32_t>(cur.child, E, device_bytes)))
            return rc;
        spans.begin("mx_emit");
        rc = dispatch_kw(k, W, [&](auto kc, auto wc) -> int {
            constexpr int K = decltype(kc)::value, WW = decltype(wc)::value;
            hipLaunchKernelGGL((k_mx_expand<K, WW>), dim3(blocks_for_threads(cur.n)), dim3(kBlock), lds_expand, st, gv,
                               d_nbr_mask, nodes, static_cast<uint32_t>(d), pl.cnt.as<uint32_t>(),
                               cur.val.as<uint8_t>(), static_cast<const unsigned long long*>(off),
                               pl.cand_cfg.as<uint32_t>(), pl.cand_cloud.as<uint64_t>(), ctr + 2);
            return 0;
        });
        spans.end();
        if (rc) return rc;
        const uint32_t e32 = static_cast<uint32_t>(E);
        HIP_TRY(hipMemsetAsync(pl.slots.p, 0xff, slots * sizeof(uint32_t), st));
        spans.begin("mx_insert");
        rc = dispatch_w(W, [&](auto wc) -> int {
            constexpr int WW = decltype(wc)::value;
            hipLaunchKernelGGL((k_mx_insert<WW>), dim3(blocks_for_threads(E, 8192)), dim3(kBlock), 0, st,
                               pl.cand_cfg.as<uint32_t>(), pl.cand_cloud.as<uint64_t>(), e32, pl.slots.as<uint32_t>(),
                               slots - 1, pl.rep.as<uint32_t>());
            return 0;
        });
        spans.end();
        if (rc) return rc;
        hipLaunchKernelGGL(k_mx_flag, dim3(blocks_for_threads(E, 8192)), dim3(kBlock), 0, st, pl.rep.as<uint32_t>(), e32,
                           pl.flag.as<uint32_t>());
        if ((rc = scan_counts(pl, pl.flag.as<uint32_t>(), E, pl.pos.as<unsigned long long>(), st, device_bytes))) return rc;
        const uint64_t U = read_u64(pl.pos.as<unsigned long long>() + E, pl.h_counters, st, rc);
        if (rc) return rc;
        if ((rc = ensure<uint32_t>(nxt.cfg, U, device_bytes)) || (rc = ensure<uint64_t>(nxt.cloud, U * W, device_bytes)) ||
            (rc = ensure<uint8_t>(nxt.val, U, device_bytes)))
            return rc;
        spans.begin("mx_compact");
        rc = dispatch_w(W, [&](auto wc) -> int {
            constexpr int WW = decltype(wc)::value;
            hipLaunchKernelGGL((k_mx_compact<WW>), dim3(blocks_for_threads(E, 8192)), dim3(kBlock), 0, st,
                               pl.cand_cfg.as<uint32_t>(), pl.cand_cloud.as<uint64_t>(), pl.rep.as<uint32_t>(),
                               pl.pos.as<unsigned long long>(), e32, nxt.cfg.as<uint32_t>(), nxt.cloud.as<uint64_t>(),
                               cur.child.as<uint32_t>());
            return 0;
        });
        spans.end();
        if (rc) return rc;
        HIP_TRY(hipGetLastError());
        cur.edges = E;
        nxt.n = U;
        prof.edges[d - 1] = E;
    }
    for (int d = 0; d < p; ++d) prof.nodes[d] = pl.steps[d].n;
    HIP_TRY(hipEventRecord(e1, st));

    // ---- passes -------------------------------------------------------------
    if ((rc = ensure<uint64_t>(pl.planes, C * kPlanes * W, device_bytes))) return rc;
    const MacroPlan::Step& leaf = pl.steps[p - 1];
    const Nodes leaf_nodes{leaf.n, p == 1 ? nullptr : leaf.cfg.as<uint32_t>(), p == 1 ? nullptr : leaf.cloud.as<uint64_t>()};
    uint64_t up_edges = 0;
    for (int d = 1; d < p; ++d) up_edges += pl.steps[d - 1].edges;
    for (int pass = 1;; ++pass) {
        if (pass > (1 << 20)) return fail(CR_ERR_DEVICE, "pass loop did not settle");
        HIP_TRY(hipMemsetAsync(ctr, 0, sizeof(unsigned long long), st));
        spans.begin("mx_planes");
        rc = dispatch_w(W, [&](auto wc) -> int {
            constexpr int WW = decltype(wc)::value;
            hipLaunchKernelGGL((k_mx_planes<WW>), dim3(blocks_for_waves(C * WW)), dim3(kBlock), 0, st, n,
                               static_cast<uint32_t>(C), d_nbr_mask, static_cast<const uint8_t*>(table), occ,
                               pl.planes.as<uint64_t>());
            return 0;
        });
        spans.end();
        if (rc) return rc;
        if (leaf.n) {
            spans.begin("mx_leaf");
            rc = dispatch_kw(k, W, [&](auto kc, auto wc) -> int {
                constexpr int K = decltype(kc)::value, WW = decltype(wc)::value;
                if (int r = raise_lds_limit(reinterpret_cast<const void*>(k_mx_leaf<K, WW>), gv.image_bytes)) return r;
                hipLaunchKernelGGL((k_mx_leaf<K, WW>), dim3(blocks_for_threads(leaf.n)), dim3(kBlock), gv.image_bytes, st,
                                   gv, leaf_nodes, static_cast<const uint64_t*>(pl.planes.as<uint64_t>()),
                                   static_cast<uint32_t>(p), pl.steps[p - 1].val.as<uint8_t>(),
                                   static_cast<const uint8_t*>(table), ctr + 1);
                return 0;
            });
            spans.end();
            if (rc) return rc;
        }
        for (int d = p - 1; d >= 1; --d) {
            const MacroPlan::Step& s = pl.steps[d - 1];
            if (s.edges == 0) continue;
            spans.begin("mx_up");
            hipLaunchKernelGGL(k_mx_up, dim3(blocks_for_threads(s.n, 8192)), dim3(kBlock), 0, st, s.n,
                               static_cast<const unsigned long long*>(s.off.as<unsigned long long>()),
                               static_cast<const uint32_t*>(s.child.as<uint32_t>()),
                               static_cast<const uint8_t*>(pl.steps[d].val.as<uint8_t>()), s.val.as<uint8_t>());
            spans.end();
        }
        spans.begin("mx_mark");
        hipLaunchKernelGGL(k_mx_mark, dim3(blocks_for_threads(S, 4096)), dim3(kBlock), 0, st, S,
                           static_cast<const uint8_t*>(pl.steps[0].val.as<uint8_t>()), table, ctr);
        spans.end();
        HIP_TRY(hipGetLastError());
        const uint64_t found = read_u64(ctr, pl.h_counters, st, rc);
        if (rc) return rc;
        spans.flush();
        passes.push_back(found);
        prof.up_edges += up_edges;
        if (found == 0) break;
    }
    rc = dispatch_w(W, [&](auto wc) -> int {
        constexpr int WW = decltype(wc)::value;
        hipLaunchKernelGGL((k_mx_marked<WW>), dim3(blocks_for_waves(C * WW)), dim3(kBlock), 0, st, n,
                           static_cast<uint32_t>(C), static_cast<const uint8_t*>(table), marked);
        return 0;
    });
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e2, st));
    HIP_TRY(hipMemcpyAsync(pl.h_counters, ctr, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    prof.leaf_moves = pl.h_counters[1];
    prof.build_moves = pl.h_counters[2];
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    prof.build_ms = ms;
    hipEventElapsedTime(&ms, e1, e2);
    prof.passes_ms = ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(e2);
    return 0;
}

}  // namespace crd
