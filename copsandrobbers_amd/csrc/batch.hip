// batch.hip -- many small plain games (p = 1) in one persistent launch, and a
// continuous-batching server on top of it.
//
// Each workgroup takes graphs from a device work queue (largest first) and
// solves one graph at a time with the whole workgroup, so a graph's depth
// levels need only workgroup barriers and graphs never wait on each other.
// Per graph (n <= 64 vertices: one 64-bit robber mask per configuration row),
// in the workgroup's scratch slot:
//   won0 / won1   states won with the cops / the robber to move
//   fr1           robber-turn states of the current depth
//   nx0 / nx1     new cop-turn / robber-turn states of the next depth
//   H0 / H1       intermediates of the factorised cop-turn pull
// Level d (states of depth d -> depth d+1; the level-synchronous form of
// k_cops_visibility.cpp:129-299, as in the single-graph solver):
//   cop turn     nx0[C] = OR_{C' in T(C)} fr1[C'] & ~won0[C], taken one cop
//                at a time: H_i[A][B] = OR_{v in N(max A)} H_{i-1}[A - max A][B + v]
//                (A: i cops still at their old vertices, B: K-i moved cops),
//                H_0 = fr1, H_K = nx0;
//   robber turn  nx1[C] = { r not won : every move of r lands in won0[C] };
//   finalize     table bytes at depth d+1, won |= new, fr1 = nx1.
// Levels start at d = 0 from the captures and stop at the first level that
// wins nothing or before depth 127 (the 7-bit cap, :189, :280).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <deque>
#include <mutex>
#include <numeric>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cr_device.h"
#include "cr_host.h"

namespace {

constexpr int kMaxBK = CR_BATCH_MAX_K;
constexpr int kMaxBN = CR_BATCH_MAX_N;
constexpr int kBinomCols = kMaxBN + kMaxBK;  // x in [0, n + K - 1]

struct BatchGraph {
    uint32_t n, k, C, hmax;   // hmax: largest intermediate H_i (rows)
    uint64_t mask_off;        // n out-masks at masks + mask_off
    uint64_t table_off;       // 2*C*n table bytes at tables + table_off
};

struct BatchOut {
    unsigned long long verdict;  // worst << 32 | cId of the best universal row, ~0 if none
    unsigned long long base, won, hash;
    uint32_t levels, pad;
    uint32_t hist[CR_MAX_ROUND][2];  // states won per depth and column
};

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

constexpr int kMaxBlockWaves = 16;
constexpr int32_t kPushPct = 100;   // default push threshold (% of the pull's cost)  // workgroups of up to 1024 threads
constexpr int kMlp = 8;             // row loads kept in flight per thread in the pull

// Block-wide sums with the result in every thread.
__device__ __forceinline__ void block_sum_all(unsigned long long& x, unsigned long long& y,
                                              unsigned long long* red /* [2][kMaxBlockWaves] */) {
    const int nw = blockDim.x >> 6;
    x = wave_sum(x);
    y = wave_sum(y);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = x;
        red[kMaxBlockWaves + (threadIdx.x >> 6)] = y;
    }
    __syncthreads();
    x = y = 0;
    for (int i = 0; i < nw; ++i) {
        x += red[i];
        y += red[kMaxBlockWaves + i];
    }
}

__device__ __forceinline__ unsigned long long block_min_all(unsigned long long x, unsigned long long* red) {
    const int nw = blockDim.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(x, off);
        x = o < x ? o : x;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    x = ~0ull;
    for (int i = 0; i < nw; ++i) x = red[i] < x ? red[i] : x;
    return x;
}

// C(x, y) for the small arguments of one graph (exact: every prefix product
// divides).
__device__ __forceinline__ uint32_t small_binom(int x, int y) {
    if (y < 0 || x < y) return 0;
    uint64_t c = 1;
    for (int i = 0; i < y; ++i) c = c * static_cast<uint64_t>(x - i) / static_cast<uint64_t>(i + 1);
    return static_cast<uint32_t>(c);
}

// Multisets of one or two vertices: rank <-> values without the binomial
// search (lexicographic order over [0, n); pairs from a per-graph table).
constexpr int kPairTab = (kMaxBN * (kMaxBN + 1)) / 2;
__device__ __forceinline__ uint32_t rank_pair_n(uint32_t a, uint32_t b, uint32_t n) {
    return a * n - (a * (a - 1)) / 2 + (b - a);
}
// Multisets of three vertices: s_t3[a] is the rank of (a, a, a), i.e. the
// triples with a smaller first element, C(n+2, 3) - C(n-a+2, 3). The triples
// (a, b, c) with first element a, in order, are the pairs (b, c) over [a, n),
// and those are the suffix of the pair order from (a, a): a triple costs a
// search over the n starts and one pair-table read.
template <int J>
__device__ __forceinline__ void unrank_small(uint32_t r, int n, const uint32_t* s_binom, int Mimg,
                                             const uint16_t* s_pair, const uint32_t* s_t3, uint32_t* out) {
    if constexpr (J == 1) {
        out[0] = r;
    } else if constexpr (J == 2) {
        const uint32_t p = s_pair[r];
        out[0] = p & 0xffu;
        out[1] = p >> 8;
    } else if constexpr (J == 3) {
        int lo = 0, hi = n - 1;  // the last a with s_t3[a] <= r
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_t3[mid] <= r) lo = mid; else hi = mid - 1;
        }
        const uint32_t a = static_cast<uint32_t>(lo);
        const uint32_t p = s_pair[rank_pair_n(a, a, static_cast<uint32_t>(n)) + (r - s_t3[a])];
        out[0] = a;
        out[1] = p & 0xffu;
        out[2] = p >> 8;
    } else if constexpr (J > 3) {
        unrank_j(r, J, n, s_binom, Mimg, out);
    }
}
template <int J>
__device__ __forceinline__ uint32_t rank_small(const uint32_t* t, int n, const uint32_t* s_binom, int Mimg) {
    if constexpr (J == 0) {
        return 0;
    } else if constexpr (J == 1) {
        return t[0];
    } else if constexpr (J == 2) {
        return rank_pair_n(t[0], t[1], static_cast<uint32_t>(n));
    } else {
        return rank_j(t, J, n, s_binom, Mimg);
    }
}

// Stage I of a graph's factorised pull: H_I[A][B] = OR_{v in N(max A)}
// H_{I-1}[A - max A][B + v] (H_0 = fr1; stage K writes the candidates masked
// with ~won0), rows o = rank(A) * cnt(K - I) + rank(B), lanes over rows.
template <int K, int I>
__device__ __forceinline__ void pull_stage(const uint64_t* __restrict__ in, uint64_t* __restrict__ dst, const uint32_t* cnt,
                                           int n, int Mimg, const uint32_t* s_binom, const uint16_t* s_pair,
                                           const uint32_t* s_t3, const uint64_t* s_mask,
                                           const uint64_t* __restrict__ won0, uint64_t full) {
    const uint32_t T = blockDim.x, tid = threadIdx.x;
    const uint32_t cb = cnt[K - I], cb_in = cnt[K - I + 1], rows = cnt[I] * cb;
    for (uint32_t o = tid; o < rows; o += T) {
        const uint32_t ra = o / cb, rb = o - ra * cb;
        uint32_t A[K], B[K + 1];
        unrank_small<I>(ra, n, s_binom, Mimg, s_pair, s_t3, A);
        unrank_small<K - I>(rb, n, s_binom, Mimg, s_pair, s_t3, B);
        const uint64_t row_in = static_cast<uint64_t>(rank_small<I - 1>(A, n, s_binom, Mimg)) * cb_in;
        // rank terms of B for its (J+1)-multisets B + v (J = K - I): with T(p, x)
        // the term of value x at position p, rank(B + v) = rank0 - sum_{B_t < v}
        // T(t, B_t) - T(q, v) - sum_{B_t >= v} T(t + 1, B_t), q = #{B_t < v}
        constexpr int J = K - I, J1 = J + 1;
        const int Mp = Mimg + 1, M1 = n + J;
        uint32_t U[J > 0 ? J : 1], S[J > 0 ? J : 1];
        uint32_t rank0 = 0;
        if constexpr (J1 >= 3) {
            rank0 = s_binom[J1 * Mp + M1] - 1;
#pragma unroll
            for (int t = 0; t < J; ++t) {
                U[t] = s_binom[(J1 - t) * Mp + (M1 - 1 - static_cast<int>(B[t]) - t)];
                S[t] = s_binom[(J - t) * Mp + (M1 - 2 - static_cast<int>(B[t]) - t)];
            }
        }
        uint64_t acc = 0, bits = s_mask[A[I - 1]];
        while (bits) {
            // ranks of up to kMlp successor rows first, then their loads together
            uint32_t idx[kMlp];
            int m = 0;
#pragma unroll
            for (int u = 0; u < kMlp; ++u) {
                if (bits == 0) break;
                const uint32_t v = static_cast<uint32_t>(__builtin_ctzll(bits));
                bits &= bits - 1;
                if constexpr (J1 >= 3) {
                    uint32_t sum = 0;
                    int q = 0;
#pragma unroll
                    for (int t = 0; t < J; ++t) {
                        const bool lt = B[t] < v;
                        sum += lt ? U[t] : S[t];
                        q += lt ? 1 : 0;
                    }
                    idx[u] = rank0 - sum - s_binom[(J1 - q) * Mp + (M1 - 1 - static_cast<int>(v) - q)];
                } else {
                    uint32_t M2[K + 1];
                    int pos = 0;
                    bool placed = false;
#pragma unroll
                    for (int t = 0; t < K - I; ++t) {
                        if (!placed && v <= B[t]) {
                            M2[pos++] = v;
                            placed = true;
                        }
                        M2[pos++] = B[t];
                    }
                    if (!placed) M2[pos++] = v;
                    idx[u] = rank_small<K - I + 1>(M2, n, s_binom, Mimg);
                }
                m = u + 1;
            }
            uint64_t got[kMlp];
#pragma unroll
            for (int u = 0; u < kMlp; ++u) got[u] = u < m ? in[row_in + idx[u]] : 0ull;
#pragma unroll
            for (int u = 0; u < kMlp; ++u) acc |= got[u];
        }
        dst[o] = I == K ? (acc & ~won0[o] & full) : acc;
    }
    __syncthreads();
}

template <int K>
__device__ void solve_graph(const BatchGraph& g, const uint64_t* __restrict__ masks, uint8_t* __restrict__ tables,
                            BatchOut* __restrict__ out, uint64_t* __restrict__ scratch, uint64_t* s_mask,
                            uint32_t* s_binom, unsigned long long* s_red, uint16_t* s_pair, uint32_t* s_t3,
                            uint64_t* s_occ, uint64_t* s_nib, uint8_t* s_inl, uint32_t push_pct) {
    const int n = static_cast<int>(g.n), Mimg = n + K - 1;
    const uint32_t C = g.C, T = blockDim.x, tid = threadIdx.x;
    const uint64_t full = n == 64 ? ~0ull : ((1ull << n) - 1);
    for (int i = tid; i < n; i += T) s_mask[i] = masks[g.mask_off + i];
    __syncthreads();  // s_mask, for the robber table
    // robber-step table: s_nib[p * 16 + x] = the robbers with a move onto one
    // of the vertices 4p + j for the set bits j of x
    const int nnib = (n + 3) / 4;
    for (int e = tid; e < nnib * 16; e += T) {
        const uint64_t vs = static_cast<uint64_t>(e & 15) << (4 * (e >> 4));
        uint64_t acc = 0;
        for (int r = 0; r < n; ++r)
            if (s_mask[r] & vs) acc |= 1ull << r;
        s_nib[e] = acc;
    }
    for (int i = tid; i < (K + 1) * (Mimg + 1); i += T) s_binom[i] = small_binom(i % (Mimg + 1), i / (Mimg + 1));
    // the pair table: row a of the lexicographic order holds (a, a..n-1)
    for (int a = 0; a < n; ++a) {
        const uint32_t base = rank_pair_n(a, a, n);
        for (int b = a + static_cast<int>(tid); b < n; b += T) s_pair[base + (b - a)] = static_cast<uint16_t>(a | (b << 8));
    }
    __syncthreads();
    // push: row v of s_inl holds v's in-degree, then the vertices with a move
    // onto v (the single-vertex entries of the nibble table)
    for (int v = tid; v < n; v += T) {
        uint8_t* row = s_inl + v * (kMaxBN + 1);
        uint32_t deg = 0;
        for (uint64_t m = s_nib[(v >> 2) * 16 + (1 << (v & 3))]; m; m &= m - 1) row[1 + deg++] = static_cast<uint8_t>(__builtin_ctzll(m));
        row[0] = static_cast<uint8_t>(deg);
    }
    uint32_t cnt[K + 1];
#pragma unroll
    for (int j = 0; j <= K; ++j) cnt[j] = s_binom[j * (Mimg + 1) + (n + j - 1)];  // C(n+j-1, j)
    // the starts of the triples by first element (K >= 3)
    const uint32_t* trip = s_t3;
    if constexpr (K >= 3) {
        for (int a = tid; a < n; a += T) s_t3[a] = cnt[3] - small_binom(n - a + 2, 3);
        __syncthreads();
    }
    uint64_t* won0 = scratch;
    uint64_t* won1 = won0 + C;
    uint64_t* fr1 = won1 + C;
    uint64_t* nx0 = fr1 + C;
    uint64_t* nx1 = nx0 + C;
    uint64_t* hbuf[2] = {nx1 + C, nx1 + C + g.hmax};
    uint8_t* last0 = reinterpret_cast<uint8_t*>(nx1 + C + 2ull * g.hmax);
    uint8_t* table = tables + g.table_off;

    // captures: depth 0 in both columns (k_cops_visibility.cpp:85-127). The
    // rows of a wave's 64 configurations are one contiguous span of the table
    // (64 * 2n bytes from a 16-byte aligned start where the table is): lanes
    // write consecutive 16-byte chunks of it, reading each row's cop mask
    // from the wave's slot in s_occ. Otherwise each lane writes its row.
    unsigned long long base = 0, unused = 0;
    const bool spans = (reinterpret_cast<uintptr_t>(table) & 15u) == 0;  // uniform
    const uint32_t rowb = 2u * static_cast<uint32_t>(n), lane = tid & 63u;
    uint64_t* occ_w = s_occ + (tid >> 6) * 64;
    for (uint32_t c0 = 0; c0 < C; c0 += T) {
        const uint32_t c = c0 + tid;
        uint64_t occ = 0;
        if (c < C) {
            uint32_t cops[K];
            unrank_small<K>(c, n, s_binom, Mimg, s_pair, trip, cops);
#pragma unroll
            for (int j = 0; j < K; ++j) occ |= 1ull << cops[j];
            won0[c] = occ;
            won1[c] = occ;
            fr1[c] = occ;
            last0[c] = 0;
            base += __builtin_popcountll(occ);
            if (!spans) {
                uint8_t* row = table + static_cast<uint64_t>(c) * rowb;
                for (int r = 0; r < n; ++r) {
                    const uint8_t b = ((occ >> r) & 1) ? won_byte(0) : kUnmarked;
                    row[2 * r] = b;
                    row[2 * r + 1] = b;
                }
            }
        }
        if (spans) {
            occ_w[lane] = occ;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const uint32_t r0 = c0 + (tid & ~63u);  // the wave's first configuration
            const uint32_t rows = r0 < C ? min(64u, C - r0) : 0u;
            const uint32_t span = rows * rowb, full16 = span / 16;
            uint8_t* dst = table + static_cast<uint64_t>(r0) * rowb;
            for (uint32_t q = lane; q < full16; q += 64) {
                uint32_t o = 16 * q, rl = o / rowb;
                o -= rl * rowb;
                uint64_t m = occ_w[rl];
                uint32_t v[4] = {0, 0, 0, 0};
#pragma unroll
                for (int t = 0; t < 16; ++t) {
                    const uint32_t b = ((m >> (o >> 1)) & 1) ? won_byte(0) : kUnmarked;
                    v[t >> 2] |= b << (8 * (t & 3));
                    if (++o == rowb) {
                        o = 0;
                        ++rl;
                        m = rl < 64 ? occ_w[rl] : 0ull;
                    }
                }
                reinterpret_cast<uint4*>(dst)[q] = make_uint4(v[0], v[1], v[2], v[3]);
            }
            for (uint32_t i = 16 * full16 + lane; i < span; i += 64) {  // the last wave's tail bytes
                const uint32_t rl = i / rowb, r = (i - rl * rowb) >> 1;
                dst[i] = ((occ_w[rl] >> r) & 1) ? won_byte(0) : kUnmarked;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    block_sum_all(base, unused, s_red);
    if (tid == 0) {
        out->hist[0][0] = static_cast<uint32_t>(base);
        out->hist[0][1] = static_cast<uint32_t>(base);
        for (int d = 1; d < CR_MAX_ROUND; ++d) out->hist[d][0] = out->hist[d][1] = 0;
    }
    // push and pull costs per level (row loads and relaxations)
    uint32_t maxdeg = 0, edges = 0;
    for (int v = 0; v < n; ++v) {
        maxdeg = max(maxdeg, static_cast<uint32_t>(s_inl[v * (kMaxBN + 1)]));
        edges += static_cast<uint32_t>(__builtin_popcountll(s_mask[v]));
    }
    unsigned long long push_unit = 1, pull_rows = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) push_unit *= maxdeg;
#pragma unroll
    for (int i = 1; i <= K; ++i) pull_rows += static_cast<unsigned long long>(cnt[i]) * cnt[K - i];
    const unsigned long long pull_cost = pull_rows * edges / static_cast<unsigned long long>(n);
    unsigned long long won = 2 * base, prev0 = base, prev1 = base;
    uint32_t levels = 0;
    for (uint32_t d = 0; d + 1 < CR_MAX_ROUND; ++d) {
        const bool do_cop = prev1 != 0, do_rob = prev0 != 0;
        // cop turn: the factorised pull from the robber frontier, or a push from
        // it where the frontier is small: at most prev1 frontier rows times
        // maxdeg^K predecessors against the pull's rows times the mean degree
        const bool push = do_cop && push_pct != 0 && prev1 * push_unit * 100 <= static_cast<unsigned long long>(push_pct) * pull_cost;
        if (do_cop && push) {
            for (uint32_t c = tid; c < C; c += T) nx0[c] = 0;
            __syncthreads();
            // waves take 64 configurations at a time; the lanes enumerate the
            // predecessors of each frontier row among them (duplicates only OR
            // the same bits again)
            for (uint32_t cb = (tid & ~63u); cb < C; cb += T) {
                const uint32_t c = cb + lane;
                const uint64_t f = c < C ? fr1[c] : 0ull;
                for (uint64_t live = __ballot(f != 0); live; live &= live - 1) {
                    const int j = __builtin_ctzll(live);
                    const uint64_t bits = __shfl(f, j);
                    uint32_t cops[K], deg[K];
                    unrank_small<K>(cb + static_cast<uint32_t>(j), n, s_binom, Mimg, s_pair, trip, cops);
                    uint32_t P = 1;
#pragma unroll
                    for (int i = 0; i < K; ++i) {
                        deg[i] = s_inl[cops[i] * (kMaxBN + 1)];
                        P *= deg[i];
                    }
                    for (uint32_t t = lane; t < P; t += 64) {
                        uint32_t x[K], q = t;
#pragma unroll
                        for (int i = K - 1; i >= 0; --i) {
                            const uint32_t qi = q / deg[i];
                            x[i] = s_inl[cops[i] * (kMaxBN + 1) + 1 + (q - qi * deg[i])];
                            q = qi;
                        }
#pragma unroll
                        for (int i = 1; i < K; ++i)  // insertion sort
#pragma unroll
                            for (int m = i; m > 0; --m)
                                if (x[m - 1] > x[m]) {
                                    const uint32_t tmp = x[m - 1];
                                    x[m - 1] = x[m];
                                    x[m] = tmp;
                                }
                        atomicOr(reinterpret_cast<unsigned long long*>(nx0 + rank_small<K>(x, n, s_binom, Mimg)),
                                 static_cast<unsigned long long>(bits));
                    }
                }
            }
            __syncthreads();
        } else if (do_cop) {
            pull_stage<K, 1>(fr1, K == 1 ? nx0 : hbuf[0], cnt, n, Mimg, s_binom, s_pair, trip, s_mask, won0, full);
            if constexpr (K >= 2) pull_stage<K, 2>(hbuf[0], K == 2 ? nx0 : hbuf[1], cnt, n, Mimg, s_binom, s_pair, trip, s_mask, won0, full);
            if constexpr (K >= 3) pull_stage<K, 3>(hbuf[1], K == 3 ? nx0 : hbuf[0], cnt, n, Mimg, s_binom, s_pair, trip, s_mask, won0, full);
            if constexpr (K >= 4) pull_stage<K, 4>(hbuf[0], nx0, cnt, n, Mimg, s_binom, s_pair, trip, s_mask, won0, full);
        }
        // robber turn: every move of r lands on a won cop-turn state (depth <= d).
        // Each robber not won yet is tested against its moves; where a lane of
        // the wave has more of them than nibbles, the wave takes the robbers
        // with a move onto an open cop-turn state from the table instead (a
        // lookup per nibble of the open set).
        if (do_rob) {
            uint64_t w0_nx = tid < C ? won0[tid] : 0, w1_nx = tid < C ? won1[tid] : 0;  // loaded a row ahead
            for (uint32_t c = tid; c < C; c += T) {
                const uint64_t w0 = w0_nx, w1 = w1_nx;
                if (c + T < C) {
                    w0_nx = won0[c + T];
                    w1_nx = won1[c + T];
                }
                uint64_t cand = full & ~w1, fresh = 0;
                if (__ballot(__builtin_popcountll(cand) > nnib) != 0) {
                    const uint64_t open = full & ~w0;
                    uint64_t stay = 0;
                    for (int p = 0; p < nnib; ++p) stay |= s_nib[p * 16 + ((open >> (4 * p)) & 15u)];
                    fresh = cand & ~stay;
                } else {
                    while (cand) {
                        const int r = __builtin_ctzll(cand);
                        cand &= cand - 1;
                        if ((s_mask[r] & ~w0) == 0) fresh |= 1ull << r;
                    }
                }
                nx1[c] = fresh;
            }
        }
        __syncthreads();
        // finalize: depth d+1
        unsigned long long c0 = 0, c1 = 0;
        const uint8_t byte = won_byte(d + 1);
        // the next configuration's rows are loaded while this one's bytes are stored
        uint64_t a_nx = 0, b_nx = 0, w0_nx = 0, w1_nx = 0;
        if (tid < C) {
            a_nx = do_cop ? nx0[tid] : 0;
            b_nx = do_rob ? nx1[tid] : 0;
            w0_nx = do_cop ? won0[tid] : 0;
            w1_nx = do_rob ? won1[tid] : 0;
        }
        for (uint32_t c = tid; c < C; c += T) {
            uint64_t a = a_nx;
            const uint64_t b = b_nx, w0 = w0_nx, w1 = w1_nx;
            if (c + T < C) {
                a_nx = do_cop ? nx0[c + T] : 0;
                b_nx = do_rob ? nx1[c + T] : 0;
                w0_nx = do_cop ? won0[c + T] : 0;
                w1_nx = do_rob ? won1[c + T] : 0;
            }
            uint8_t* row = table + static_cast<uint64_t>(c) * n * 2;
            a &= ~w0 & full;  // the push's candidates may include won states (the pull's are masked)
            if (a) {
                won0[c] = w0 | a;
                last0[c] = static_cast<uint8_t>(d + 1);
                uint64_t m = a;
                while (m) {
                    row[2 * __builtin_ctzll(m)] = byte;
                    m &= m - 1;
                }
                c0 += __builtin_popcountll(a);
            }
            if (b) {
                won1[c] = w1 | b;
                uint64_t m = b;
                while (m) {
                    row[2 * __builtin_ctzll(m) + 1] = byte;
                    m &= m - 1;
                }
                c1 += __builtin_popcountll(b);
            }
            fr1[c] = b;
        }
        block_sum_all(c0, c1, s_red);
        if (tid == 0) {
            out->hist[d + 1][0] = static_cast<uint32_t>(c0);
            out->hist[d + 1][1] = static_cast<uint32_t>(c1);
        }
        won += c0 + c1;
        prev0 = c0;
        prev1 = c1;
        levels = d + 1;
        if (c0 + c1 == 0) break;
    }
    // findFinalResult (k_cops_visibility.cpp:301-354): universal rows, least
    // worst depth, lowest cId
    unsigned long long best = ~0ull;
    for (uint32_t c = tid; c < C; c += T)
        if (won0[c] == full) {
            const unsigned long long key = (static_cast<unsigned long long>(last0[c]) << 32) | c;
            best = key < best ? key : best;
        }
    best = block_min_all(best, s_red);
    // order-independent table hash, sum over bytes i of (b_i + 1) * mix64(i):
    // 16 bytes per lane and step where the table is 16-byte aligned (tables
    // start at 256-byte offsets of the table memory), then the tail bytes
    unsigned long long h = 0;
    const uint64_t bytes = 2ull * C * n;
    const uint64_t chunks = (reinterpret_cast<uintptr_t>(table) & 15u) == 0 ? bytes / 16 : 0;
    const uint4* table16 = reinterpret_cast<const uint4*>(table);
    for (uint64_t j = tid; j < chunks; j += T) {
        const uint4 v = table16[j];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int t = 0; t < 16; ++t)
            h += (static_cast<uint64_t>((w[t >> 2] >> (8 * (t & 3))) & 0xffu) + 1) * mix64(16 * j + t);
    }
    for (uint64_t i = 16 * chunks + tid; i < bytes; i += T) h += (static_cast<uint64_t>(table[i]) + 1) * mix64(i);
    unused = 0;
    block_sum_all(h, unused, s_red);
    if (tid == 0) {
        out->verdict = best;
        out->base = base;
        out->won = won;
        out->hash = h;
        out->levels = levels;
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_batch_solve(const BatchGraph* __restrict__ graphs,
                                                        const uint32_t* __restrict__ order, uint32_t count,
                                                        const uint64_t* __restrict__ masks, uint8_t* tables,
                                                        BatchOut* out, uint64_t* scratch, uint64_t slot_words,
                                                        unsigned int* next, uint32_t push_pct) {
    __shared__ uint64_t s_mask[kMaxBN];
    __shared__ uint32_t s_binom[(kMaxBK + 1) * (kBinomCols + 1)];
    __shared__ unsigned long long s_red[2 * kMaxBlockWaves];
    __shared__ uint16_t s_pair[kPairTab];
    __shared__ uint32_t s_t3[kMaxBN];
    __shared__ uint64_t s_occ[kMaxBlockWaves * 64];  // per wave: the cop masks of its 64 configurations
    __shared__ uint64_t s_nib[(kMaxBN + 3) / 4 * 16];  // robber step: robbers by nibble of target vertices
    __shared__ uint8_t s_inl[kMaxBN * (kMaxBN + 1)];  // push: per vertex, its in-degree and in-neighbours
    __shared__ uint32_t s_item;
    uint64_t* slot = scratch + static_cast<uint64_t>(blockIdx.x) * slot_words;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(next, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();  // every thread has read s_item before it is rewritten
        if (item >= count) break;
        const uint32_t gi = order[item];
        const BatchGraph g = graphs[gi];
        switch (g.k) {
            case 1: solve_graph<1>(g, masks, tables, out + gi, slot, s_mask, s_binom, s_red, s_pair, s_t3, s_occ, s_nib, s_inl, push_pct); break;
            case 2: solve_graph<2>(g, masks, tables, out + gi, slot, s_mask, s_binom, s_red, s_pair, s_t3, s_occ, s_nib, s_inl, push_pct); break;
            case 3: solve_graph<3>(g, masks, tables, out + gi, slot, s_mask, s_binom, s_red, s_pair, s_t3, s_occ, s_nib, s_inl, push_pct); break;
            case 4: solve_graph<4>(g, masks, tables, out + gi, slot, s_mask, s_binom, s_red, s_pair, s_t3, s_occ, s_nib, s_inl, push_pct); break;
            default: break;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
using Clock = std::chrono::steady_clock;

uint64_t now_ns() {
    return static_cast<uint64_t>(std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now().time_since_epoch()).count());
}

struct Buf {
    void* p = nullptr;
    uint64_t cap = 0;
    bool pinned = false;
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

int grow(Buf& b, uint64_t bytes, bool pinned) {
    bytes = std::max<uint64_t>(bytes, 64);
    if (b.cap >= bytes) return 0;
    if (b.p) {
        if (b.pinned) hipHostFree(b.p);
        else hipFree(b.p);
    }
    b.p = nullptr;
    b.cap = 0;
    b.pinned = pinned;
    if (pinned) {
        HIP_TRY(hipHostMalloc(&b.p, bytes, hipHostMallocDefault));
    } else {
        size_t freeb = 0, totalb = 0;
        HIP_TRY(hipMemGetInfo(&freeb, &totalb));
        if (bytes + (512ull << 20) > freeb)
            return fail(CR_ERR_TOO_LARGE, "batch needs %llu more device bytes; %zu free",
                        static_cast<unsigned long long>(bytes), freeb);
        HIP_TRY(hipMalloc(&b.p, bytes));
    }
    b.cap = bytes;
    return 0;
}

struct TableRefView {
    const void* base;
    uint64_t offset, bytes;
};

void drop(Buf& b) {
    if (b.p) {
        if (b.pinned) hipHostFree(b.p);
        else hipFree(b.p);
    }
    b = Buf{};
}

// One graph of a batch, validated and sized on the host.
struct Item {
    int32_t status = CR_OK;
    uint32_t n = 0, k = 0, C = 0, hmax = 0;
    uint64_t table_bytes = 0, slot_words = 0;
    double cost = 0;
    uint64_t masks[kMaxBN] = {};
};

// Sizes and validates one graph; returns its status (the message stays in
// the thread's last error).
int prepare(Item& it, int32_t n, int32_t k, const uint8_t* adj) {
    if (n <= 0) return it.status = fail(CR_ERR_EMPTY_GRAPH, "graph is empty");
    if (n > kMaxBN) return it.status = fail(CR_ERR_GRAPH_LIMIT, "batched graphs have at most %d vertices (got %d)", kMaxBN, n);
    if (k <= 0 || k > kMaxBK) return it.status = fail(CR_ERR_K_RANGE, "batched graphs have 1..%d cops (got %d)", kMaxBK, k);
    if (adj == nullptr) return it.status = fail(CR_ERR_ARG, "null matrix");
    uint64_t edges = 0;
    for (int u = 0; u < n; ++u) {
        uint64_t m = 0;
        for (int v = 0; v < n; ++v)
            if (adj[static_cast<size_t>(u) * n + v]) m |= 1ull << v;
        if (m == 0) return it.status = fail(CR_ERR_ISOLATED_VERTEX, "vertex %d has no move", u);
        it.masks[u] = m;
        edges += static_cast<uint64_t>(__builtin_popcountll(m));
    }
    it.n = static_cast<uint32_t>(n);
    it.k = static_cast<uint32_t>(k);
    it.C = static_cast<uint32_t>(cr::config_count(n, k));
    uint64_t hmax = 0, pull = 0;
    for (int i = 1; i <= k; ++i) {
        const uint64_t rows = cr::binom(n + i - 1, i) * cr::binom(n + k - i - 1, k - i);
        if (i < k) hmax = std::max(hmax, rows);
        pull += rows;
    }
    it.hmax = static_cast<uint32_t>(hmax);
    it.table_bytes = 2ull * it.C * n;
    it.slot_words = 5ull * it.C + 2 * hmax + (it.C + 7) / 8;
    it.cost = static_cast<double>(pull) * edges / n + static_cast<double>(it.C) * n;
    return it.status = CR_OK;
}

// Device buffers and staging of one launch slot; they only grow.
struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    Buf d_graphs, d_order, d_masks, d_tables, d_out, d_scratch, d_next;
    Buf h_graphs, h_order, h_masks, h_out;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint32_t count = 0, grid = 0;
    std::vector<uint32_t> index;  // launched graph -> caller's index
    std::vector<BatchGraph> graphs;
    uint64_t total_table_bytes = 0, footprint = 0;
    int cus = 1;
    int32_t threads = 0;        // threads per workgroup: 256, 512, 1024, or 0 = by the largest graph
    int32_t blocks_per_cu = 0;  // resident workgroups per CU (0 = 2048 threads per CU)
    int32_t push_pct = kPushPct;  // a level pushes when its push costs at most push_pct % of the pull
    uint32_t block = kBlock;    // the last launch's workgroup size
    int32_t share = 1;          // launches in flight together (server slots): each gets 1/share of the CUs

    int init(int dev) {
        device = dev;
        HIP_TRY(hipSetDevice(dev));
        HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&ev0));
        HIP_TRY(hipEventCreate(&ev1));
        hipDeviceProp_t prop{};
        HIP_TRY(hipGetDeviceProperties(&prop, dev));
        cus = std::max(1, prop.multiProcessorCount);
        return 0;
    }
    void release() {
        for (Buf* b : {&d_graphs, &d_order, &d_masks, &d_tables, &d_out, &d_scratch, &d_next, &h_graphs, &h_order,
                       &h_masks, &h_out})
            drop(*b);
        if (ev0) hipEventDestroy(ev0);
        if (ev1) hipEventDestroy(ev1);
        if (stream) hipStreamDestroy(stream);
        ev0 = ev1 = nullptr;
        stream = nullptr;
    }

    // Enqueues one launch over the valid items (asynchronous; finish() reads
    // the results after the stream completes). sel holds caller indices.
    int launch(const std::vector<const Item*>& items, const std::vector<uint32_t>& sel) {
        HIP_TRY(hipSetDevice(device));
        count = static_cast<uint32_t>(items.size());
        index = sel;
        graphs.assign(count, BatchGraph{});
        uint64_t table_off = 0, mask_off = 0, slot_words = 64;  // bytes, masks, words
        for (uint32_t i = 0; i < count; ++i) {
            const Item& it = *items[i];
            graphs[i] = BatchGraph{it.n, it.k, it.C, it.hmax, mask_off, table_off};
            table_off += (it.table_bytes + 255) & ~255ull;
            mask_off += it.n;
            slot_words = std::max<uint64_t>(slot_words, (it.slot_words + 31) & ~31ull);
        }
        total_table_bytes = table_off;
        std::vector<uint32_t> order(count);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return items[a]->cost > items[b]->cost; });
        // workgroup size: a graph is solved by one workgroup, so large graphs get more waves
        uint64_t max_c = 0;
        for (uint32_t i = 0; i < count; ++i) max_c = std::max<uint64_t>(max_c, items[i]->C);
        block = threads ? static_cast<uint32_t>(threads) : (max_c >= 4096 ? 1024u : 256u);
        const uint64_t per_cu = blocks_per_cu ? static_cast<uint64_t>(blocks_per_cu) : 2048u / block;
        const uint64_t max_grid = std::max<uint64_t>(1, static_cast<uint64_t>(cus) * per_cu / share);
        // workgroups: enough to fill the device, bounded by scratch memory
        size_t freeb = 0, totalb = 0;
        HIP_TRY(hipMemGetInfo(&freeb, &totalb));
        const uint64_t avail = freeb > table_off + (1ull << 30) ? (freeb - table_off - (1ull << 30)) / 2 : 0;
        uint64_t g = std::min<uint64_t>(count, max_grid);
        g = std::min<uint64_t>(g, std::max<uint64_t>(1, avail / (slot_words * 8)));
        grid = static_cast<uint32_t>(std::max<uint64_t>(g, count ? 1 : 0));
        int rc = 0;
        if ((rc = grow(d_graphs, count * sizeof(BatchGraph), false)) || (rc = grow(d_order, count * 4ull, false)) ||
            (rc = grow(d_masks, mask_off * 8, false)) || (rc = grow(d_tables, table_off, false)) ||
            (rc = grow(d_out, count * sizeof(BatchOut), false)) ||
            (rc = grow(d_scratch, static_cast<uint64_t>(grid) * slot_words * 8, false)) ||
            (rc = grow(d_next, 64, false)) || (rc = grow(h_graphs, count * sizeof(BatchGraph), true)) ||
            (rc = grow(h_order, count * 4ull, true)) || (rc = grow(h_masks, mask_off * 8, true)) ||
            (rc = grow(h_out, count * sizeof(BatchOut), true)))
            return rc;
        footprint = table_off + static_cast<uint64_t>(grid) * slot_words * 8;
        std::memcpy(h_graphs.p, graphs.data(), count * sizeof(BatchGraph));
        std::memcpy(h_order.p, order.data(), count * 4ull);
        uint64_t* hm = h_masks.as<uint64_t>();
        for (uint32_t i = 0; i < count; ++i) std::memcpy(hm + graphs[i].mask_off, items[i]->masks, items[i]->n * 8ull);
        if (count == 0) return 0;
        HIP_TRY(hipMemcpyAsync(d_graphs.p, h_graphs.p, count * sizeof(BatchGraph), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(d_order.p, h_order.p, count * 4ull, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(d_masks.p, h_masks.p, mask_off * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemsetAsync(d_next.p, 0, sizeof(unsigned int), stream));
        HIP_TRY(hipEventRecord(ev0, stream));
        auto args = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, stream, d_graphs.as<BatchGraph>(),
                               d_order.as<uint32_t>(), count, d_masks.as<uint64_t>(), d_tables.as<uint8_t>(),
                               d_out.as<BatchOut>(), d_scratch.as<uint64_t>(), slot_words, d_next.as<unsigned int>(),
                               static_cast<uint32_t>(push_pct));
        };
        if (block == 1024) args(k_batch_solve<1024>);
        else if (block == 512) args(k_batch_solve<512>);
        else args(k_batch_solve<256>);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(ev1, stream));
        HIP_TRY(hipMemcpyAsync(h_out.p, d_out.p, count * sizeof(BatchOut), hipMemcpyDeviceToHost, stream));
        return 0;
    }

    // Converts launched graph i's output (after the stream has completed).
    void result(uint32_t i, const Item& it, cr_graph_result& r) const { to_result(h_out.as<BatchOut>()[i], it, r); }

    static void to_result(const BatchOut& o, const Item& it, cr_graph_result& r) {
        std::memset(&r, 0, sizeof r);
        r.status = CR_OK;
        r.base_captures = o.base;
        r.won_states = o.won;
        r.table_hash = o.hash;
        r.table_bytes = it.table_bytes;
        const uint32_t worst = static_cast<uint32_t>(o.verdict >> 32);
        if (o.verdict != ~0ull && worst < CR_MAX_ROUND) {
            r.win = 1;
            r.plys = static_cast<int32_t>(worst);
            r.best_cid = o.verdict & 0xffffffffull;
            cr::unrank_tuple(r.best_cid, static_cast<int32_t>(it.n), static_cast<int32_t>(it.k), r.positions);
        } else {
            r.plys = -1;
        }
        r.n_passes = static_cast<int32_t>(pass_counts(o).size());
    }

    // The reference's per-pass counts from the depth histogram: pass i
    // reports cop states of depth 2i-2, 2i-1 and robber states of depth 2i-1,
    // 2i (depth > 0); the pass that finds nothing ends the loop.
    static std::vector<uint64_t> pass_counts(const BatchOut& o) {
        std::vector<uint64_t> pc;
        for (int i = 1;; ++i) {
            uint64_t c = 0;
            const int c0 = 2 * i - 2, c1 = 2 * i - 1, r0 = 2 * i - 1, r1 = 2 * i;
            if (c0 > 0 && c0 < CR_MAX_ROUND) c += o.hist[c0][0];
            if (c1 < CR_MAX_ROUND) c += o.hist[c1][0];
            if (r0 < CR_MAX_ROUND) c += o.hist[r0][1];
            if (r1 < CR_MAX_ROUND) c += o.hist[r1][1];
            pc.push_back(c);
            if (c == 0) break;
        }
        return pc;
    }
};

}  // namespace

struct cr_batch {
    Engine eng;
    std::vector<BatchOut> last;  // outputs of the last solve, caller order
    std::vector<uint8_t> valid;
    double device_ms = 0;
};

struct cr_server {
    struct Request {
        uint64_t ticket;
        uint64_t arrival_ns;
        Item item;
    };
    // The state tables of one finished launch, kept in HBM for read-back:
    // the launch slot's table buffer itself (the slot grows a new one).
    struct Retained {
        Buf buf;
        uint32_t refs = 0;  // tickets not released yet
        uint64_t seq = 0;   // completion order (oldest evicted first)
    };
    struct TableRef {
        std::shared_ptr<Retained> owner;
        uint64_t offset = 0, bytes = 0;
    };
    struct Slot {
        Engine eng;
        bool busy = false;
        std::vector<Request> reqs;
    };
    int device = 0;
    uint64_t budget = 0, in_use = 0, peak = 0;
    int32_t max_batch = 1024;
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> pending;
    std::deque<cr_graph_result> done;
    uint64_t next_ticket = 1, submitted = 0, finished = 0, launches = 0;
    bool retain = false;                 // keep finished tables resident (cr_server_retain_tables)
    std::unordered_map<uint64_t, TableRef> tables;  // ticket -> its resident table
    std::deque<std::shared_ptr<Retained>> retained;  // in completion order
    uint64_t retained_bytes = 0, retained_seq = 0, evicted = 0;
    bool stopping = false;

    void drop_retained(const std::shared_ptr<Retained>& r);
    void evict_for(uint64_t need);
    std::thread th;
    std::string thread_error;

    void run();
};

// Frees a retained launch's tables (every ticket of it loses its table).
void cr_server::drop_retained(const std::shared_ptr<Retained>& r) {
    for (auto it = tables.begin(); it != tables.end();)
        it = it->second.owner == r ? tables.erase(it) : std::next(it);
    retained_bytes -= r->buf.cap;
    in_use -= r->buf.cap;
    drop(r->buf);
    retained.erase(std::remove(retained.begin(), retained.end(), r), retained.end());
}

// Retained tables live within the HBM budget: before admitting `need` more
// bytes that would not fit, the oldest retained launches are evicted.
void cr_server::evict_for(uint64_t need) {
    while (!retained.empty() && in_use + need > budget) {
        ++evicted;
        drop_retained(retained.front());
    }
}

void cr_server::run() {
    hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        bool progress = false;
        // harvest finished launches
        for (auto& sp : slots) {
            Slot& s = *sp;
            if (!s.busy || hipEventQuery(s.eng.ev1) != hipSuccess) continue;
            if (hipStreamQuery(s.eng.stream) != hipSuccess) continue;  // results still copying
            const uint64_t t = now_ns();
            std::shared_ptr<Retained> keep;
            if (retain && !s.reqs.empty()) {  // the slot's table buffer becomes the launch's resident tables
                keep = std::make_shared<Retained>();
                keep->buf = s.eng.d_tables;
                s.eng.d_tables = Buf{};
                keep->refs = static_cast<uint32_t>(s.reqs.size());
                keep->seq = ++retained_seq;
                retained.push_back(keep);
                retained_bytes += keep->buf.cap;
            }
            for (uint32_t i = 0; i < s.reqs.size(); ++i) {
                cr_graph_result r;
                s.eng.result(i, s.reqs[i].item, r);
                r.ticket = s.reqs[i].ticket;
                r.latency_ms = static_cast<double>(t - s.reqs[i].arrival_ns) / 1e6;
                done.push_back(r);
                if (keep) tables[r.ticket] = TableRef{keep, s.eng.graphs[i].table_off, s.reqs[i].item.table_bytes};
            }
            finished += s.reqs.size();
            in_use -= s.eng.footprint;
            if (keep) in_use += keep->buf.cap;
            s.reqs.clear();
            s.busy = false;
            progress = true;
        }
        if (progress) cv_done.notify_all();
        // admit pending requests into a free slot
        for (auto& sp : slots) {
            Slot& s = *sp;
            if (s.busy || pending.empty()) continue;
            std::vector<Request> take;
            uint64_t bytes = 0;
            while (!pending.empty() && static_cast<int32_t>(take.size()) < max_batch) {
                const Item& it = pending.front().item;
                const uint64_t need = it.table_bytes + it.slot_words * 8;
                evict_for(bytes + need);
                if (in_use + bytes + need > budget && !take.empty()) break;
                if (in_use + bytes + need > budget && in_use > 0) break;  // wait for memory to free up
                bytes += need;
                take.push_back(std::move(pending.front()));
                pending.pop_front();
            }
            if (take.empty()) break;
            s.reqs = std::move(take);
            std::vector<const Item*> items;
            std::vector<uint32_t> sel;
            for (uint32_t i = 0; i < s.reqs.size(); ++i) {
                items.push_back(&s.reqs[i].item);
                sel.push_back(i);
            }
            lk.unlock();
            const int rc = s.eng.launch(items, sel);
            lk.lock();
            if (rc != 0) {  // report every request of the failed launch
                for (auto& rq : s.reqs) {
                    cr_graph_result r{};
                    r.ticket = rq.ticket;
                    r.status = rc;
                    r.plys = -1;
                    done.push_back(r);
                }
                finished += s.reqs.size();
                s.reqs.clear();
                cv_done.notify_all();
                continue;
            }
            s.busy = true;
            in_use += s.eng.footprint;
            peak = std::max(peak, in_use);
            ++launches;
            progress = true;
        }
        bool busy = false;
        for (auto& sp : slots) busy = busy || sp->busy;
        if (stopping && !busy && pending.empty()) break;
        if (!progress) {
            if (busy) cv_work.wait_for(lk, std::chrono::microseconds(20));
            else cv_work.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
}

extern "C" {

int cr_batch_create(int device, cr_batch** out) {
    if (out == nullptr) return fail(CR_ERR_ARG, "null output pointer");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(CR_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count) return fail(CR_ERR_ARG, "device %d of %d", device, count);
    auto b = std::make_unique<cr_batch>();
    if (int rc = b->eng.init(device)) return rc;
    *out = b.release();
    return CR_OK;
}

void cr_batch_destroy(cr_batch* b) {
    if (!b) return;
    hipSetDevice(b->eng.device);
    if (b->eng.stream) hipStreamSynchronize(b->eng.stream);
    b->eng.release();
    delete b;
}

int cr_batch_solve(cr_batch* b, int32_t count, const int32_t* n, const int32_t* k, const uint8_t* adj,
                   const uint64_t* adj_offset, cr_graph_result* results, uint8_t* tables, const uint64_t* table_offset) {
    if (b == nullptr) return fail(CR_ERR_ARG, "null batch");
    if (count < 0 || (count > 0 && (n == nullptr || k == nullptr || adj == nullptr || adj_offset == nullptr ||
                                    results == nullptr)))
        return fail(CR_ERR_ARG, "null argument");
    if (tables != nullptr && table_offset == nullptr) return fail(CR_ERR_ARG, "tables without offsets");
    std::vector<Item> items(count);
    std::vector<const Item*> ok;
    std::vector<uint32_t> sel;
    int first_bad = CR_OK;
    char first_msg[512] = "";
    for (int32_t i = 0; i < count; ++i) {
        const int st = prepare(items[i], n[i], k[i], adj + adj_offset[i]);
        std::memset(&results[i], 0, sizeof results[i]);
        results[i].ticket = static_cast<uint64_t>(i);
        results[i].status = st;
        results[i].plys = -1;
        if (st == CR_OK) {
            ok.push_back(&items[i]);
            sel.push_back(static_cast<uint32_t>(i));
        } else if (first_bad == CR_OK) {
            first_bad = st;
            std::snprintf(first_msg, sizeof first_msg, "graph %d: %s", i, crd::g_err);
        }
    }
    b->last.assign(count, BatchOut{});
    b->valid.assign(count, 0);
    Engine& e = b->eng;
    if (int rc = e.launch(ok, sel)) return rc;
    HIP_TRY(hipStreamSynchronize(e.stream));
    float ms = 0;
    if (!ok.empty()) hipEventElapsedTime(&ms, e.ev0, e.ev1);
    b->device_ms = ms;
    for (uint32_t i = 0; i < ok.size(); ++i) {
        const uint32_t ci = sel[i];
        e.result(i, *ok[i], results[ci]);
        results[ci].ticket = ci;
        b->last[ci] = e.h_out.as<BatchOut>()[i];
        b->valid[ci] = 1;
        if (tables)
            HIP_TRY(hipMemcpy(tables + table_offset[ci], e.d_tables.as<uint8_t>() + e.graphs[i].table_off,
                              ok[i]->table_bytes, hipMemcpyDeviceToHost));
    }
    if (first_bad != CR_OK) return fail(CR_ERR_ARG, "%s", first_msg);
    return CR_OK;
}

namespace {
int set_engine_tuning(Engine& e, int32_t key, int64_t value) {
    switch (key) {
        case CR_BATCH_TUNE_THREADS:
            if (value != 0 && value != 256 && value != 512 && value != 1024)
                return fail(CR_ERR_ARG, "threads per workgroup must be 0, 256, 512 or 1024");
            e.threads = static_cast<int32_t>(value);
            return CR_OK;
        case CR_BATCH_TUNE_BLOCKS_PER_CU:
            if (value < 0 || value > 32) return fail(CR_ERR_ARG, "workgroups per CU must be 0..32");
            e.blocks_per_cu = static_cast<int32_t>(value);
            return CR_OK;
        case CR_BATCH_TUNE_PUSH_PCT:
            if (value < 0 || value > 1000000000) return fail(CR_ERR_ARG, "push threshold must be 0..10^9 %%");
            e.push_pct = static_cast<int32_t>(value);
            return CR_OK;
        default: return fail(CR_ERR_ARG, "unknown batch tuning key %d", key);
    }
}
}  // namespace

int cr_batch_set_tuning(cr_batch* b, int32_t key, int64_t value) {
    if (b == nullptr) return fail(CR_ERR_ARG, "null batch");
    return set_engine_tuning(b->eng, key, value);
}

int cr_server_set_tuning(cr_server* s, int32_t key, int64_t value) {
    if (s == nullptr) return fail(CR_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& sl : s->slots) {
        if (sl->busy) return fail(CR_ERR_ORDER, "tune the server before submitting");
        if (int rc = set_engine_tuning(sl->eng, key, value)) return rc;
    }
    return CR_OK;
}

int cr_batch_pass_counts(cr_batch* b, int32_t index, uint64_t* out, int32_t cap, int32_t* n_passes) {
    if (b == nullptr) return fail(CR_ERR_ARG, "null batch");
    if (index < 0 || index >= static_cast<int32_t>(b->last.size()) || !b->valid[index])
        return fail(CR_ERR_ARG, "no solved graph %d in the last batch", index);
    const std::vector<uint64_t> pc = Engine::pass_counts(b->last[index]);
    for (int32_t i = 0; out && i < cap && i < static_cast<int32_t>(pc.size()); ++i) out[i] = pc[i];
    if (n_passes) *n_passes = static_cast<int32_t>(pc.size());
    return CR_OK;
}

int cr_batch_last_timing(cr_batch* b, double* device_ms, int32_t* workgroups) {
    if (b == nullptr) return fail(CR_ERR_ARG, "null batch");
    if (device_ms) *device_ms = b->device_ms;
    if (workgroups) *workgroups = static_cast<int32_t>(b->eng.grid);
    return CR_OK;
}

uint64_t cr_clock_ns(void) { return now_ns(); }

int cr_server_create(int device, uint64_t hbm_budget, int32_t max_batch, int32_t streams, cr_server** out) {
    if (out == nullptr) return fail(CR_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (streams < 1 || streams > 4) return fail(CR_ERR_ARG, "streams must be 1..4 (got %d)", streams);
    if (max_batch < 1) return fail(CR_ERR_ARG, "max_batch must be positive");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(CR_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count) return fail(CR_ERR_ARG, "device %d of %d", device, count);
    auto s = std::make_unique<cr_server>();
    s->device = device;
    s->budget = hbm_budget;
    s->max_batch = max_batch;
    for (int i = 0; i < streams; ++i) {
        auto sl = std::make_unique<cr_server::Slot>();
        if (int rc = sl->eng.init(device)) return rc;
        sl->eng.share = streams;  // concurrent launches split the device instead of queueing behind each other
        s->slots.push_back(std::move(sl));
    }
    cr_server* raw = s.get();
    s->th = std::thread([raw] { raw->run(); });
    *out = s.release();
    return CR_OK;
}

int cr_server_submit(cr_server* s, int32_t n, int32_t k, const uint8_t* adj, uint64_t arrival_ns, uint64_t* ticket) {
    if (s == nullptr) return fail(CR_ERR_ARG, "null server");
    cr_server::Request rq;
    if (int rc = prepare(rq.item, n, k, adj)) return rc;
    const uint64_t need = rq.item.table_bytes + rq.item.slot_words * 8;
    if (need > s->budget) return fail(CR_ERR_TOO_LARGE, "request needs %llu bytes; the budget is %llu",
                                      static_cast<unsigned long long>(need), static_cast<unsigned long long>(s->budget));
    rq.arrival_ns = arrival_ns ? arrival_ns : now_ns();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) return fail(CR_ERR_ORDER, "server is stopping");
        rq.ticket = s->next_ticket++;
        if (ticket) *ticket = rq.ticket;
        s->pending.push_back(std::move(rq));
        ++s->submitted;
    }
    s->cv_work.notify_one();
    return CR_OK;
}

int cr_server_submit_many(cr_server* s, int32_t count, const int32_t* n, const int32_t* k, const uint8_t* adj,
                          const uint64_t* adj_offset, const uint64_t* arrival_ns, uint64_t* tickets,
                          int32_t* status) {
    if (s == nullptr) return fail(CR_ERR_ARG, "null server");
    if (count < 0 || (count > 0 && (n == nullptr || k == nullptr || adj == nullptr || adj_offset == nullptr ||
                                    tickets == nullptr)))
        return fail(CR_ERR_ARG, "null argument");
    std::vector<cr_server::Request> reqs;
    reqs.reserve(count);
    std::vector<int32_t> st(count, CR_OK);
    const uint64_t now = now_ns();
    for (int32_t i = 0; i < count; ++i) {
        cr_server::Request rq;
        st[i] = prepare(rq.item, n[i], k[i], adj + adj_offset[i]);
        if (st[i] == CR_OK && rq.item.table_bytes + rq.item.slot_words * 8 > s->budget)
            st[i] = CR_ERR_TOO_LARGE;
        rq.arrival_ns = arrival_ns && arrival_ns[i] ? arrival_ns[i] : now;
        rq.ticket = 0;
        reqs.push_back(std::move(rq));
    }
    bool rejected = false;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) return fail(CR_ERR_ORDER, "server is stopping");
        for (int32_t i = 0; i < count; ++i) {
            if (status) status[i] = st[i];
            if (st[i] != CR_OK) {
                tickets[i] = 0;
                rejected = true;
                continue;
            }
            reqs[i].ticket = s->next_ticket++;
            tickets[i] = reqs[i].ticket;
            s->pending.push_back(std::move(reqs[i]));
            ++s->submitted;
        }
    }
    s->cv_work.notify_one();
    return rejected ? fail(CR_ERR_ARG, "some requests were rejected (see status)") : CR_OK;
}

int cr_server_poll(cr_server* s, cr_graph_result* out, int32_t cap, int32_t wait_us, int32_t* count) {
    if (s == nullptr || (cap > 0 && out == nullptr)) return fail(CR_ERR_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_us > 0)
        s->cv_done.wait_for(lk, std::chrono::microseconds(wait_us), [s] { return !s->done.empty(); });
    int32_t m = 0;
    while (m < cap && !s->done.empty()) {
        out[m++] = s->done.front();
        s->done.pop_front();
    }
    if (count) *count = m;
    return CR_OK;
}

int cr_server_stats(cr_server* s, uint64_t* submitted, uint64_t* finished, uint64_t* launches, uint64_t* peak_bytes) {
    if (s == nullptr) return fail(CR_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (submitted) *submitted = s->submitted;
    if (finished) *finished = s->finished;
    if (launches) *launches = s->launches;
    if (peak_bytes) *peak_bytes = s->peak;
    return CR_OK;
}

int cr_server_retain_tables(cr_server* s, int32_t on) {
    if (s == nullptr) return fail(CR_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->retain = on != 0;
    return CR_OK;
}

int cr_server_copy_table(cr_server* s, uint64_t ticket, uint8_t* dst, uint64_t cap, uint64_t* bytes) {
    if (s == nullptr) return fail(CR_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tables.find(ticket);
    if (it == s->tables.end())
        return fail(CR_ERR_ARG, "no resident table for ticket %llu (not finished, not retained, released or evicted)",
                    static_cast<unsigned long long>(ticket));
    const TableRefView ref{it->second.owner->buf.p, it->second.offset, it->second.bytes};
    if (bytes) *bytes = ref.bytes;
    if (dst == nullptr) return CR_OK;  // size query
    if (cap < ref.bytes) return fail(CR_ERR_ARG, "table of ticket %llu has %llu bytes; buffer holds %llu",
                                     static_cast<unsigned long long>(ticket), static_cast<unsigned long long>(ref.bytes),
                                     static_cast<unsigned long long>(cap));
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMemcpy(dst, static_cast<const uint8_t*>(ref.base) + ref.offset, ref.bytes, hipMemcpyDeviceToHost));
    return CR_OK;
}

int cr_server_release_table(cr_server* s, uint64_t ticket) {
    if (s == nullptr) return fail(CR_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tables.find(ticket);
    if (it == s->tables.end()) return fail(CR_ERR_ARG, "no resident table for ticket %llu", static_cast<unsigned long long>(ticket));
    std::shared_ptr<cr_server::Retained> owner = it->second.owner;
    s->tables.erase(it);
    if (--owner->refs == 0) s->drop_retained(owner);
    return CR_OK;
}

int cr_server_table_stats(cr_server* s, uint64_t* resident_tables, uint64_t* resident_bytes, uint64_t* evicted_launches) {
    if (s == nullptr) return fail(CR_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (resident_tables) *resident_tables = s->tables.size();
    if (resident_bytes) *resident_bytes = s->retained_bytes;
    if (evicted_launches) *evicted_launches = s->evicted;
    return CR_OK;
}

void cr_server_destroy(cr_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    hipSetDevice(s->device);
    for (auto& sl : s->slots) {
        if (sl->eng.stream) hipStreamSynchronize(sl->eng.stream);
        sl->eng.release();
    }
    for (auto& r : s->retained) drop(r->buf);
    delete s;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// cr_batch_solve_noalloc: the batch launch with every work buffer in one
// file-scope device array, for callers that cannot allocate.
// ---------------------------------------------------------------------------
namespace {

constexpr uint64_t kNoAllocScratchBytes = CR_NOALLOC_SCRATCH_BYTES;
constexpr int32_t kNoAllocMaxGraphs = CR_NOALLOC_MAX_GRAPHS;
constexpr int kNoAllocMaxDevices = 64;

// The device side of every call on a device: graph descriptors, work order,
// robber masks, the work counter, per-graph outputs and the workgroups' work
// planes. Each device holds its own copy until the process ends.
__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

// Host staging in one block, uploaded by one copy (the counter is zeroed by
// that copy). Static, so the call allocates nothing on the host either.
struct NoAllocStage {
    BatchGraph graphs[kNoAllocMaxGraphs];
    uint32_t order[kNoAllocMaxGraphs];
    uint32_t next[64];  // the work counter, alone in its 256 bytes
    uint64_t masks[static_cast<size_t>(kNoAllocMaxGraphs) * kMaxBN];
};
static_assert(sizeof(BatchGraph) % 8 == 0 && offsetof(NoAllocStage, next) % 8 == 0 &&
              offsetof(NoAllocStage, masks) % 8 == 0, "staging layout");

NoAllocStage g_stage;
BatchOut g_out[kNoAllocMaxGraphs];
Item g_items[kNoAllocMaxGraphs];
uint32_t g_valid[kNoAllocMaxGraphs];  // caller index of each launched graph

// Layout of the device array (byte offsets; each 256-aligned).
constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }
constexpr uint64_t kOffStage = 0;
constexpr uint64_t kOffOut = align256(kOffStage + sizeof(NoAllocStage));
constexpr uint64_t kOffSlots = align256(kOffOut + sizeof(BatchOut) * kNoAllocMaxGraphs);
static_assert(kOffSlots < kNoAllocScratchBytes, "scratch holds the staging, outputs and some work planes");

std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on the current device `dev`, looked up once
// per device. The array's host-side name is not a device address; only this
// pointer is passed to kernels and copies.
int scratch_on(int dev, uint8_t** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr)
            return fail(CR_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = static_cast<uint8_t*>(g_scratch_ptr[dev]);
    return CR_OK;
}

}  // namespace

extern "C" int cr_batch_solve_noalloc(int device, int32_t count, const int32_t* n, const int32_t* k,
                                      const uint8_t* adj, const uint64_t* adj_offset, cr_graph_result* results,
                                      uint8_t* d_tables, uint64_t tables_cap, uint64_t* table_offsets) {
    if (count < 0 || (count > 0 && (n == nullptr || k == nullptr || adj == nullptr || adj_offset == nullptr ||
                                    results == nullptr)) || table_offsets == nullptr)
        return fail(CR_ERR_ARG, "null argument");
    if (count > kNoAllocMaxGraphs)
        return fail(CR_ERR_TOO_LARGE, "at most %d graphs per no-allocation call (got %d)", kNoAllocMaxGraphs, count);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(CR_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= ndev || device >= kNoAllocMaxDevices) return fail(CR_ERR_ARG, "device %d of %d", device, ndev);
    // validate and size every graph; tables are laid out 256-byte aligned in caller order
    int first_bad = CR_OK;
    char first_msg[512] = "";
    uint32_t nvalid = 0;
    uint64_t table_off = 0, mask_off = 0, slot_words = 64, max_c = 0;
    for (int32_t i = 0; i < count; ++i) {
        Item& it = g_items[i];
        it = Item{};
        const int st = prepare(it, n[i], k[i], adj + adj_offset[i]);
        std::memset(&results[i], 0, sizeof results[i]);
        results[i].ticket = static_cast<uint64_t>(i);
        results[i].status = st;
        results[i].plys = -1;
        table_offsets[i] = table_off;
        if (st != CR_OK) {
            if (first_bad == CR_OK) {
                first_bad = st;
                std::snprintf(first_msg, sizeof first_msg, "graph %d: %s", i, crd::g_err);
            }
            continue;
        }
        g_stage.graphs[nvalid] = BatchGraph{it.n, it.k, it.C, it.hmax, mask_off, table_off};
        std::memcpy(g_stage.masks + mask_off, it.masks, it.n * 8ull);
        g_valid[nvalid++] = static_cast<uint32_t>(i);
        table_off += align256(it.table_bytes);
        mask_off += it.n;
        slot_words = std::max<uint64_t>(slot_words, (it.slot_words + 31) & ~31ull);
        max_c = std::max<uint64_t>(max_c, it.C);
    }
    table_offsets[count] = table_off;
    if (d_tables == nullptr) return first_bad == CR_OK ? CR_OK : fail(CR_ERR_ARG, "%s", first_msg);  // sizing only
    if (tables_cap < table_off)
        return fail(CR_ERR_TOO_LARGE, "the tables need %llu bytes; d_tables holds %llu",
                    static_cast<unsigned long long>(table_off), static_cast<unsigned long long>(tables_cap));
    const uint64_t slot_bytes = slot_words * 8;
    const uint64_t slot_room = kNoAllocScratchBytes - kOffSlots;
    if (nvalid > 0 && slot_bytes > slot_room)
        return fail(CR_ERR_TOO_LARGE, "a graph needs %llu bytes of work planes; the no-allocation scratch has %llu",
                    static_cast<unsigned long long>(slot_bytes), static_cast<unsigned long long>(slot_room));
    // largest first, ties by position (std::sort: no temporary buffer)
    for (uint32_t i = 0; i < nvalid; ++i) g_stage.order[i] = i;
    std::sort(g_stage.order, g_stage.order + nvalid, [](uint32_t a, uint32_t b) {
        const double ca = g_items[g_valid[a]].cost, cb = g_items[g_valid[b]].cost;
        return ca != cb ? ca > cb : a < b;
    });
    std::memset(g_stage.next, 0, sizeof g_stage.next);
    HIP_TRY(hipSetDevice(device));
    uint8_t* base = nullptr;
    if (int rc = scratch_on(device, &base)) return rc;
    if (nvalid > 0) {
        int cus = 1;
        HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        const uint32_t block = max_c >= 4096 ? 1024u : 256u;
        uint64_t grid = std::min<uint64_t>(nvalid, static_cast<uint64_t>(std::max(cus, 1)) * (2048u / block));
        grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, slot_room / slot_bytes));
        // the staging block up to the masks in use, in one copy (graphs, order, zeroed counter, masks)
        const uint64_t stage_bytes = offsetof(NoAllocStage, masks) + mask_off * 8;
        HIP_TRY(hipMemcpyAsync(base + kOffStage, &g_stage, stage_bytes, hipMemcpyHostToDevice, nullptr));
        NoAllocStage* ds = reinterpret_cast<NoAllocStage*>(base + kOffStage);
        BatchOut* dout = reinterpret_cast<BatchOut*>(base + kOffOut);
        uint64_t* slots = reinterpret_cast<uint64_t*>(base + kOffSlots);
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(static_cast<uint32_t>(grid)), dim3(block), 0, nullptr, ds->graphs, ds->order,
                               nvalid, ds->masks, d_tables, dout, slots, slot_words, ds->next, kPushPct);
        };
        if (block == 1024) go(k_batch_solve<1024>);
        else go(k_batch_solve<256>);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(g_out, dout, nvalid * sizeof(BatchOut), hipMemcpyDeviceToHost, nullptr));
        HIP_TRY(hipStreamSynchronize(nullptr));
    }
    for (uint32_t i = 0; i < nvalid; ++i) {
        const uint32_t ci = g_valid[i];
        Engine::to_result(g_out[i], g_items[ci], results[ci]);
        results[ci].ticket = ci;
    }
    if (first_bad != CR_OK) return fail(CR_ERR_ARG, "%s", first_msg);
    return CR_OK;
}
