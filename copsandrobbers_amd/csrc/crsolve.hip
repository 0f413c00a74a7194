// crsolve.hip -- MI355X (gfx950) k-cop retrograde solver behind the C ABI of
// include/crsolve.h.
//
// State space (reference: lib/include/AuxGraph.h:59-113): a configuration C is
// a sorted k-tuple of cop vertices, identified by its lexicographic rank cId
// (AuxGraph.h:117-177 enumeration order, computed here in closed form); a
// state is (cId, r, turn).
//
// Device layout for the plain game (p = 1, k_cops_visibility.cpp with two
// columns):
//   * bit planes, one row of W = ceil(N/64) 64-bit words per configuration,
//     bit r = robber vertex r:
//       won0 / won1  states already won with the cops / the robber to move
//       fr0          cop-turn states won at the latest depth
//       fr1[2]       robber-turn states won at the current depth (frontier),
//                    by depth parity: read by the cop step, written by the
//                    robber step of the level before
//       nx0          cop-turn candidates of the next depth (atomicOr target)
//   * table: the reference's AoS DataItem bytes, index (cId*N + r)*2 + turn,
//     written once per state when it is won.
//   * row lists of configurations holding new states (compacted with
//     wavefront ballots and one atomic per wavefront).
//
// One depth level d -> d+1 (level-synchronous retrograde analysis; equal to
// the reference's pass sweep, whose pass i marks cop states of depth 2i-2 and
// 2i-1 and robber states of depth 2i-1 and 2i), robber step first (both
// halves read only depths <= d):
//   k_cop_push      every configuration C' with new robber-turn states at
//                   depth d ORs its robber mask into nx0 of each predecessor C
//                   (each cop steps back along an in-edge; the product is
//                   enumerated, sorted and ranked in registers). Cop turn =
//                   "exists a move to a won state", depth = min + 1
//                   (k_cops_visibility.cpp:160-197).
//   k_robber_step   for every configuration with new cop-turn states at depth
//                   d, robber r is won iff every move of r lands on a won
//                   cop-turn state: out_mask[r] subset of won0 row; depth =
//                   max + 1 (k_cops_visibility.cpp:235-288 with p = 1).
//   k_cop_finalize  nx0 & ~won0 becomes the cop-turn states of depth d+1.
// States whose depth would reach 127 stay unmarked (the 7-bit cap, :189,
// :220, :280), so levels run d = 0..125.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "cr_device.h"
#include "cr_columns.h"
#include "cr_game.h"
#include "cr_host.h"
#include "cr_macro.h"
#include "crsolve.h"

namespace crd {
thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}
}  // namespace crd

namespace {

struct Timer;  // per-launch HIP-event pair (host side)

struct Ctl {
    uint32_t level;    // frontier depth d of the level being processed
    uint32_t done;     // set when no level remains
    uint32_t cnt0[2];  // per parity: length of the list of configurations with new cop-turn states
    uint32_t cnt1[2];  // ... with new robber-turn states
    uint32_t pull;     // the level's cop turn gathers (pull) instead of pushing
    uint32_t dir[128]; // per level: 1 = pulled
    // depth log: the lists of new states of each depth with their robber
    // bits, applied to the table by k_apply_log (see there)
    uint32_t log_now;               // the current depth's lists go to the log (else straight to the table)
    uint32_t log_open;              // the lists of the depth that ended the loop are still to be written
    unsigned long long log_used;    // entries in use
    unsigned long long log_cap;     // capacity in entries (0: no log)
    uint32_t log_first;             // the first depth of the log since the last application
    unsigned long long log_off[128][2];  // per depth and turn: first entry
    uint32_t log_len[128][2];            // per depth and turn: entries (0 when written directly)
    unsigned long long lvl[128][2];
    uint32_t rows[128][2];  // per depth: configurations listed with new cop-turn / robber-turn states
    unsigned long long hist[2][128];  // retro-edges by depth: [0] cop (robber states), [1] robber (cop states)
    unsigned long long verdict;
    unsigned long long retro[3];
    // k_end_level (single-pass level end): per level, the tile tickets and the
    // spaces that have finished; per space, the states of its tiles and its total
    uint32_t lb_ticket[128];
    uint32_t lb_arrived[128];
    unsigned long long lb_states[2];
    uint32_t lb_total[2];
    uint32_t done_level;  // the level whose end set `done`
};

// Row sets are compacted without contended atomics. A producer kernel walks an
// index space in groups of kGroupWords bitmap words (one workgroup per group,
// one wavefront per word at a time): it stores one flag word per 64 indices and
// the group's flag and state counts with plain stores. k_advance scans the
// group counts into offsets, and k_write_lists turns the flag words into a
// sorted dense list.
constexpr int kGroupWords = kWavesPerBlock;  // one flag word per wavefront and group
constexpr int kWordsPerWave = kGroupWords / kWavesPerBlock;

struct Compact {
    uint64_t* flags;              // one bit per index of the producer's space
    uint32_t* gcnt;               // per group: set flags; after the scan: offset
    unsigned long long* gstates;  // per group: states won
};

struct LevelArgs {
    uint64_t* won0;
    uint64_t* won1;
    // robber-turn frontier by depth parity: the cop step of depth d reads (and
    // clears) fr1s[d & 1] and the robber step writes depth d + 1 into
    // fr1s[(d + 1) & 1]. Two buffers let the robber step run first (fused
    // finalize); otherwise both are one buffer and the cop step runs first.
    uint64_t* fr1s[2];
    uint64_t* nx0;
    uint8_t* table;
    uint32_t* list0[2];  // per parity: configurations with new cop-turn states (sorted)
    uint32_t* list1[2];  // ... with new robber-turn states
    Compact c0;          // cop-turn rows (index space: configurations)
    Compact c1;          // robber-turn rows (index space: entries of list0, or configurations at init)
    unsigned long long* gcaps;  // per group of k_init: capture states
    Ctl* ctl;
    const uint64_t* out_mask;  // [n][4] robber moves
    const uint64_t* in_mask;   // [n][4] robber positions with a move onto each vertex
    uint64_t* fr0;             // per configuration: the cop-turn states won at the latest depth
    uint32_t nbatch;           // ceil(C / 64)
    int32_t filter;
    uint32_t pull_rows;        // a level pulls when its robber frontier has more rows than this
    uint32_t* log_rows;        // depth log: configuration of each entry
    uint64_t* log_bits;        // depth log: W robber words of each entry
    uint32_t* log_beg;         // depth log: per list and chunk of log_R rows, the first and one past the
    uint32_t* log_end;         //   last entry with rows in the chunk (end 0: none), [kMaxLogLists][log_chunks]
    uint32_t log_R, log_chunks;
    int32_t n, W;
    int32_t fuse;              // pulled levels: the last pull stage does k_cop_finalize's work
    uint32_t open_factor;      // robber step: rows with open <= open_factor * new cop-turn states take the open-state path
    uint8_t* dirty;            // per group of 256 configurations: the push wrote nx0 rows there (null: not kept)
    // two-level scan of the group counts (0: one workgroup scans them all):
    // k_scan_tiles scans tiles of scan_tile groups, k_advance the tile totals
    uint32_t scan_tile;
    uint32_t tile_stride;                 // entries per list in tile_cnt / tile_states
    uint32_t* tile_cnt;                   // [2][tile_stride]: tile totals, then offsets
    unsigned long long* tile_states;      // [2][tile_stride]
    unsigned long long* lb_status;        // k_end_level: [2][lb_stride] published tile sums (epoch, flag, value)
    uint32_t lb_stride;
};

// Work items [0, n) in 8 contiguous segments, one per XCD (workgroups are
// dispatched round-robin over the XCDs, so workgroup b runs on XCD b % 8 and
// the grid is a multiple of 8): neighbouring configurations, whose moves
// reach overlapping rows, are processed by one XCD and share its L2.
constexpr uint32_t kXcds = 8;
struct XcdSplit {
    uint32_t end, first, stride;
};
__device__ __forceinline__ uint32_t wave_id();
__device__ __forceinline__ XcdSplit xcd_split(uint32_t n) {
    const uint32_t x = blockIdx.x % kXcds, bpx = gridDim.x / kXcds, lb = blockIdx.x / kXcds;
    const uint32_t begin = static_cast<uint32_t>(static_cast<uint64_t>(n) * x / kXcds);
    const uint32_t end = static_cast<uint32_t>(static_cast<uint64_t>(n) * (x + 1) / kXcds);
    return XcdSplit{end, begin + lb * kWavesPerBlock + wave_id(), bpx * kWavesPerBlock};
}

// Whether the level starting with `frontier_rows` robber-frontier rows pulls.
__device__ __forceinline__ uint32_t choose_pull(uint32_t frontier_rows, uint32_t threshold) {
    return frontier_rows > threshold ? 1u : 0u;
}

// Loads / stores of one W-word row: 16-byte vector accesses where the row is
// 16-byte aligned (W = 2, 4).
template <int W>
__device__ __forceinline__ void load_row(const uint64_t* __restrict__ src, uint64_t* v) {
    if constexpr (W == 4) {
        const ulonglong2 x = reinterpret_cast<const ulonglong2*>(src)[0];
        const ulonglong2 y = reinterpret_cast<const ulonglong2*>(src)[1];
        v[0] = x.x;
        v[1] = x.y;
        v[2] = y.x;
        v[3] = y.y;
    } else if constexpr (W == 2) {
        const ulonglong2 x = reinterpret_cast<const ulonglong2*>(src)[0];
        v[0] = x.x;
        v[1] = x.y;
    } else {
#pragma unroll
        for (int q = 0; q < W; ++q) v[q] = src[q];
    }
}

template <int W>
__device__ __forceinline__ void store_row(uint64_t* __restrict__ dst, const uint64_t* v) {
    if constexpr (W == 4) {
        reinterpret_cast<ulonglong2*>(dst)[0] = make_ulonglong2(v[0], v[1]);
        reinterpret_cast<ulonglong2*>(dst)[1] = make_ulonglong2(v[2], v[3]);
    } else if constexpr (W == 2) {
        reinterpret_cast<ulonglong2*>(dst)[0] = make_ulonglong2(v[0], v[1]);
    } else {
#pragma unroll
        for (int q = 0; q < W; ++q) dst[q] = v[q];
    }
}

// ---------------------------------------------------------------------------
// init: reset every state, seed captures (k_cops_visibility.cpp:85-127) and
// resolve depth 1 in closed form. With no stay moves and no vertex without a
// move, a cop-turn state (C, r) is won at depth 1 iff some cop of C has r as
// an out-neighbour (that cop steps onto the robber; the others move anywhere),
// and a robber-turn state iff every move of r lands on a cop of C. This is the
// push of level 0 (every capture state to all of its predecessors) done per
// row without touching any other row. Lanes = configurations.
// ---------------------------------------------------------------------------
template <int K, int W>
__global__ __launch_bounds__(kBlock) void k_init(GraphView gv, LevelArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint64_t lds_om[CR_MAX_N * W];
    __shared__ uint64_t stage[kWavesPerBlock][3][W][64];  // per wave: its rows' cm, c1, r1 words
    __shared__ uint64_t lds_im[CR_MAX_N * W];
    __shared__ uint64_t lnm[W];  // robbers without any move
    load_image(lds, gv.image, gv.image_bytes);
    const uint64_t* lom = stage_out_masks<W>(lds_om, a.out_mask, gv.n);
    const uint64_t* lim = stage_out_masks<W>(lds_im, a.in_mask, gv.n);
    if (threadIdx.x < W) {
        uint64_t m = 0;
        for (int r = 64 * threadIdx.x; r < min(gv.n, 64 * static_cast<int>(threadIdx.x) + 64); ++r) {
            bool none = true;
            for (int q = 0; q < W; ++q) none = none && lom[r * W + q] == 0;
            if (none) m |= 1ull << (r & 63);
        }
        lnm[threadIdx.x] = m;
    }
    __syncthreads();
    const uint32_t* binom = reinterpret_cast<const uint32_t*>(lds + gv.off_binom);
    const int lane = threadIdx.x & 63;
    const uint32_t wid = wave_id();
    const int n = gv.n;
    const uint32_t ngroups = (a.nbatch + kGroupWords - 1) / kGroupWords;
    for (uint32_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        unsigned long long caps = 0, cop1 = 0, rob1 = 0, rows0 = 0, rows1 = 0;
        if (a.dirty && threadIdx.x == 0) a.dirty[g] = 0;
        for (int j = 0; j < kWordsPerWave; ++j) {
            const uint32_t b = g * kGroupWords + wid * kWordsPerWave + j;
            if (b >= a.nbatch) break;
            const uint32_t row = b * 64 + lane;
            const bool live = row < gv.C;
            uint64_t cm[W], am[W], r1[W];
#pragma unroll
            for (int q = 0; q < W; ++q) cm[q] = am[q] = r1[q] = 0;
            uint32_t cops[K];
            if (live) {
                unrank<K>(row, gv, binom, cops);
#pragma unroll
                for (int i = 0; i < K; ++i) {
#pragma unroll
                    for (int q = 0; q < W; ++q) {
                        if ((cops[i] >> 6) == static_cast<uint32_t>(q)) cm[q] |= 1ull << (cops[i] & 63);
                        am[q] |= lom[cops[i] * W + q];
                    }
                }
            }
            // a robber-turn state is won at depth 1 iff every move of r lands on a
            // cop: r has a move onto some cop (candidates from the transposed
            // moves) and all its moves are onto cops; robbers without any move
            // have none onto a cop and are tested in a uniform pass
            uint64_t cand[W];
#pragma unroll
            for (int q = 0; q < W; ++q) cand[q] = lnm[q];
            if (live) {
#pragma unroll
                for (int i = 0; i < K; ++i)
#pragma unroll
                    for (int q = 0; q < W; ++q) cand[q] |= lim[cops[i] * W + q];
            }
#pragma unroll
            for (int w = 0; w < W; ++w) {
                uint64_t x = cand[w] & ~cm[w];
                while (x) {
                    const int rb = __builtin_ctzll(x);
                    x &= x - 1;
                    const uint64_t* om = lom + (64 * w + rb) * W;
                    uint64_t bad = 0;
#pragma unroll
                    for (int q = 0; q < W; ++q) bad |= om[q] & ~cm[q];
                    r1[w] |= static_cast<uint64_t>(bad == 0) << rb;
                }
            }
            bool has0 = false, has1 = false;
            uint32_t lc = 0, l0 = 0, l1 = 0;
            uint64_t c1[W];
#pragma unroll
            for (int q = 0; q < W; ++q) c1[q] = am[q] & ~cm[q];
            // table bytes of position rb: (cap 0x01 | depth-1 0x03 | unmarked 0xFE) per turn
            auto pos_pair = [&](uint64_t cw, uint64_t c1w, uint64_t r1w, int rb) -> uint32_t {
                const uint64_t bit = 1ull << rb;
                const bool cap = cw & bit;
                const uint32_t b0 = cap ? 0x01u : ((c1w & bit) ? 0x03u : kUnmarked);
                const uint32_t b1 = cap ? 0x01u : ((r1w & bit) ? 0x03u : kUnmarked);
                return b0 | (b1 << 8);
            };
            if (live) {  // the rows' bitsets by 16-byte stores where the rows allow
                const uint64_t o = static_cast<uint64_t>(row) * W;
                uint64_t v0[W], v1[W], z[W];
#pragma unroll
                for (int q = 0; q < W; ++q) {
                    v0[q] = cm[q] | c1[q];
                    v1[q] = cm[q] | r1[q];
                    z[q] = 0;
                    lc += __popcll(cm[q]);
                    l0 += __popcll(c1[q]);
                    l1 += __popcll(r1[q]);
                    has0 |= c1[q] != 0;
                    has1 |= r1[q] != 0;
                }
                store_row<W>(a.won0 + o, v0);
                store_row<W>(a.won1 + o, v1);
                store_row<W>(a.fr1s[1] + o, r1);  // depth 1, read by the cop step of level 1
                if (a.fr1s[0] != a.fr1s[1]) store_row<W>(a.fr1s[0] + o, z);
                store_row<W>(a.fr0 + o, c1);
                store_row<W>(a.nx0 + o, z);
            }
            if ((n & 7) == 0) {
                // the wave's rows are one contiguous span of 64 * 2n bytes: lanes
                // write consecutive 16-byte chunks (8 positions) of it, reading the
                // owning row's masks from the wave's LDS slot
#pragma unroll
                for (int q = 0; q < W; ++q) {
                    stage[wid][0][q][lane] = cm[q];
                    stage[wid][1][q][lane] = c1[q];
                    stage[wid][2][q][lane] = r1[q];
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const uint32_t rows_here = min(64u, gv.C - b * 64);
                const uint32_t cpr = static_cast<uint32_t>(n) / 8;  // chunks per row
                uint4* dst = reinterpret_cast<uint4*>(a.table + static_cast<uint64_t>(b) * 64 * n * 2);
                for (uint32_t i = lane; i < rows_here * cpr; i += 64) {
                    const uint32_t rl = i / cpr, r0 = (i - rl * cpr) * 8;
                    const int q = r0 >> 6, sh = r0 & 63;
                    const uint64_t cw = stage[wid][0][q][rl] >> sh, c1w = stage[wid][1][q][rl] >> sh,
                                   r1w = stage[wid][2][q][rl] >> sh;
                    uint32_t d[4];
#pragma unroll
                    for (int h = 0; h < 4; ++h)
                        d[h] = pos_pair(cw, c1w, r1w, 2 * h) | (pos_pair(cw, c1w, r1w, 2 * h + 1) << 16);
                    dst[i] = make_uint4(d[0], d[1], d[2], d[3]);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            } else if (live) {
                uint8_t* trow = a.table + static_cast<uint64_t>(row) * n * 2;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    const int rend = min(64, n - 64 * w);
                    uint8_t* tw = trow + 128 * w;
                    if ((n & 3) == 0) {  // rows start 8-byte aligned
                        for (int rb = 0; rb < rend; rb += 4) {
                            const uint64_t v = static_cast<uint64_t>(pos_pair(cm[w], c1[w], r1[w], rb)) |
                                               (static_cast<uint64_t>(pos_pair(cm[w], c1[w], r1[w], rb + 1)) << 16) |
                                               (static_cast<uint64_t>(pos_pair(cm[w], c1[w], r1[w], rb + 2)) << 32) |
                                               (static_cast<uint64_t>(pos_pair(cm[w], c1[w], r1[w], rb + 3)) << 48);
                            *reinterpret_cast<uint64_t*>(tw + 2 * rb) = v;
                        }
                    } else if ((n & 1) == 0) {  // 4-byte aligned
                        for (int rb = 0; rb < rend; rb += 2)
                            *reinterpret_cast<uint32_t*>(tw + 2 * rb) =
                                pos_pair(cm[w], c1[w], r1[w], rb) | (pos_pair(cm[w], c1[w], r1[w], rb + 1) << 16);
                    } else {
                        for (int rb = 0; rb < rend; ++rb)
                            *reinterpret_cast<uint16_t*>(tw + 2 * rb) =
                                static_cast<uint16_t>(pos_pair(cm[w], c1[w], r1[w], rb));
                    }
                }
            }
            const uint64_t f0 = __ballot(has0), f1 = __ballot(has1);
            if (lane == 0) {
                a.c0.flags[b] = f0;
                a.c1.flags[b] = f1;
            }
            rows0 += __popcll(f0);
            rows1 += __popcll(f1);
            caps += wave_sum(lc);
            cop1 += wave_sum(l0);
            rob1 += wave_sum(l1);
        }
        unsigned long long r0 = rows0, r1s = rows1;
        block_sum2(r0, r1s);
        block_sum2(caps, cop1);
        unsigned long long rb1 = rob1, unused = 0;
        block_sum2(rb1, unused);
        if (threadIdx.x == 0) {
            a.c0.gcnt[g] = static_cast<uint32_t>(r0);
            a.c1.gcnt[g] = static_cast<uint32_t>(r1s);
            a.gcaps[g] = caps;
            a.c0.gstates[g] = cop1;
            a.c1.gstates[g] = rb1;
        }
    }
}

// ---------------------------------------------------------------------------
// cop turn, push: frontier robber-turn rows -> predecessors' nx0. One
// wavefront per list row: lanes enumerate the team moves back along in-edges,
// rank them and OR the row's robber mask into nx0 of every predecessor.
// ---------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(kBlock) void k_cop_push(GraphView gv, LevelArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    if (a.ctl->done || a.ctl->pull) return;  // uniform: before any barrier
    const uint32_t cur = a.ctl->level & 1;
    const uint32_t count = a.ctl->cnt1[cur];
    const XcdSplit sp = xcd_split(count);
    // a workgroup whose waves have no frontier row skips the image copy (the
    // grid is sized for the largest frontier; small levels use few workgroups)
    if (sp.first - wave_id() >= sp.end) return;  // uniform: the workgroup's first wave
    load_image(lds, gv.image, gv.image_bytes);
    const uint32_t* rank_tab = reinterpret_cast<const uint32_t*>(lds);
    const uint32_t* binom = reinterpret_cast<const uint32_t*>(lds + gv.off_binom);
    const uint16_t* in_off = reinterpret_cast<const uint16_t*>(lds + gv.off_in_off);
    const uint8_t* in_nbr = lds + gv.off_in_nbr;
    const int lane = threadIdx.x & 63;
    const int W = gv.W;
    uint64_t* fr1 = a.fr1s[cur];
    const uint32_t* list = a.list1[cur];
    for (uint32_t it = sp.first; it < sp.end; it += sp.stride) {
        const uint32_t row = list[it];
        uint64_t m[4] = {0, 0, 0, 0};
        for (int w = 0; w < W; ++w) m[w] = fr1[static_cast<uint64_t>(row) * W + w];
        uint32_t cops[K], deg[K], base[K];
        unrank<K>(row, gv, binom, cops);
        uint32_t P = 1;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            base[i] = in_off[cops[i]];
            deg[i] = in_off[cops[i] + 1] - base[i];
            P *= deg[i];
        }
        // two predecessors per lane per iteration keep two filter loads in flight
        for (uint32_t t = lane; t < P; t += 128) {
            uint32_t tup[2][K];
            bool keep[2], hit[2] = {false, false};
            uint64_t pr[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const uint32_t tt = t + 64 * u;
                const bool in = tt < P;  // decode a valid index either way; drop it when out of range
                keep[u] = product_tuple<K>(in ? tt : t, cops, deg, base, in_nbr, tup[u]) && in;
                pr[u] = rank_sorted<K>(tup[u], rank_tab, gv.n, gv.ctot_m1);
            }
            // every filter load of the iteration before any atomic: the
            // atomics may alias won0 as far as the compiler knows, so loads
            // interleaved with them would wait one after another
            uint64_t wv[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
            if (a.filter) {
#pragma unroll
                for (int u = 0; u < 2; ++u)
                    if (keep[u]) {
                        const uint64_t* src = a.won0 + pr[u] * W;
#pragma unroll
                        for (int w = 0; w < 4; ++w)
                            if (w < W && m[w]) wv[u][w] = src[w];
                    }
            }
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                if (w >= W || m[w] == 0) continue;
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const uint64_t mw = keep[u] ? m[w] & ~wv[u][w] : 0ull;
                    if (mw) {
                        atomicOr(reinterpret_cast<unsigned long long*>(a.nx0 + pr[u] * W + w),
                                 static_cast<unsigned long long>(mw));
                        hit[u] = true;
                    }
                }
            }
            if (a.dirty) {  // k_cop_finalize reads only the groups marked here
#pragma unroll
                for (int u = 0; u < 2; ++u)
                    if (hit[u]) a.dirty[pr[u] >> 8] = 1;
            }
        }
        if (lane < W) fr1[static_cast<uint64_t>(row) * W + lane] = 0;
    }
}

// ---------------------------------------------------------------------------
// cop turn, factorised pull (dense levels). The cop-turn candidates of C are
// OR_{C' in T(C)} fr1[C'] with T(C) = { sort(n_1..n_K) : n_i in N(c_i) }
// (AuxGraph.h:194-254). That OR is a K-fold product of per-cop
// neighbourhoods, so it is taken one cop at a time: stage i turns
//   H_{i-1}[A'][B'] = OR over the first i-1 cops already moved
// into H_i[A][B], where A is the multiset of the i moved cops, B of the K-i
// others (H_0 = fr1, H_K = candidates of C), by
//   H_i[A][B] = OR_{n in N(max A)} H_{i-1}[A \ {max A}][B + {n}].
// H_i is symmetric within A and within B, so rows are indexed by the pair of
// multiset ranks (lexicographic, AuxGraph.h:117-177 order) and the stages
// read O(sum |H_i| * degree) rows instead of O(C * |T|) random ones; for a
// fixed A the reads for increasing B are increasing rows of H_{i-1}.
// ---------------------------------------------------------------------------
struct SepStage {
    const uint64_t* in;   // H_{i-1}, rows [cnt(i-1) * cnt(K-i+1)]
    const uint64_t* in_odd;  // stage 1: the frontier of odd depths (in holds the even ones)
    uint64_t* out;        // H_i, rows [cnt(i) * cnt(K-i)] (stage K: nx0)
    uint32_t cnt_b;       // cnt(K-i): B-multisets of H_i
    uint32_t cnt_b_in;    // cnt(K-i+1): B-multisets of H_{i-1}
    uint32_t cnt_a;       // cnt(i): A-multisets of H_i
    uint32_t cnt_a_in;    // cnt(i-1): A-multisets of H_{i-1}
    uint64_t rows;        // rows of H_i
    int32_t stage;        // i, 1..K
    uint32_t r;           // k_pull_tile: rows of each input block per tile
    uint32_t chunk;       // k_pull_tile: tiles a workgroup takes at a time
    int32_t fuse;         // stage K: do k_cop_finalize's work instead of writing nx0
    int32_t clear;        // stage 2: clear the frontier rows of fr1 (k_clear_frontier's work)
    // k_pull_pair (stages K-1 and K in one pass): tiles (A-range R, c), R-major
    const uint32_t* tstart;  // [nR + 1]: first tile of A-range R
    const uint32_t* cstart;  // [nR]: least max A over R (the first c of R's tiles)
    uint32_t nR;             // A-ranges of r consecutive A
};

// The stage's input rows: stage 1 reads the robber frontier of the level's depth.
__device__ __forceinline__ const uint64_t* stage_input(const SepStage& st, const Ctl* ctl) {
    return st.in_odd && (ctl->level & 1) ? st.in_odd : st.in;
}

constexpr int kPullMlp = 8;  // input rows whose loads a thread keeps in flight

// Lexicographic successor of a non-decreasing J-tuple over [0, n), in place
// (the order rank_j counts in); false after the last tuple.
template <int J>
__device__ __forceinline__ bool ms_next(uint32_t* t, uint32_t n) {
#pragma unroll
    for (int i = J - 1; i >= 0; --i) {
        if (t[i] + 1 < n) {
            const uint32_t v = t[i] + 1;
#pragma unroll
            for (int j = i; j < J; ++j) t[j] = v;
            return true;
        }
    }
    return false;
}

// Stage I of the pull (J = K - I cops not moved yet), one row of H_I per
// lane. The intermediate tables H_1..H_{K-1} are laid out with B major: row
// rank(B) * cnt(I) + rank(A) (H_0 and H_K are indexed by configuration). Rows
// are taken 64 at a time; consecutive rows differ only in A's last element as
// long as it does not wrap, so a wave unranks the first row of its window once
// and walks these runs, handing each lane its (A, B) pair. A neighbour's
// input row then costs J compares and one binomial lookup: with T(p, x) the
// rank term of value x at position p of a (J+1)-multiset,
//   rank(B + v) = rank0 - sum_{B_i < v} T(i, B_i) - T(q, v) - sum_{B_i >= v} T(i + 1, B_i)
// where q counts the B_i below v (rank_j's formula with v merged in).
template <int I, int J, int W>
__global__ __launch_bounds__(kBlock) void k_sep_stage(GraphView gv, LevelArgs a, SepStage st) {
    extern __shared__ __align__(16) uint8_t lds[];
    if (a.ctl->done || !a.ctl->pull) return;  // uniform: before any barrier
    const uint64_t* sin = stage_input(st, a.ctl);
    load_image(lds, gv.image, gv.image_bytes);
    const uint32_t* binom = reinterpret_cast<const uint32_t*>(lds + gv.off_binom);
    const uint16_t* out_off = reinterpret_cast<const uint16_t*>(lds + gv.off_out_off);
    const uint8_t* out_nbr = lds + gv.off_out_nbr;
    const uint32_t n = gv.n;
    const uint32_t Mp = gv.M + 1;  // row stride of the binomial table
    const int lane = threadIdx.x & 63;
    uint64_t full[W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const int bits = min(64, max(0, static_cast<int>(n) - 64 * q));
        full[q] = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    }
    const uint32_t nwin = static_cast<uint32_t>((st.rows + 63) / 64);
    const XcdSplit xs = xcd_split(nwin);
    for (uint32_t win = xs.first; win < xs.end; win += xs.stride) {
        const uint32_t o0 = win * 64;
        uint32_t A[I], B[J > 0 ? J : 1];
        {  // runs over A's last element; A wraps into the next B
            uint32_t ca[I], cb[J > 0 ? J : 1];
            uint32_t ra0 = o0;
            if constexpr (J > 0) {
                const uint32_t rb0 = o0 / st.cnt_a;
                ra0 = o0 - rb0 * st.cnt_a;
                unrank_j(rb0, J, n, binom, gv.M, cb);
            }
            unrank_j(ra0, I, n, binom, gv.M, ca);
            for (int s = 0;;) {
                const uint32_t x0 = ca[I - 1];
                const int len = static_cast<int>(n - x0);
                if (lane >= s && lane < s + len) {
#pragma unroll
                    for (int t = 0; t < I; ++t) A[t] = ca[t];
                    A[I - 1] = x0 + static_cast<uint32_t>(lane - s);
#pragma unroll
                    for (int t = 0; t < J; ++t) B[t] = cb[t];
                }
                s += len;
                if (s >= 64) break;
                ca[I - 1] = n - 1;
                if (!ms_next<I>(ca, n)) {
#pragma unroll
                    for (int t = 0; t < I; ++t) ca[t] = 0;
                    if constexpr (J > 0) ms_next<J>(cb, n);
                }
            }
        }
        const uint64_t o = static_cast<uint64_t>(o0) + lane;
        if (o >= st.rows) continue;
        const uint32_t c = A[I - 1];  // the cop moved by this stage: A's largest element
        // input row of (A without c, B + v): row_in + rank(B + v) * in_stride
        uint64_t row_in = 0;
        const uint64_t in_stride = st.cnt_a_in;
        if constexpr (I > 1) row_in = rank_j(A, I - 1, n, binom, gv.M);
        // rank terms of B's elements at their own and at the next position
        constexpr int J1 = J + 1;
        const uint32_t M1 = n + J;  // M of (J+1)-multisets
        uint32_t U[J > 0 ? J : 1], S[J > 0 ? J : 1];
#pragma unroll
        for (int t = 0; t < J; ++t) {
            U[t] = binom[(J1 - t) * Mp + (M1 - 1 - B[t] - t)];
            S[t] = binom[(J - t) * Mp + (M1 - 2 - B[t] - t)];
        }
        const uint32_t rank0 = binom[J1 * Mp + M1] - 1;
        uint64_t acc[W];
#pragma unroll
        for (int q = 0; q < W; ++q) acc[q] = 0;
        const int beg = out_off[c], end = out_off[c + 1];
        for (int e0 = beg; e0 < end; e0 += kPullMlp) {
            uint32_t idx[kPullMlp];
            const int m = min(kPullMlp, end - e0);
#pragma unroll
            for (int u = 0; u < kPullMlp; ++u) {
                idx[u] = 0;
                if (u >= m) continue;
                const uint32_t v = out_nbr[e0 + u];
                if constexpr (J == 0) {
                    idx[u] = v;
                } else {
                    uint32_t sum = 0, q = 0;
#pragma unroll
                    for (int t = 0; t < J; ++t) {
                        const bool lt = B[t] < v;
                        sum += lt ? U[t] : S[t];
                        q += lt ? 1u : 0u;
                    }
                    idx[u] = rank0 - sum - binom[(J1 - q) * Mp + (M1 - 1 - v - q)];
                }
            }
            uint64_t got[kPullMlp][W];
#pragma unroll
            for (int u = 0; u < kPullMlp; ++u) {
                if (u < m) {
                    load_row<W>(sin + (row_in + idx[u] * in_stride) * W, got[u]);
                } else {
#pragma unroll
                    for (int q = 0; q < W; ++q) got[u][q] = 0;
                }
            }
#pragma unroll
            for (int u = 0; u < kPullMlp; ++u)
#pragma unroll
                for (int q = 0; q < W; ++q) acc[q] |= got[u][q];
        }
        if constexpr (J == 0) {  // candidates of configuration o: only states not won yet
            uint64_t w0[W];
            load_row<W>(a.won0 + o * W, w0);
#pragma unroll
            for (int q = 0; q < W; ++q) acc[q] &= ~w0[q] & full[q];
        }
        store_row<W>(st.out + o * W, acc);
    }
}

// The row flags and group counts of 64 consecutive configurations starting
// at `base` (one wave; bit i of `bits` = configuration base + i has new
// states, `pop` = this lane's new states), added into zeroed words with at
// most two atomics each: 64 rows touch at most two flag words and two groups
// of kGroupWords * 64 rows.
__device__ __forceinline__ void fuse_flags(const LevelArgs& a, uint64_t base, uint64_t bits, uint32_t pop) {
    const int lane = threadIdx.x & 63;
    constexpr uint32_t kGroupRows = kGroupWords * 64;
    const uint64_t g0 = base / kGroupRows;
    const uint32_t split = static_cast<uint32_t>(min<uint64_t>(64, (g0 + 1) * kGroupRows - base));  // lanes in g0
    const bool first = static_cast<uint32_t>(lane) < split;
    unsigned long long s0 = first ? pop : 0u, s1 = first ? 0u : pop;
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    if (lane != 0 || bits == 0) return;
    const uint64_t w = base / 64;
    const uint32_t sh = static_cast<uint32_t>(base % 64);
    atomicOr(reinterpret_cast<unsigned long long*>(a.c0.flags + w), static_cast<unsigned long long>(bits << sh));
    if (sh && (bits >> (64 - sh)))
        atomicOr(reinterpret_cast<unsigned long long*>(a.c0.flags + w + 1),
                 static_cast<unsigned long long>(bits >> (64 - sh)));
    const uint64_t lo = split >= 64 ? bits : (bits & ((1ull << split) - 1));
    const uint32_t c0 = static_cast<uint32_t>(__popcll(lo)), c1 = static_cast<uint32_t>(__popcll(bits)) - c0;
    if (c0) {
        atomicAdd(a.c0.gcnt + g0, c0);
        atomicAdd(a.c0.gstates + g0, s0);
    }
    if (c1) {
        atomicAdd(a.c0.gcnt + g0 + 1, c1);
        atomicAdd(a.c0.gstates + g0 + 1, s1);
    }
}

// Zeroes the cop-turn flags and group counts before a fused last pull stage.
__global__ __launch_bounds__(kBlock) void k_zero_groups(LevelArgs a) {
    if (a.ctl->done || !a.ctl->pull) return;
    const uint32_t ngroups = (a.nbatch + kGroupWords - 1) / kGroupWords;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.nbatch; i += gridDim.x * blockDim.x)
        a.c0.flags[i] = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ngroups; i += gridDim.x * blockDim.x) {
        a.c0.gcnt[i] = 0;
        a.c0.gstates[i] = 0;
    }
}

// Stage I of the pull by LDS tiles. With B major, the input rows of a fixed B
// are n contiguous blocks of H_{I-1} (block B + v holds the rows of every A'),
// so a tile (B, A'-range [ra0, ra0 + r)) loads rows ra0.. of all n blocks
// into LDS with coalesced reads, then every output row (A' + c, B) of the tile,
// c >= max A', ORs its neighbours' rows out of LDS:
//   H_I[A' + c][B] = OR_{v in N(c)} tile[v][A'].
// For fixed A' the outputs of increasing c are consecutive rows and the
// output runs of consecutive A' are adjacent, so consecutive tiles write
// consecutive ranges of H_I. Each input row is loaded once per B that it
// extends (at most J + 1 times).
// A workgroup walks a contiguous chunk of tiles: B, A' and the output row
// advance incrementally (one unrank per chunk), and the loads of the next
// tile are held in registers while the current tile is computed from its
// LDS buffer (n * r * W words after the graph image): the bytes in flight per
// CU are bounded by registers, not by a second LDS buffer.
constexpr int kTileMaxRows = 16;
constexpr int kTileBlock = 256;  // threads of a tile workgroup
constexpr int kTileLoads = 16;   // input words a thread stages per tile: n * r * W <= kTileLoads * kTileBlock
constexpr uint32_t kLdsBytes = 160 * 1024;  // LDS per workgroup on gfx950


template <int I, int J, int W>
__global__ __launch_bounds__(kTileBlock) void k_pull_tile(GraphView gv, LevelArgs a, SepStage st) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t t_lo[2][kTileMaxRows], t_pre[2][kTileMaxRows + 1];
    __shared__ unsigned long long t_out[2];
    if (a.ctl->done || !a.ctl->pull) return;  // uniform: before any barrier
    const uint64_t* sin = stage_input(st, a.ctl);
    load_image(lds, gv.image, gv.image_bytes);
    const uint32_t* binom = reinterpret_cast<const uint32_t*>(lds + gv.off_binom);
    const uint16_t* out_off = reinterpret_cast<const uint16_t*>(lds + gv.off_out_off);
    const uint8_t* out_nbr = lds + gv.off_out_nbr;
    const uint32_t n = gv.n;
    const uint32_t Mp = gv.M + 1;
    const uint32_t r = st.r, rW = r * W, words = n * rW;
    uint64_t* buf = reinterpret_cast<uint64_t*>(lds + ((gv.image_bytes + 15) & ~15u));
    uint64_t full[W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const int bits = min(64, max(0, static_cast<int>(n) - 64 * q));
        full[q] = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    }
    const uint32_t cnt_in = st.cnt_a_in;        // rows per block: cnt(I - 1)
    const uint32_t ntr = (cnt_in + r - 1) / r;  // A'-ranges per B
    const uint32_t cnt_b = st.cnt_b;
    const uint64_t ntiles = static_cast<uint64_t>(cnt_b) * ntr;
    // Tiles are ordered A'-range major (tile = tr * cnt_b + rb) and handed out
    // in chunks of st.chunk, chunk c to workgroup c mod gridDim: the resident
    // workgroups work on neighbouring tiles, which share their A'-range, so
    // the J + 1 loads of an input row fall close together in time and the
    // repeats find it in the caches.
    const uint64_t nchunks = (ntiles + st.chunk - 1) / st.chunk;
    uint64_t ch = blockIdx.x;
    if (st.clear) {  // stage 2: stage 1 has read the frontier; clear its rows (k_clear_frontier's work)
        const uint32_t cur = a.ctl->level & 1;
        const uint64_t cw = static_cast<uint64_t>(a.ctl->cnt1[cur]) * W;
        const uint32_t* list = a.list1[cur];
        for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < cw;
             i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
            a.fr1s[cur][static_cast<uint64_t>(list[i / W]) * W + i % W] = 0;
    }
    if (ch >= nchunks) return;  // uniform per workgroup

    // walk state (uniform): the tile, its B with rank terms and first A' row;
    // thread 0 also keeps the A' of the range and the output row
    constexpr int J1 = J + 1;
    const uint32_t M1 = n + J;  // M of (J+1)-multisets
    const uint32_t rank0 = binom[J1 * Mp + M1] - 1;
    uint64_t tix = 0, ch_end = 0;
    uint32_t rb = 0, ra0 = 0;
    uint32_t B[J > 0 ? J : 1], U[J > 0 ? J : 1], S[J > 0 ? J : 1];
    uint32_t Ap[I > 1 ? I - 1 : 1], Ap0[I > 1 ? I - 1 : 1];  // thread 0: A' walk, first A' of the range
    unsigned long long base_tr = 0;  // thread 0: output row of (A'-range first A' + its max) in block 0
    uint32_t last_total = 0;         // thread 0: outputs per tile of the current A'-range
    auto terms = [&]() {
#pragma unroll
        for (int t = 0; t < J; ++t) {
            U[t] = binom[(J1 - t) * Mp + (M1 - 1 - B[t] - t)];
            S[t] = binom[(J - t) * Mp + (M1 - 2 - B[t] - t)];
        }
    };
    auto start_at = [&](uint64_t t) {  // walk state from scratch at tile t
        tix = t;
        ch_end = min(ntiles, (t / st.chunk + 1) * st.chunk);
        const uint32_t tr = static_cast<uint32_t>(t / cnt_b);
        rb = static_cast<uint32_t>(t - static_cast<uint64_t>(tr) * cnt_b);
        ra0 = tr * r;
        if constexpr (J > 0) unrank_j(rb, J, n, binom, gv.M, B);
        terms();
        if (threadIdx.x == 0) {
            uint32_t Af[I];
            if constexpr (I > 1) {
                unrank_j(ra0, I - 1, n, binom, gv.M, Ap0);
#pragma unroll
                for (int q = 0; q < I - 1; ++q) Af[q] = Ap0[q];
                Af[I - 1] = Ap0[I - 2];
            } else {
                Af[0] = 0;
            }
            base_tr = rank_j(Af, I, n, binom, gv.M);
        }
    };
    // the next tile of this workgroup's sequence; false at the end
    auto advance = [&]() -> bool {
        if (tix + 1 < ch_end) {
            ++tix;
            if (++rb == cnt_b) {  // next A'-range, first B
                rb = 0;
                ra0 += r;
                if constexpr (J > 0) {
#pragma unroll
                    for (int q = 0; q < J; ++q) B[q] = 0;
                    terms();
                }
                if (threadIdx.x == 0) {
                    if constexpr (I > 1) {
#pragma unroll
                        for (int q = 0; q < I - 1; ++q) Ap0[q] = Ap[q];  // walked past the previous range
                    }
                    base_tr += last_total;
                }
            } else if constexpr (J > 0) {
                ms_next<J>(B, n);
                terms();
            }
            return true;
        }
        ch += gridDim.x;
        if (ch >= nchunks) return false;
        start_at(ch * st.chunk);
        return true;
    };
    // thread 0: the tile's rows t -> last element of A'_t and output offsets
    auto tile_table = [&](int b, uint32_t rc) {
        if constexpr (I > 1) {
#pragma unroll
            for (int q = 0; q < I - 1; ++q) Ap[q] = Ap0[q];
        }
        uint32_t pre = 0;
        for (uint32_t t = 0; t < rc; ++t) {
            const uint32_t lo = I > 1 ? Ap[I > 1 ? I - 2 : 0] : 0u;
            t_lo[b][t] = lo;
            t_pre[b][t] = pre;
            pre += n - lo;
            if constexpr (I > 1) ms_next<I - 1>(Ap, n);
        }
        t_pre[b][rc] = pre;
        last_total = pre;
        t_out[b] = static_cast<unsigned long long>(rb) * st.cnt_a + base_tr;
    };
    // the tile's input words (rows ra0.. of the n blocks) into registers
    uint64_t stage_w[kTileLoads];
    auto load_tile = [&](uint32_t rc) {
#pragma unroll
        for (int m = 0; m < kTileLoads; ++m) {
            const uint32_t i = threadIdx.x + m * kTileBlock;
            stage_w[m] = 0;
            if (i < words) {
                const uint32_t v = i / rW, w = i - v * rW;
                if (w < rc * W) {
                    uint32_t idx = v;
                    if constexpr (J > 0) {
                        uint32_t sum = 0, q = 0;
#pragma unroll
                        for (int t = 0; t < J; ++t) {
                            const bool lt = B[t] < v;
                            sum += lt ? U[t] : S[t];
                            q += lt ? 1u : 0u;
                        }
                        idx = rank0 - sum - binom[(J1 - q) * Mp + (M1 - 1 - v - q)];
                    }
                    stage_w[m] = sin[(static_cast<uint64_t>(idx) * cnt_in + ra0) * W + w];
                }
            }
        }
    };
    auto store_tile = [&]() {
#pragma unroll
        for (int m = 0; m < kTileLoads; ++m) {
            const uint32_t i = threadIdx.x + m * kTileBlock;
            if (i < words) buf[i] = stage_w[m];
        }
    };

    start_at(ch * st.chunk);
    uint32_t rc = min(r, cnt_in - ra0);
    if (threadIdx.x == 0) tile_table(0, rc);
    load_tile(rc);
    store_tile();
    for (uint32_t k = 0;; ++k) {
        const int cur = k & 1, nxt = cur ^ 1;
        __syncthreads();  // tile k is in buf
        const uint32_t rc_cur = rc;
        const bool more = advance();  // uniform
        if (more) {  // start the next tile's loads
            rc = min(r, cnt_in - ra0);
            if (threadIdx.x == 0) tile_table(nxt, rc);
            load_tile(rc);
        }
        const uint64_t* tb = buf;
        const uint32_t total = t_pre[cur][rc_cur];
        const uint64_t o0 = t_out[cur];
        for (uint32_t j0 = 0; j0 < total; j0 += blockDim.x) {  // every lane runs every round (ballots)
            const uint32_t j = j0 + threadIdx.x;
            const bool valid = j < total;
            uint64_t acc[W];
#pragma unroll
            for (int q = 0; q < W; ++q) acc[q] = 0;
            if (valid) {
                uint32_t t = 0;
                while (t + 1 < rc_cur && t_pre[cur][t + 1] <= j) ++t;
                const uint32_t c = t_lo[cur][t] + (j - t_pre[cur][t]);
                const int beg = out_off[c], end = out_off[c + 1];
                for (int e = beg; e < end; ++e) {
                    const uint64_t* src = tb + out_nbr[e] * rW + t * W;
#pragma unroll
                    for (int q = 0; q < W; ++q) acc[q] |= src[q];
                }
            }
            const uint64_t o = o0 + j;
            if constexpr (J == 0) {  // candidates of configuration o: only states not won yet
                uint64_t w0[W];
                uint32_t pop = 0;
                if (valid) {
                    load_row<W>(a.won0 + o * W, w0);
#pragma unroll
                    for (int q = 0; q < W; ++q) {
                        acc[q] &= ~w0[q] & full[q];
                        pop += __popcll(acc[q]);
                    }
                }
                if (st.fuse) {
                    // k_cop_finalize's work for a pulled level (the level's
                    // robber step has already read fr0): rows with new states
                    // write them to fr0 (st.out; fr0 is read only for the rows
                    // listed), and the row flags and group counts (zeroed
                    // before the stage) take this wave's rows. The next
                    // level's robber step adds them to won0.
                    if (pop) store_row<W>(st.out + o * W, acc);
                    fuse_flags(a, o0 + j0 + (threadIdx.x & ~63u), __ballot(pop != 0), pop);
                } else if (valid) {
                    store_row<W>(st.out + o * W, acc);
                }
            } else if (valid) {
                store_row<W>(st.out + o * W, acc);
            }
        }
        if (!more) break;
        __syncthreads();  // every thread is done with tile k: the next one replaces it
        store_tile();
    }
}

// Stage 1 of the pull (A' empty, J = K - 1 >= 1) by tiles of r consecutive
// B: H_1[B][c] = OR_{v in N(c)} fr1[B + v]. A tile loads the rows B_t + v
// (t < r, every v) into LDS as tile[v][t]. Consecutive B differ only in
// their last element while it does not wrap, so for every v <= max B the
// rows B_t + v of consecutive t are adjacent: the tile's reads come in runs
// of r rows instead of single rows. Its outputs, rows rank(B_t) * n + c, are
// r * n consecutive rows of H_1. Thread x always loads word x mod W of B_t,
// t = (x mod rW) / W, for rows v = x / rW + m * (blockDim / rW), so each
// thread keeps one B_t and its rank terms. Same chunked walk and register
// staging as k_pull_tile, with kTile1Block threads: a run of r rows starts at
// any 32-byte offset, so longer runs waste less of the 64-byte sectors they
// touch, and their loads need the larger workgroup. At W = 4, r = 9 (a 288-byte
// LDS row) measured best: r = 8 gives 256-byte LDS rows, which start on the
// same banks, and the lanes gather rows of random neighbours.
constexpr int kTile1Block = 512;

template <int J, int W>
__global__ __launch_bounds__(kTile1Block) void k_pull_tile1(GraphView gv, LevelArgs a, SepStage st) {
    static_assert(J > 0, "stage 1 of a single-cop pull is its last stage");
    extern __shared__ __align__(16) uint8_t lds[];
    if (a.ctl->done || !a.ctl->pull) return;  // uniform: before any barrier
    const uint64_t* sin = stage_input(st, a.ctl);
    load_image(lds, gv.image, gv.image_bytes);
    const uint32_t* binom = reinterpret_cast<const uint32_t*>(lds + gv.off_binom);
    const uint16_t* out_off = reinterpret_cast<const uint16_t*>(lds + gv.off_out_off);
    const uint8_t* out_nbr = lds + gv.off_out_nbr;
    const uint32_t n = gv.n;
    const uint32_t Mp = gv.M + 1;
    const uint32_t r = st.r, rW = r * W;
    const uint32_t per = kTile1Block / rW;  // input rows v that one pass of the workgroup loads
    const bool loader = threadIdx.x < per * rW;
    const uint32_t tq = threadIdx.x % rW, my_t = tq / W, my_w = tq % W, v0 = threadIdx.x / rW;
    uint64_t* buf = reinterpret_cast<uint64_t*>(lds + ((gv.image_bytes + 15) & ~15u));
    const uint32_t cnt_b = st.cnt_b;
    const uint64_t ntiles = (cnt_b + r - 1) / r;
    const uint64_t nchunks = (ntiles + st.chunk - 1) / st.chunk;
    uint64_t ch = blockIdx.x;
    if (ch >= nchunks) return;  // uniform per workgroup

    constexpr int J1 = J + 1;
    const uint32_t M1 = n + J;  // M of (J+1)-multisets
    const uint32_t rank0 = binom[J1 * Mp + M1] - 1;
    uint64_t tix = 0, ch_end = 0;
    uint32_t rb0 = 0;              // rank of the tile's first B
    uint32_t B[J];                 // the tile's first B
    uint32_t Bt[J], U[J], S[J];    // this thread's B_t and its rank terms
    bool mine = false;             // B_t exists
    auto terms = [&]() {
#pragma unroll
        for (int q = 0; q < J; ++q) Bt[q] = B[q];
        mine = rb0 + my_t < cnt_b;
        if (mine)
            for (uint32_t t = 0; t < my_t; ++t) ms_next<J>(Bt, n);
#pragma unroll
        for (int t = 0; t < J; ++t) {
            U[t] = binom[(J1 - t) * Mp + (M1 - 1 - Bt[t] - t)];
            S[t] = binom[(J - t) * Mp + (M1 - 2 - Bt[t] - t)];
        }
    };
    auto start_at = [&](uint64_t t) {  // walk state from scratch at tile t
        tix = t;
        ch_end = min(ntiles, (t / st.chunk + 1) * st.chunk);
        rb0 = static_cast<uint32_t>(t * r);
        unrank_j(rb0, J, n, binom, gv.M, B);
        terms();
    };
    auto advance = [&]() -> bool {  // the next tile of this workgroup's sequence; false at the end
        if (tix + 1 < ch_end) {
            ++tix;
            for (uint32_t t = 0; t < r; ++t) ms_next<J>(B, n);
            rb0 += r;
            terms();
            return true;
        }
        ch += gridDim.x;
        if (ch >= nchunks) return false;
        start_at(ch * st.chunk);
        return true;
    };
    uint64_t stage_w[kTileLoads];
    auto load_tile = [&]() {  // rows B_t + v of this thread's B_t, word my_w
#pragma unroll
        for (int m = 0; m < kTileLoads; ++m) {
            const uint32_t v = v0 + m * per;
            stage_w[m] = 0;
            if (loader && mine && v < n) {
                uint32_t sum = 0, q = 0;
#pragma unroll
                for (int t = 0; t < J; ++t) {
                    const bool lt = Bt[t] < v;
                    sum += lt ? U[t] : S[t];
                    q += lt ? 1u : 0u;
                }
                const uint32_t idx = rank0 - sum - binom[(J1 - q) * Mp + (M1 - 1 - v - q)];
                stage_w[m] = sin[static_cast<uint64_t>(idx) * W + my_w];
            }
        }
    };
    auto store_tile = [&]() {
#pragma unroll
        for (int m = 0; m < kTileLoads; ++m) {
            const uint32_t v = v0 + m * per;
            if (loader && v < n) buf[v * rW + tq] = stage_w[m];
        }
    };

    start_at(ch * st.chunk);
    load_tile();
    store_tile();
    for (;;) {
        __syncthreads();  // the current tile is in buf
        const uint32_t rc = min(r, cnt_b - rb0);
        const uint64_t o0 = static_cast<uint64_t>(rb0) * n;
        const bool more = advance();  // uniform
        if (more) load_tile();        // the next tile's loads stay in flight during the compute
        const uint32_t total = rc * n;
        for (uint32_t j = threadIdx.x; j < total; j += blockDim.x) {
            const uint32_t t = j / n, c = j - t * n;
            uint64_t acc[W];
#pragma unroll
            for (int q = 0; q < W; ++q) acc[q] = 0;
            const int beg = out_off[c], end = out_off[c + 1];
            for (int e = beg; e < end; ++e) {
                const uint64_t* src = buf + out_nbr[e] * rW + t * W;
#pragma unroll
                for (int q = 0; q < W; ++q) acc[q] |= src[q];
            }
            store_row<W>(st.out + (o0 + j) * W, acc);
        }
        if (!more) break;
        __syncthreads();  // every thread is done with the current tile: the next one replaces it
        store_tile();
    }
}

// Stages K-1 and K of the pull in one pass (K >= 3). With I2 = K - 2 cops in
// A and the pair B2 = {b, m}:
//   H_{K-1}[A + c][{b}] = OR_{m in N(c)} H_{K-2}[A][{b, m}]      (c >= max A)
//   H_K[A + c + a]      = OR_{b in N(a)} H_{K-1}[A + c][{b}]       (a >= c)
// For a fixed A and c the first line is only n rows, so a tile (A-range R of
// r consecutive A, one c) builds them in LDS (Y[t][b], phase 1) and ORs the
// outputs out of LDS (phase 2): H_{K-1} is never written to memory. The
// outputs of a tile are, per A_t, the consecutive configurations
// A_t + c + c .. A_t + c + (n - 1). Phase 1 reads, for each m in N(c) and
// every b, the rows of R in block {b, m} of H_{K-2} (B major: r consecutive
// rows). Tiles are numbered R-major, c-minor, and split into 8 contiguous
// ranges, one per XCD (workgroup w runs on XCD w mod 8): the workgroups of
// an XCD work on neighbouring tiles, which share R, so the rows of R's blocks
// (n (n + 1) / 2 * r rows) are re-read from that XCD's L2.
constexpr int kPairBlock = 256;
constexpr int kPairMaxRows = 8;
constexpr int kPairItems = 8;   // phase-1 items per thread: r * n * W / (2 or 1) <= kPairItems * kPairBlock
constexpr int kPairUnroll = 2;  // neighbours whose loads a phase-1 step issues together

template <int I2, int W>
__global__ __launch_bounds__(kPairBlock) void k_pull_pair(GraphView gv, LevelArgs a, SepStage st) {
    static_assert(I2 >= 1, "k_pull_pair takes K >= 3");
    constexpr int K = I2 + 2;
    constexpr int WH = (W % 2 == 0) ? 2 : 1;  // words per phase-1 item (16-byte loads where rows allow)
    constexpr int NI = W / WH;                // items per (t, b)
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ unsigned long long s_base[kPairMaxRows];  // configuration of A_t + c + c, ~0 if max A_t > c
    if (a.ctl->done || !a.ctl->pull) return;  // uniform: before any barrier
    const uint64_t* sin = st.in;
    load_image(lds, gv.image, gv.image_bytes);
    const uint32_t* binom = reinterpret_cast<const uint32_t*>(lds + gv.off_binom);
    const uint16_t* out_off = reinterpret_cast<const uint16_t*>(lds + gv.off_out_off);
    const uint8_t* out_nbr = lds + gv.off_out_nbr;
    const uint32_t n = gv.n;
    const uint32_t r = st.r;
    uint64_t* Y = reinterpret_cast<uint64_t*>(lds + ((gv.image_bytes + 15) & ~15u));
    uint64_t full[W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const int bits = min(64, max(0, static_cast<int>(n) - 64 * q));
        full[q] = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    }
    if (st.clear) {  // K = 3: this is stage 2, after stage 1 has read the frontier (k_clear_frontier's work)
        const uint32_t cur = a.ctl->level & 1;
        const uint64_t cw = static_cast<uint64_t>(a.ctl->cnt1[cur]) * W;
        const uint32_t* list = a.list1[cur];
        for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < cw;
             i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
            a.fr1s[cur][static_cast<uint64_t>(list[i / W]) * W + i % W] = 0;
    }
    const uint32_t cnt_a = st.cnt_a_in;  // A-multisets of H_{K-2}: rows per block
    const uint32_t T = st.tstart[st.nR];
    const uint32_t xcd = blockIdx.x % kXcds, per_xcd = gridDim.x / kXcds, l = blockIdx.x / kXcds;
    const uint32_t t_lo = static_cast<uint32_t>(static_cast<uint64_t>(T) * xcd / kXcds);
    const uint32_t t_hi = static_cast<uint32_t>(static_cast<uint64_t>(T) * (xcd + 1) / kXcds);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t R = 0;
    for (uint32_t tile = t_lo + l; tile < t_hi; tile += per_xcd) {
        // the tile's A-range: the last R with tstart[R] <= tile (tiles only move forward)
        for (uint32_t hi = st.nR - 1; R < hi;) {
            const uint32_t mid = (R + hi + 1) >> 1;
            if (st.tstart[mid] <= tile) R = mid; else hi = mid - 1;
        }
        const uint32_t c = st.cstart[R] + (tile - st.tstart[R]);
        const uint32_t ra0 = R * r, rc = min(r, cnt_a - ra0);
        if (threadIdx.x < rc) {
            uint32_t A[K];
            unrank_j(ra0 + threadIdx.x, I2, n, binom, gv.M, A);
            unsigned long long base = ~0ull;
            if (A[I2 - 1] <= c) {
                A[I2] = c;
                A[I2 + 1] = c;
                base = rank_j(A, K, n, binom, gv.M);
            }
            s_base[threadIdx.x] = base;
        }
        __syncthreads();
        // phase 1: Y[t][b] = OR_{m in N(c)} H_{K-2}[A_t][{b, m}]. Each thread
        // keeps up to kPairItems (t, b, word pair) items and, per neighbour m,
        // issues the loads of all of them before ORing (kPairItems loads in flight).
        const int beg = out_off[c], end = out_off[c + 1];
        const uint32_t items = rc * n * NI;
        uint64_t acc[kPairItems][WH];
        uint32_t ib[kPairItems];   // b of the item, n if the thread has no such item
        uint64_t ioff[kPairItems]; // word offset of the item's row within a block of H_{K-2}
#pragma unroll
        for (int it = 0; it < kPairItems; ++it) {
            const uint32_t i = threadIdx.x + it * kPairBlock;
            const uint32_t q = i % NI, t = (i / NI) % rc, b = i / (NI * rc);
            const bool ok = i < items && s_base[t] != ~0ull;
            ib[it] = ok ? b : n;
            ioff[it] = static_cast<uint64_t>(ra0 + t) * W + q * WH;
#pragma unroll
            for (int h = 0; h < WH; ++h) acc[it][h] = 0;
        }
        // kPairUnroll neighbours per step (a step past the last neighbour repeats
        // it: ORing a row twice changes nothing), so that their loads are in
        // flight together
        for (int e0 = beg; e0 < end; e0 += kPairUnroll) {
#pragma unroll
            for (int u = 0; u < kPairUnroll; ++u) {
                const uint32_t m = out_nbr[min(e0 + u, end - 1)];
#pragma unroll
                for (int it = 0; it < kPairItems; ++it) {
                    const uint32_t b = ib[it];
                    if (b < n) {
                        const uint32_t x = min(b, m), y = max(b, m);
                        const uint32_t pr = x * n - (x * (x - 1)) / 2 + (y - x);  // rank of the pair {x <= y}
                        const uint64_t* src = sin + static_cast<uint64_t>(pr) * cnt_a * W + ioff[it];
                        if constexpr (WH == 2) {
                            const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(src);
                            acc[it][0] |= v.x;
                            acc[it][1] |= v.y;
                        } else {
                            acc[it][0] |= src[0];
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int it = 0; it < kPairItems; ++it) {
            if (ib[it] < n) {
                const uint32_t i = threadIdx.x + it * kPairBlock;
                const uint32_t q = i % NI, t = (i / NI) % rc;
                uint64_t* dst = Y + (static_cast<uint64_t>(t) * n + ib[it]) * W + q * WH;
                if constexpr (WH == 2) {
                    *reinterpret_cast<ulonglong2*>(dst) = make_ulonglong2(acc[it][0], acc[it][1]);
                } else {
                    dst[0] = acc[it][0];
                }
            }
        }
        __syncthreads();
        // phase 2: one wave per A_t, lanes over consecutive outputs a = c + j
        for (uint32_t t = wave; t < rc; t += kPairBlock / 64) {
            const unsigned long long o0 = s_base[t];
            if (o0 == ~0ull) continue;  // uniform per wave
            const uint64_t* yt = Y + static_cast<uint64_t>(t) * n * W;
            for (uint32_t j0 = 0; j0 < n - c; j0 += 64) {
                const uint32_t av = c + j0 + lane;
                const bool valid = av < n;
                uint64_t acc[W];
#pragma unroll
                for (int q = 0; q < W; ++q) acc[q] = 0;
                uint32_t pop = 0;
                const uint64_t o = o0 + j0 + lane;
                if (valid) {
                    uint64_t w0[W];
                    load_row<W>(a.won0 + o * W, w0);  // in flight during the ORs
                    const int b0 = out_off[av], b1 = out_off[av + 1];
                    for (int e = b0; e < b1; ++e) {
                        uint64_t v[W];
                        load_row<W>(yt + static_cast<uint32_t>(out_nbr[e]) * W, v);
#pragma unroll
                        for (int q = 0; q < W; ++q) acc[q] |= v[q];
                    }
#pragma unroll
                    for (int q = 0; q < W; ++q) {
                        acc[q] &= ~w0[q] & full[q];
                        pop += __popcll(acc[q]);
                    }
                }
                if (st.fuse) {  // as the tiled last stage: new states to fr0, flags and group counts
                    if (pop) store_row<W>(st.out + o * W, acc);
                    fuse_flags(a, o0 + j0, __ballot(pop != 0), pop);
                } else if (valid) {
                    store_row<W>(st.out + o * W, acc);
                }
            }
        }
        __syncthreads();  // Y and s_base are rewritten by the next tile
    }
}

// After a pull, the frontier rows still hold their depth-d robber bits;
// clear them: the robber step writes depth d+2 into this buffer one level
// later (or depth d+1 right after, where both parities share one buffer).
__global__ __launch_bounds__(kBlock) void k_clear_frontier(LevelArgs a, int32_t W) {
    if (a.ctl->done || !a.ctl->pull) return;
    const uint32_t cur = a.ctl->level & 1;
    const uint64_t n = static_cast<uint64_t>(a.ctl->cnt1[cur]) * W;
    const uint32_t* list = a.list1[cur];
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        a.fr1s[cur][static_cast<uint64_t>(list[i / W]) * W + i % W] = 0;
}

// ---------------------------------------------------------------------------
// robber turn: for every configuration with new cop-turn states at depth d,
// robber r is won iff every move of r lands on a won cop-turn state
// (out_mask[r] within won0); depth = d + 1. Lanes = entries of list0.
// ---------------------------------------------------------------------------
// Robber positions that can stay open in a row, OR_{v open} in_mask[v], by
// nibbles of the open set: lds_nib[(p * 16 + x) * W + q] = word q of the OR of
// the in-masks of the vertices 4p + j for the set bits j of x. A row then
// costs one table row per nibble (n / 4) whatever its density.
constexpr uint32_t kNibbleMin = 24;  // open states above which a row takes the nibble table
// 512-thread workgroups: two groups of rows share one copy of the masks and the
// nibble table, which bound the step's occupancy
constexpr int kRobberBlock = 2 * kBlock;
constexpr int kRobberGroups = kRobberBlock / 64 / kWavesPerBlock;

template <int W>
__global__ __launch_bounds__(kRobberBlock, 8) void k_robber_step(GraphView gv, LevelArgs a) {
    // out-masks, in-masks and the nibble table, sized for this graph (the
    // occupancy of the step is bound by this LDS; LaunchRobber sizes it)
    extern __shared__ __align__(16) uint64_t lds_rob[];
    uint64_t* lds_om = lds_rob;
    uint64_t* lds_im = lds_rob + gv.n * W;
    uint64_t* lds_nib = lds_rob + 2 * gv.n * W;
    const int lane = threadIdx.x & 63;
    if (a.ctl->done) return;
    const uint32_t level = a.ctl->level, cur = level & 1;
    const uint32_t count = a.ctl->cnt0[cur];
    const uint32_t* list = a.list0[cur];
    const uint32_t nwords = (count + 63) / 64;
    const uint32_t ngroups = (nwords + kGroupWords - 1) / kGroupWords;
    if (blockIdx.x * kRobberGroups >= ngroups) return;  // no rows here: skip the mask staging (uniform)
    const uint64_t* lom = stage_out_masks<W>(lds_om, a.out_mask, gv.n);
    const uint64_t* lim = stage_out_masks<W>(lds_im, a.in_mask, gv.n);
    const int nnib = (gv.n + 3) / 4;
    for (int e = threadIdx.x; e < nnib * 16; e += blockDim.x) {
        const int p = e >> 4, x = e & 15;
        uint64_t acc[W];
#pragma unroll
        for (int q = 0; q < W; ++q) acc[q] = 0;
        for (int j = 0; j < 4; ++j)
            if (((x >> j) & 1) && 4 * p + j < gv.n) {
#pragma unroll
                for (int q = 0; q < W; ++q) acc[q] |= lim[(4 * p + j) * W + q];
            }
#pragma unroll
        for (int q = 0; q < W; ++q) lds_nib[e * W + q] = acc[q];
    }
    __syncthreads();
    // kRobberGroups groups per pass, kWavesPerBlock waves (one flag word each) per group
    const uint32_t wid = wave_id() % kWavesPerBlock, half = wave_id() / kWavesPerBlock;
    __shared__ unsigned long long s_red[2][kRobberBlock / 64];
    const int n = gv.n;
    const uint8_t nb = won_byte(level + 1);
    uint64_t full[W];  // the row's robber positions
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const int bits = min(64, max(0, n - 64 * q));
        full[q] = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    }
    for (uint32_t g0 = blockIdx.x * kRobberGroups; g0 < ngroups; g0 += gridDim.x * kRobberGroups) {
        const uint32_t g = g0 + half;  // past the end for the second half of the last pass: no word
        unsigned long long rows = 0, found = 0;
        for (int j = 0; j < kWordsPerWave; ++j) {
            const uint32_t b = g * kGroupWords + wid * kWordsPerWave + j;
            if (b >= nwords) break;
            const uint32_t i = b * 64 + lane;
            const bool live = i < count;
            const uint32_t row = live ? list[i] : 0u;
            const uint64_t o = static_cast<uint64_t>(row) * W;
            uint64_t w0[W], w1[W], f0[W], nw[W], open[W];
            if (live) {  // 16-byte row loads (rows are 16-byte aligned for even W)
                load_row<W>(a.won0 + o, w0);
                load_row<W>(a.won1 + o, w1);
                load_row<W>(a.fr0 + o, f0);
            }
            // After a pulled level with a fused last stage the new cop-turn
            // states are in fr0 only: take them into won0 here (a no-op after
            // other levels), before anything of this level reads won0.
            if (live) {
                bool add0 = false;
#pragma unroll
                for (int q = 0; q < W; ++q) add0 |= (f0[q] & ~w0[q]) != 0;
                if (add0) {
#pragma unroll
                    for (int q = 0; q < W; ++q) w0[q] |= f0[q];
                    store_row<W>(a.won0 + o, w0);
                }
            }
            uint32_t nopen = 0, nnew = 0;
#pragma unroll
            for (int q = 0; q < W; ++q) {
                if (!live) {
                    w0[q] = 0;
                    w1[q] = ~0ull;
                    f0[q] = 0;
                }
                nw[q] = 0;
                open[q] = ~w0[q] & full[q];
                nopen += __popcll(open[q]);
                nnew += __popcll(f0[q]);
            }
            const bool by_open = nopen <= a.open_factor * nnew;
            const bool by_nibble = !by_open && nopen > kNibbleMin && nnew > kNibbleMin / 4;
            if (by_open || by_nibble) {
                // A robber with a move onto an open cop-turn state stays open, and
                // every other robber not yet won is won at this depth (had all its
                // moves been won by an earlier depth, an earlier level would have
                // taken it). The open set is walked bit by bit where it is small
                // and nibble by nibble through the table where it is not.
                uint64_t stay[W];
#pragma unroll
                for (int q = 0; q < W; ++q) stay[q] = 0;
                if (by_open && nopen <= kNibbleMin) {
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        uint64_t x = open[w];
                        while (x) {
                            const uint64_t* im = lim + (64 * w + __builtin_ctzll(x)) * W;
                            x &= x - 1;
#pragma unroll
                            for (int q = 0; q < W; ++q) stay[q] |= im[q];
                        }
                    }
                } else {
                    for (int p = 0; p < nnib; ++p) {
                        const uint32_t x = static_cast<uint32_t>(open[p >> 4] >> (4 * (p & 15))) & 15u;
                        if (x == 0) continue;
                        uint64_t v[W];
                        load_row<W>(lds_nib + (p * 16 + x) * W, v);
#pragma unroll
                        for (int q = 0; q < W; ++q) stay[q] |= v[q];
                    }
                }
#pragma unroll
                for (int q = 0; q < W; ++q) nw[q] = ~stay[q] & ~w1[q] & full[q];
            } else {
                // Only a robber with a move onto a cop-turn state won at this depth
                // can change: its other moves were already counted at earlier depths.
                uint64_t cand[W];
#pragma unroll
                for (int q = 0; q < W; ++q) cand[q] = 0;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    uint64_t x = f0[w];
                    while (x) {
                        const uint64_t* im = lim + (64 * w + __builtin_ctzll(x)) * W;
                        x &= x - 1;
#pragma unroll
                        for (int q = 0; q < W; ++q) cand[q] |= im[q];
                    }
                }
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    uint64_t x = cand[w] & ~w1[w];
                    while (x) {
                        const int rb = __builtin_ctzll(x);
                        x &= x - 1;
                        const uint64_t* om = lom + (64 * w + rb) * W;
                        uint64_t bad = 0;
#pragma unroll
                        for (int q = 0; q < W; ++q) bad |= om[q] & ~w0[q];
                        nw[w] |= static_cast<uint64_t>(bad == 0) << rb;
                    }
                }
            }
            uint32_t pop = 0;
#pragma unroll
            for (int q = 0; q < W; ++q) pop += __popcll(nw[q]);
            const bool append = pop != 0;
            if (append) {
                uint64_t u[W];
#pragma unroll
                for (int q = 0; q < W; ++q) u[q] = w1[q] | nw[q];
                store_row<W>(a.won1 + o, u);
                store_row<W>(a.fr1s[cur ^ 1] + o, nw);  // depth level + 1
#pragma unroll
                for (int q = 0; q < W; ++q) {
                    if (a.log_rows) continue;  // the depth bytes go through the log (k_log_lists)
                    uint64_t bits = nw[q];
                    while (bits) {
                        const int r = q * 64 + __builtin_ctzll(bits);
                        bits &= bits - 1;
                        a.table[(static_cast<uint64_t>(row) * n + r) * 2 + 1] = nb;
                    }
                }
            }
            const uint64_t f = __ballot(append);
            if (lane == 0) a.c1.flags[b] = f;
            rows += __popcll(f);
            found += wave_sum(pop);
        }
        __syncthreads();  // s_red of the previous pass is read
        if (lane == 0) {
            s_red[0][wave_id()] = rows;
            s_red[1][wave_id()] = found;
        }
        __syncthreads();
        if (wid == 0 && lane == 0 && g < ngroups) {
            rows = found = 0;
            for (int i = 0; i < kWavesPerBlock; ++i) {
                rows += s_red[0][half * kWavesPerBlock + i];
                found += s_red[1][half * kWavesPerBlock + i];
            }
            a.c1.gcnt[g] = static_cast<uint32_t>(rows);
            a.c1.gstates[g] = found;
        }
    }
}

// ---------------------------------------------------------------------------
// cop turn, finalise: nx0 & ~won0 -> new cop-turn states (full scan of nx0,
// lanes = configurations)
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(kBlock) void k_cop_finalize(GraphView gv, LevelArgs a) {
    const int lane = threadIdx.x & 63;
    if (a.ctl->done || (a.fuse && a.ctl->pull)) return;  // fused into the last pull stage
    const uint8_t nb = won_byte(a.ctl->level + 1);
    const uint32_t ngroups = (a.nbatch + kGroupWords - 1) / kGroupWords;
    const uint32_t wid = wave_id();
    // after a push, groups the push did not mark have no candidates
    const bool marked_only = a.dirty != nullptr && !a.ctl->pull;
    for (uint32_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        if (marked_only && !a.dirty[g]) {  // uniform per workgroup
            const uint32_t b = g * kGroupWords + wid;
            if (lane == 0 && b < a.nbatch) a.c0.flags[b] = 0;
            if (threadIdx.x == 0) {
                a.c0.gcnt[g] = 0;
                a.c0.gstates[g] = 0;
            }
            continue;
        }
        unsigned long long rows = 0, found = 0;
        for (int j = 0; j < kWordsPerWave; ++j) {
            const uint32_t b = g * kGroupWords + wid * kWordsPerWave + j;
            if (b >= a.nbatch) break;
            const uint32_t row = b * 64 + lane;
            const uint64_t o = static_cast<uint64_t>(row) * W;
            uint64_t x[W];
            bool any = false;
            if (row < gv.C) {
                load_row<W>(a.nx0 + o, x);
#pragma unroll
                for (int q = 0; q < W; ++q) any |= x[q] != 0;
            }
            uint32_t pop = 0;
            if (any) {  // this lane's row has candidates: settle them
                uint64_t nw[W], wv[W], z[W];
                load_row<W>(a.won0 + o, wv);
#pragma unroll
                for (int q = 0; q < W; ++q) {
                    nw[q] = x[q] & ~wv[q];
                    pop += __popcll(nw[q]);
                    wv[q] |= nw[q];
                    z[q] = 0;
                }
                if (pop) store_row<W>(a.won0 + o, wv);
                store_row<W>(a.nx0 + o, z);
                if (pop) {  // a listed row: its new states for the robber step and k_log_lists
                    store_row<W>(a.fr0 + o, nw);
#pragma unroll
                    for (int q = 0; q < W; ++q) {
                        if (a.log_rows) continue;  // the depth bytes go through the log
                        uint64_t bits = nw[q];
                        while (bits) {
                            const int r = q * 64 + __builtin_ctzll(bits);
                            bits &= bits - 1;
                            a.table[(static_cast<uint64_t>(row) * gv.n + r) * 2] = nb;
                        }
                    }
                }
            }
            const uint64_t f = __ballot(pop != 0);
            if (lane == 0) a.c0.flags[b] = f;
            rows += __popcll(f);
            found += wave_sum(pop);
        }
        block_sum2(rows, found);
        if (threadIdx.x == 0) {
            a.c0.gcnt[g] = static_cast<uint32_t>(rows);
            a.c0.gstates[g] = found;
            if (marked_only) a.dirty[g] = 0;  // every thread has read the mark (block_sum2 synchronised)
        }
    }
}

// Block-wide exclusive scan of gcnt[0, ng) in place; returns the total and the
// sum of gstates[0, ng) to thread 0 (one workgroup of kScanThreads threads).
// Tiles of kScanThreads * kScanPer elements: each thread loads kScanPer
// consecutive counts (independent loads), the tile is scanned with wave
// shuffles and one pass over the wave totals.
constexpr int kScanPer = 8;
constexpr uint32_t kScanTileMax = kScanThreads * kScanPer;  // groups one pass of scan_groups_n takes
constexpr uint32_t kScanTileMin = 4;

// L lists of group counts are scanned in the same passes, so their loads are
// in flight together (k_advance scans the cop-turn and robber-turn lists).
template <int L>
__device__ void scan_groups_n(uint32_t* const* gcnt, const unsigned long long* const* gstates, const uint32_t* ng,
                              uint32_t* total, unsigned long long* states) {
    constexpr int kWaves = kScanThreads / 64;
    __shared__ uint32_t wsum[L][kWaves];
    __shared__ unsigned long long wst[L][kWaves];
    const uint32_t t = threadIdx.x, lane = t & 63, wid = t >> 6;
    uint32_t run[L];
    unsigned long long st[L];
    uint32_t ngmax = 0;
#pragma unroll
    for (int l = 0; l < L; ++l) {
        run[l] = 0;
        st[l] = 0;
        ngmax = max(ngmax, ng[l]);
    }
    for (uint32_t base = 0; base < ngmax; base += kScanThreads * kScanPer) {
        const uint32_t i0 = base + t * kScanPer;
        uint32_t v[L][kScanPer], sum[L], incl[L];
#pragma unroll
        for (int l = 0; l < L; ++l) {
            sum[l] = 0;
#pragma unroll
            for (int e = 0; e < kScanPer; ++e) {
                const bool in = i0 + e < ng[l];
                v[l][e] = in ? gcnt[l][i0 + e] : 0u;
                st[l] += in ? gstates[l][i0 + e] : 0ull;
                sum[l] += v[l][e];
            }
            incl[l] = sum[l];
        }
        for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
            for (int l = 0; l < L; ++l) {
                const uint32_t u = __shfl_up(incl[l], off);
                if (lane >= static_cast<uint32_t>(off)) incl[l] += u;
            }
        }
        if (lane == 63) {
#pragma unroll
            for (int l = 0; l < L; ++l) wsum[l][wid] = incl[l];
        }
        __syncthreads();
#pragma unroll
        for (int l = 0; l < L; ++l) {
            uint32_t pre = 0, tile = 0;
            for (int w = 0; w < kWaves; ++w) {
                if (static_cast<uint32_t>(w) < wid) pre += wsum[l][w];
                tile += wsum[l][w];
            }
            uint32_t r = run[l] + pre + incl[l] - sum[l];
#pragma unroll
            for (int e = 0; e < kScanPer; ++e) {
                if (i0 + e < ng[l]) gcnt[l][i0 + e] = r;
                r += v[l][e];
            }
            run[l] += tile;
        }
        __syncthreads();  // wsum is reused by the next tile
    }
#pragma unroll
    for (int l = 0; l < L; ++l) {
        st[l] = wave_sum(st[l]);
        if (lane == 0) wst[l][wid] = st[l];
    }
    __syncthreads();
    if (t == 0) {
#pragma unroll
        for (int l = 0; l < L; ++l) {
            total[l] = run[l];
            states[l] = 0;
            for (int w = 0; w < kWaves; ++w) states[l] += wst[l][w];
        }
    }
    __syncthreads();
}

__device__ void scan_groups(uint32_t* gcnt, const unsigned long long* gstates, uint32_t ng,
                            uint32_t& total, unsigned long long& states) {
    scan_groups_n<1>(&gcnt, &gstates, &ng, &total, &states);
}

// Ends a level: scans the group counts of both producers into list offsets,
// adds the level's state counts to the histogram and moves to depth level+1;
// stops when no row holds a new state or when the next depth would have to
// reach the 127 cap (levels run d = 1..125).
// First half of the two-level scan (large tables): workgroup b scans tile
// b mod T of list b / T (T tiles per list), leaving offsets within the tile in
// gcnt and the tile's totals in tile_cnt / tile_states for k_advance.
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(LevelArgs a, uint32_t T) {
    const Ctl* ctl = a.ctl;
    if (ctl->done) return;  // uniform
    const uint32_t l = blockIdx.x / T, t = blockIdx.x - l * T;
    const uint32_t ng = l == 0 ? (a.nbatch + kGroupWords - 1) / kGroupWords
                               : ((ctl->cnt0[ctl->level & 1] + 63) / 64 + kGroupWords - 1) / kGroupWords;
    const uint32_t base = t * a.scan_tile;
    if (base >= ng) return;  // uniform
    const Compact& c = l == 0 ? a.c0 : a.c1;
    uint32_t total = 0;
    unsigned long long states = 0;
    scan_groups(c.gcnt + base, c.gstates + base, min(a.scan_tile, ng - base), total, states);
    if (threadIdx.x == 0) {
        a.tile_cnt[l * a.tile_stride + t] = total;
        a.tile_states[l * a.tile_stride + t] = states;
    }
}

__device__ void finish_level(const LevelArgs& a, uint32_t d, uint32_t t0, uint32_t t1, unsigned long long s0,
                             unsigned long long s1);

__global__ __launch_bounds__(kScanThreads) void k_advance(LevelArgs a) {
    Ctl* ctl = a.ctl;
    if (ctl->done) {  // a queued level after the last one: its k_write_lists does nothing
        if (threadIdx.x == 0) ctl->log_open = 0;
        return;
    }
    const uint32_t d = ctl->level, cur = d & 1;
    const uint32_t ng0 = (a.nbatch + kGroupWords - 1) / kGroupWords;
    const uint32_t nw1 = (ctl->cnt0[cur] + 63) / 64;
    const uint32_t ng1 = (nw1 + kGroupWords - 1) / kGroupWords;
    uint32_t tt[2] = {0, 0};
    unsigned long long ss[2] = {0, 0};
    if (a.scan_tile) {  // k_scan_tiles has scanned within tiles: scan the tile totals
        uint32_t* gc[2] = {a.tile_cnt, a.tile_cnt + a.tile_stride};
        const unsigned long long* gs[2] = {a.tile_states, a.tile_states + a.tile_stride};
        const uint32_t nt[2] = {(ng0 + a.scan_tile - 1) / a.scan_tile, (ng1 + a.scan_tile - 1) / a.scan_tile};
        scan_groups_n<2>(gc, gs, nt, tt, ss);
    } else {
        uint32_t* gc[2] = {a.c0.gcnt, a.c1.gcnt};
        const unsigned long long* gs[2] = {a.c0.gstates, a.c1.gstates};
        const uint32_t ng[2] = {ng0, ng1};
        scan_groups_n<2>(gc, gs, ng, tt, ss);
    }
    if (threadIdx.x == 0) finish_level(a, d, tt[0], tt[1], ss[0], ss[1]);
}

// The end of level d once both lists are counted: histogram, list lengths,
// direction and log room of depth d + 1, and the stop test.
__device__ void finish_level(const LevelArgs& a, uint32_t d, uint32_t t0, uint32_t t1, unsigned long long s0,
                             unsigned long long s1) {
    Ctl* ctl = a.ctl;
    const uint32_t nxt = (d & 1) ^ 1;
    {
        ctl->lvl[d + 1][0] = s0;
        ctl->lvl[d + 1][1] = s1;
        ctl->cnt0[nxt] = t0;
        ctl->cnt1[nxt] = t1;
        ctl->rows[d + 1][0] = t0;
        ctl->rows[d + 1][1] = t1;
        ctl->pull = choose_pull(t1, a.pull_rows);
        ctl->dir[d + 1] = ctl->pull;
        ctl->level = d + 1;
        // room in the depth log for the lists of depth d + 1, or direct writes
        const unsigned long long need = static_cast<unsigned long long>(t0) + t1;
        ctl->log_now = ctl->log_cap != 0 && ctl->log_used + need <= ctl->log_cap;
        ctl->log_off[d + 1][0] = ctl->log_used;
        ctl->log_off[d + 1][1] = ctl->log_used + t0;
        ctl->log_len[d + 1][0] = ctl->log_now ? t0 : 0u;
        ctl->log_len[d + 1][1] = ctl->log_now ? t1 : 0u;
        if (ctl->log_now) ctl->log_used += need;
        if ((t0 == 0 && t1 == 0) || d + 1 >= kMaxRound - 1) {
            ctl->done = 1;
            ctl->done_level = d;
            ctl->log_open = 1;  // the depth d + 1 lists still carry their states' bytes
        }
    }
}

// After k_init: offsets for the depth-1 lists (parity 1) and the histogram
// entries of depths 0 and 1.
__global__ __launch_bounds__(kScanThreads) void k_settle_init(LevelArgs a) {
    Ctl* ctl = a.ctl;
    const uint32_t ng = (a.nbatch + kGroupWords - 1) / kGroupWords;
    uint32_t t0 = 0, t1 = 0, tc = 0;
    unsigned long long s0 = 0, s1 = 0, caps = 0;
    scan_groups(a.c0.gcnt, a.c0.gstates, ng, t0, s0);
    scan_groups(a.c1.gcnt, a.c1.gstates, ng, t1, s1);
    // capture counts: sum only (the scan result is not used)
    __shared__ unsigned long long cred[kScanThreads / 64];
    unsigned long long c = 0;
    for (uint32_t i = threadIdx.x; i < ng; i += blockDim.x) c += a.gcaps[i];
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) cred[threadIdx.x >> 6] = c;
    __syncthreads();
    (void)tc;
    if (threadIdx.x == 0) {
        for (int i = 0; i < kScanThreads / 64; ++i) caps += cred[i];
        ctl->lvl[0][0] = caps;
        ctl->lvl[0][1] = caps;
        ctl->lvl[1][0] = s0;
        ctl->lvl[1][1] = s1;
        ctl->cnt0[1] = t0;
        ctl->cnt1[1] = t1;
        ctl->rows[1][0] = t0;
        ctl->rows[1][1] = t1;
        ctl->pull = choose_pull(t1, a.pull_rows);
        ctl->dir[1] = ctl->pull;
        ctl->level = 1;
        ctl->done = (t0 == 0 && t1 == 0) ? 1u : 0u;
    }
}

// The list entries of one flag word per lane (consecutive words in
// consecutive lanes; every lane of the wave calls it): bit j of word w is
// entry pos + (set bits below j), holding w * 64 + j, or map[w * 64 + j] when
// the word indexes another list. Small tables and sparse words: each lane
// writes its own word's entries; dense words in HBM: the wave writes them one
// word at a time, lane j the entry of bit j, so stores (and the map reads) are
// contiguous.
__device__ __forceinline__ void write_word_entries(uint64_t bits, uint32_t pc, uint32_t pos, uint32_t word,
                                                   uint32_t* dst, const uint32_t* map, bool small) {
    const uint32_t lane = threadIdx.x & 63;
    unsigned long long live = __ballot(bits != 0);
    if (!live) return;  // uniform
    uint32_t maxpc = pc;  // the wave's densest word
    for (int off = 32; off > 0; off >>= 1) maxpc = max(maxpc, static_cast<uint32_t>(__shfl_xor(static_cast<int>(maxpc), off)));
    if (small || 4 * maxpc <= static_cast<uint32_t>(__popcll(live))) {
        uint32_t p = pos;
        uint64_t rest = bits;
        while (rest) {
            const uint32_t idx = word * 64 + static_cast<uint32_t>(__builtin_ctzll(rest));
            rest &= rest - 1;
            dst[p++] = map ? map[idx] : idx;
        }
        return;
    }
    while (live) {
        const int w = __builtin_ctzll(live);
        live &= live - 1;
        const uint32_t lo = __shfl(static_cast<uint32_t>(bits), w);
        const uint32_t hi = __shfl(static_cast<uint32_t>(bits >> 32), w);
        const uint32_t wpos = __shfl(pos, w);
        const uint64_t wb = (static_cast<uint64_t>(hi) << 32) | lo;
        if ((wb >> lane) & 1ull) {
            const uint32_t idx = (word - lane + w) * 64 + lane;
            dst[wpos + __popcll(wb & ((1ull << lane) - 1))] = map ? map[idx] : idx;
        }
    }
}

// Dense sorted lists from the flag words: list0 of the next parity from the
// finalize flags (values = configurations) and list1 from the robber flags
// (values = the list0 entries they flag). One thread per flag word; the
// words of a group are adjacent lanes, so a segmented scan over kGroupWords
// lanes gives each word's offset inside its group. init_mode serves k_init
// (both spaces are configurations, lists go to parity 1).
__global__ __launch_bounds__(kBlock) void k_write_lists(LevelArgs a, int32_t init_mode) {
    const Ctl* ctl = a.ctl;
    // no further level reads the lists; only the depth that ended the loop
    // still has its states' bytes to record
    if (ctl->done && (init_mode || !ctl->log_open)) return;
    // after k_advance, level already names the next depth: its parity holds the new lists
    const uint32_t nxt = init_mode ? 1u : (ctl->level & 1);
    const uint32_t prev = nxt ^ 1;
    const uint32_t nw0 = a.nbatch;
    const uint32_t nw1 = init_mode ? a.nbatch : (ctl->cnt0[prev] + 63) / 64;
    const uint32_t span0 = (nw0 + 63) / 64 * 64;  // space 0 words, padded to whole waves
    const bool small = a.log_rows == nullptr;     // no depth log: a table that stays in the caches
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t - lane < span0 + nw1; t += gridDim.x * blockDim.x) {
        const bool space1 = t >= span0;  // uniform per wave
        const uint32_t word = space1 ? t - span0 : t;
        const Compact& c = space1 ? a.c1 : a.c0;
        const uint32_t nw = space1 ? nw1 : nw0;
        const uint64_t bits = word < nw ? c.flags[word] : 0ull;
        const uint32_t pc = __popcll(bits);
        uint32_t incl = pc;
        for (int off = 1; off < kGroupWords; off <<= 1) {
            const uint32_t v = __shfl_up(incl, off, kGroupWords);
            if (static_cast<int>(word % kGroupWords) >= off) incl += v;
        }
        uint32_t pos = 0;
        if (bits) {
            const uint32_t g = word / kGroupWords;
            pos = c.gcnt[g] + incl - pc;
            if (a.scan_tile && !init_mode) pos += a.tile_cnt[(space1 ? a.tile_stride : 0u) + g / a.scan_tile];
        }
        uint32_t* dst = space1 ? a.list1[nxt] : a.list0[nxt];
        write_word_entries(bits, pc, pos, word, dst, (space1 && !init_mode) ? a.list0[prev] : nullptr, small);
    }
}

// Single-pass end of level d for tables whose group counts one scan covers
// (no two-level scan): k_advance's scan and bookkeeping and k_write_lists in
// one launch. Workgroups take tiles of kEndTileGroups groups (a flag word per
// thread) in ticket order, the cop-turn space first. A tile publishes its
// count, then walks back over its predecessors' published sums (a
// predecessor's inclusive total ends the walk, its count alone does not:
// decoupled look-back), publishes its inclusive total and writes its words'
// list entries. A tile only waits on tiles with smaller tickets, which are
// running or done. The last tile of each space leaves the space's total; the
// later of the two ends the level. `d` comes from the host: levels enqueued
// together advance one each, and a workgroup that starts after the level has
// ended must not read the next level's number.
constexpr uint32_t kEndTileGroups = kBlock / kGroupWords;

__global__ __launch_bounds__(kBlock) void k_end_level(LevelArgs a, uint32_t d) {
    Ctl* ctl = a.ctl;
    if (ctl->done) {
        // a level queued after the last one (the last level's own late
        // workgroups find its tiles taken and leave the log flag alone)
        if (d > ctl->done_level && blockIdx.x == 0 && threadIdx.x == 0) ctl->log_open = 0;
        return;
    }
    if (d >= kMaxRound) return;  // never reached before `done` (levels run d = 1..125)
    const uint32_t cur = d & 1, nxt = cur ^ 1;
    const uint32_t nw0 = a.nbatch, nw1 = (ctl->cnt0[cur] + 63) / 64;
    const uint32_t ng0 = (nw0 + kGroupWords - 1) / kGroupWords, ng1 = (nw1 + kGroupWords - 1) / kGroupWords;
    const uint32_t T0 = (ng0 + kEndTileGroups - 1) / kEndTileGroups, T1 = (ng1 + kEndTileGroups - 1) / kEndTileGroups;
    const bool small = a.log_rows == nullptr;
    const unsigned long long epoch = static_cast<unsigned long long>(d + 1) << 40;
    __shared__ uint32_t s_tile, s_base;
    __shared__ uint32_t s_gpre[kEndTileGroups];
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&ctl->lb_ticket[d], 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= T0 + T1) break;  // uniform
        const uint32_t sp = tile >= T0 ? 1u : 0u, tt = sp ? tile - T0 : tile;
        const Compact& c = sp ? a.c1 : a.c0;
        const uint32_t ng = sp ? ng1 : ng0, nw = sp ? nw1 : nw0, T = sp ? T1 : T0;
        const uint32_t g0 = tt * kEndTileGroups;
        // the tile's flag words, in flight while wave 0 looks back
        const uint32_t word = g0 * kGroupWords + threadIdx.x;
        const uint64_t bits = word < nw ? c.flags[word] : 0ull;
        if (wid == 0) {  // the tile's group counts: in-tile offsets, count, states; then the look-back
            const uint32_t g = g0 + lane;
            const uint32_t cg = g < ng ? c.gcnt[g] : 0u;
            unsigned long long sg = g < ng ? c.gstates[g] : 0ull;
            uint32_t incl = cg;
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t u = __shfl_up(incl, off);
                if (lane >= static_cast<uint32_t>(off)) incl += u;
            }
            s_gpre[lane] = incl - cg;
            const uint32_t agg = __shfl(incl, 63);
            sg = wave_sum(sg);
            unsigned long long* st = a.lb_status + static_cast<uint64_t>(sp) * a.lb_stride;
            if (lane == 0 && sg) atomicAdd(&ctl->lb_states[sp], sg);  // before the count is published
            if (lane == 0 && tt > 0)
                __hip_atomic_store(&st[tt], epoch | (1ull << 32) | agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            // look-back, 64 predecessors per step: lane j reads tile p_hi - j; the walk
            // sums back to the nearest inclusive total once every tile up to it has published
            uint32_t prefix = 0;
            for (int32_t p_hi = static_cast<int32_t>(tt) - 1; p_hi >= 0;) {
                const int32_t p = p_hi - static_cast<int32_t>(lane);
                unsigned long long v = 0;
                if (p >= 0) v = __hip_atomic_load(&st[p], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                const bool ok = p < 0 || ((v >> 40) == d + 1 && ((v >> 32) & 0xffull) != 0);
                const bool inc = p >= 0 && ok && ((v >> 32) & 0xffull) == 2;
                const unsigned long long incl_mask = __ballot(inc), ready = __ballot(ok);
                const uint32_t stop = incl_mask ? static_cast<uint32_t>(__builtin_ctzll(incl_mask)) : 63u;
                const unsigned long long need = stop >= 63 ? ~0ull : ((2ull << stop) - 1);  // lanes 0..stop
                if ((ready & need) != need) {  // a tile up to the stop has not published yet
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                const uint32_t part = (lane <= stop && p >= 0) ? static_cast<uint32_t>(v) : 0u;
                prefix += static_cast<uint32_t>(wave_sum(static_cast<unsigned long long>(part)));
                if (incl_mask) break;
                p_hi -= 64;
            }
            if (lane == 0) {
                __hip_atomic_store(&st[tt], epoch | (2ull << 32) | (prefix + agg), __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
                s_base = prefix;
                if (tt + 1 == T) {  // the space's last tile: its total; the second space to finish ends the level
                    __hip_atomic_store(&ctl->lb_total[sp], prefix + agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __threadfence();
                    const uint32_t expected = (T0 > 0 ? 1u : 0u) + (T1 > 0 ? 1u : 0u);
                    if (atomicAdd(&ctl->lb_arrived[d], 1u) + 1 == expected) {
                        __threadfence();
                        auto ld32 = [](uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); };
                        auto ld64 = [](unsigned long long* p) {
                            return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        };
                        const uint32_t t0 = T0 ? ld32(&ctl->lb_total[0]) : 0u, t1 = T1 ? ld32(&ctl->lb_total[1]) : 0u;
                        const unsigned long long s0 = ld64(&ctl->lb_states[0]), s1 = ld64(&ctl->lb_states[1]);
                        ctl->lb_states[0] = 0;
                        ctl->lb_states[1] = 0;
                        finish_level(a, d, t0, t1, s0, s1);
                    }
                }
            }
        }
        __syncthreads();
        // the tile's words, one per thread, to their list positions
        const uint32_t pc = __popcll(bits);
        uint32_t incl = pc;
        for (int off = 1; off < kGroupWords; off <<= 1) {
            const uint32_t v = __shfl_up(incl, off, kGroupWords);
            if (static_cast<int>(word % kGroupWords) >= off) incl += v;
        }
        const uint32_t pos = s_base + s_gpre[threadIdx.x / kGroupWords] + incl - pc;
        write_word_entries(bits, pc, pos, word, sp ? a.list1[nxt] : a.list0[nxt], sp ? a.list0[cur] : nullptr, small);
        // a grid with a workgroup per tile takes one ticket each (the tickets
        // are same-address atomics, which serialise)
        if (gridDim.x >= T0 + T1) break;  // uniform
        __syncthreads();  // s_tile, s_base and s_gpre are rewritten by the next tile
    }
}

// The depth bytes of the states listed by k_write_lists (not in init mode):
// one thread per list entry takes the row's new robber bits (fr1 of the new
// depth's parity for the robber turn, fr0 for the cop turn) into the depth
// log, or, when the log has no room for this depth, writes the bytes into the
// table directly.
template <int W>
__global__ __launch_bounds__(kBlock) void k_log_lists(LevelArgs a) {
    const Ctl* ctl = a.ctl;
    if (ctl->done && !ctl->log_open) return;
    const uint32_t depth = ctl->level, nxt = depth & 1;
    const uint32_t n0 = ctl->cnt0[nxt], n1 = ctl->cnt1[nxt];
    const bool logging = ctl->log_now != 0;
    const uint8_t nb = won_byte(depth);
    const unsigned long long off0 = ctl->log_off[depth][0], off1 = ctl->log_off[depth][1];
    const uint64_t total = static_cast<uint64_t>(n0) + n1;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const bool robber = i >= n0;
        const uint32_t k = static_cast<uint32_t>(robber ? i - n0 : i);
        const uint32_t row = (robber ? a.list1[nxt] : a.list0[nxt])[k];
        const uint64_t o = static_cast<uint64_t>(row) * W;
        uint64_t w[W];
        load_row<W>((robber ? a.fr1s[nxt] : a.fr0) + o, w);
        if (logging) {
            const unsigned long long e = (robber ? off1 : off0) + k;
            a.log_rows[e] = row;
            store_row<W>(a.log_bits + e * W, w);
            // the list's range in each chunk: its first entry there writes the
            // begin, its last the end (the list is sorted by row)
            const uint32_t* list = robber ? a.list1[nxt] : a.list0[nxt];
            const uint32_t len = robber ? n1 : n0, c = row / a.log_R;
            const uint64_t cell = static_cast<uint64_t>(2 * (depth - ctl->log_first) + (robber ? 1 : 0)) * a.log_chunks + c;
            if (k == 0 || list[k - 1] / a.log_R != c) a.log_beg[cell] = k;
            if (k + 1 == len || list[k + 1] / a.log_R != c) a.log_end[cell] = k + 1;
        } else {
            uint8_t* trow = a.table + static_cast<uint64_t>(row) * a.n * 2 + (robber ? 1 : 0);
#pragma unroll
            for (int q = 0; q < W; ++q) {
                uint64_t x = w[q];
                while (x) {
                    trow[2 * (q * 64 + __builtin_ctzll(x))] = nb;
                    x &= x - 1;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// depth log -> table. The levels record each depth's new states as sorted
// lists of (configuration, robber bits); writing their depth bytes straight
// into the table would touch a few bytes of many 32-byte sectors of a row at
// every depth. Instead the table is taken in chunks of R rows through LDS
// (coalesced), every logged depth's entries for the chunk are applied there,
// and the chunk is written back once. k_log_lists has recorded, for every
// list and chunk, the list's entries with rows in the chunk.
// ---------------------------------------------------------------------------
constexpr int kMaxLogLists = 256;  // lists (depth, turn) of one log segment
struct LogList {
    unsigned long long off;    // first entry in the log
    uint32_t len;              // entries
    uint32_t turn_byte;        // turn << 8 | depth byte
};

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* p, uint64_t count, uint32_t v) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        p[i] = v;
}

// Exclusive scan of one value per thread over the block (kBlock threads);
// returns the thread's prefix and the block total in `total`.
__device__ __forceinline__ uint32_t block_scan_u32(uint32_t v, uint32_t& total) {
    __shared__ uint32_t wsum[kWavesPerBlock];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(incl, off);
        if (lane >= off) incl += u;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    uint32_t base = 0;
    total = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) {
        if (w < wid) base += wsum[w];
        total += wsum[w];
    }
    __syncthreads();
    return base + incl - v;
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_apply_log(const LogList* lists, int32_t nl, const uint32_t* rows,
                                                      const uint64_t* bits, const uint32_t* beg, const uint32_t* end,
                                                      uint8_t* table, uint32_t n, uint32_t C, uint32_t R) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t s_beg[256], s_pre[256];
    __shared__ unsigned long long s_off[256];
    __shared__ uint32_t s_tb[256];
    const uint32_t rowbytes = 2 * n;
    const uint32_t nchunks = (C + R - 1) / R;
    for (int L = threadIdx.x; L < nl; L += blockDim.x) {
        s_off[L] = lists[L].off;
        s_tb[L] = lists[L].turn_byte;
    }
    for (uint32_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const uint32_t R0 = ch * R, R1 = min(R0 + R, C);
        __syncthreads();  // the previous chunk is written back
        uint32_t c = 0;
        if (threadIdx.x < static_cast<uint32_t>(nl)) {
            const uint32_t e = end[static_cast<uint64_t>(threadIdx.x) * nchunks + ch];
            if (e) {
                const uint32_t b = beg[static_cast<uint64_t>(threadIdx.x) * nchunks + ch];
                s_beg[threadIdx.x] = b;
                c = e - b;
            }
        }
        uint32_t total = 0;
        const uint32_t pre = block_scan_u32(c, total);
        if (threadIdx.x < static_cast<uint32_t>(nl)) s_pre[threadIdx.x] = pre;
        if (total == 0) continue;  // uniform
        // the chunk's rows into LDS: 16-byte words (R is a multiple of 8, so
        // R0 * 2n is a multiple of 16), then the tail bytes
        uint8_t* trow = table + static_cast<uint64_t>(R0) * rowbytes;
        const uint32_t nbytes = (R1 - R0) * rowbytes, nvec = nbytes / 16;
        {
            const uint4* src = reinterpret_cast<const uint4*>(trow);
            uint4* dst = reinterpret_cast<uint4*>(lds);
            const uint32_t bd = blockDim.x;
            uint32_t i0 = threadIdx.x;
            for (; i0 + 3 * bd < nvec; i0 += 4 * bd) {  // four loads in flight, no conditions
                const uint4 x0 = src[i0], x1 = src[i0 + bd], x2 = src[i0 + 2 * bd], x3 = src[i0 + 3 * bd];
                dst[i0] = x0;
                dst[i0 + bd] = x1;
                dst[i0 + 2 * bd] = x2;
                dst[i0 + 3 * bd] = x3;
            }
            for (; i0 < nvec; i0 += bd) dst[i0] = src[i0];
        }
        for (uint32_t i = nvec * 16 + threadIdx.x; i < nbytes; i += blockDim.x) lds[i] = trow[i];
        __syncthreads();
        constexpr int kU = 4;  // entries per thread whose loads are in flight together
        for (uint32_t e0 = threadIdx.x; e0 < total; e0 += blockDim.x * kU) {
            uint32_t row[kU], tb[kU];
            uint64_t w[kU][W];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const uint32_t e = e0 + u * blockDim.x;
                tb[u] = 0xFFFFFFFFu;
                if (e >= total) continue;
                int lo = 0, hi = nl - 1;  // the list holding the chunk's entry e (the last with prefix <= e)
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_pre[mid] <= e) lo = mid; else hi = mid - 1;
                }
                const unsigned long long ent = s_off[lo] + s_beg[lo] + (e - s_pre[lo]);
                row[u] = rows[ent];
                load_row<W>(bits + ent * W, w[u]);
                tb[u] = s_tb[lo];
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                if (tb[u] == 0xFFFFFFFFu) continue;
                uint8_t* rp = lds + (row[u] - R0) * rowbytes + (tb[u] >> 8);
                const uint8_t byte = static_cast<uint8_t>(tb[u] & 0xFF);
#pragma unroll
                for (int q = 0; q < W; ++q) {
                    uint64_t x = w[u][q];
                    while (x) {
                        const uint32_t r = q * 64 + __builtin_ctzll(x);
                        x &= x - 1;
                        if (r < n) rp[2 * r] = byte;
                    }
                }
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nvec; i += blockDim.x)
            reinterpret_cast<uint4*>(trow)[i] = reinterpret_cast<const uint4*>(lds)[i];
        for (uint32_t i = nvec * 16 + threadIdx.x; i < nbytes; i += blockDim.x) trow[i] = lds[i];
    }
}

// ---------------------------------------------------------------------------
// cop turn, push over the precomputed predecessor CSR (built once by
// cr_build_aux_graph): one wavefront per frontier row streams its
// predecessor slice and ORs the row's robber mask into their nx0 words
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_cop_push_csr(GraphView gv, LevelArgs a,
                                                         const unsigned long long* heads,
                                                         const uint32_t* csr) {
    if (a.ctl->done || a.ctl->pull) return;
    const uint32_t cur = a.ctl->level & 1;
    const int lane = threadIdx.x & 63;
    const int W = gv.W;
    const uint32_t count = a.ctl->cnt1[cur];
    const uint32_t* list = a.list1[cur];
    for (uint32_t it = blockIdx.x * kWavesPerBlock + wave_id(); it < count; it += gridDim.x * kWavesPerBlock) {
        const uint32_t row = list[it];
        uint64_t m[4] = {0, 0, 0, 0};
        for (int w = 0; w < W; ++w) m[w] = a.fr1s[cur][static_cast<uint64_t>(row) * W + w];
        const unsigned long long beg = heads[row], end = heads[row + 1];
        for (unsigned long long i = beg + lane; i < end; i += 256) {
            uint64_t pr[4];
            bool keep[4], hit[4] = {false, false, false, false};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                keep[u] = i + 64 * u < end;
                pr[u] = keep[u] ? csr[i + 64 * u] : 0u;
            }
            // filter loads first, then the atomics (as in k_cop_push)
            uint64_t wv[4][4] = {};
            if (a.filter) {
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (keep[u]) {
                        const uint64_t* src = a.won0 + pr[u] * W;
#pragma unroll
                        for (int w = 0; w < 4; ++w)
                            if (w < W && m[w]) wv[u][w] = src[w];
                    }
            }
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                if (w >= W || m[w] == 0) continue;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint64_t mw = keep[u] ? m[w] & ~wv[u][w] : 0ull;
                    if (mw) {
                        atomicOr(reinterpret_cast<unsigned long long*>(a.nx0 + pr[u] * W + w),
                                 static_cast<unsigned long long>(mw));
                        hit[u] = true;
                    }
                }
            }
            if (a.dirty) {
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (hit[u]) a.dirty[pr[u] >> 8] = 1;
            }
        }
        if (lane < W) a.fr1s[cur][static_cast<uint64_t>(row) * W + lane] = 0;
    }
}

// ---------------------------------------------------------------------------
// verdict (k_cops_visibility.cpp:301-354): universal rows -> min worst depth,
// lowest cId on ties. Lanes = configurations; each lane scans its own row
// (16-byte loads where the two-column rows allow), the wave keeps the least
// packed (worst << 32 | cId) and issues one atomicMin.
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(kBlock) void k_verdict(GraphView gv, LevelArgs a, int32_t columns) {
    const int lane = threadIdx.x & 63;
    const int n = gv.n;
    uint64_t full[W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const int bits = min(64, n - 64 * q);
        full[q] = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    }
    // two-column rows of 2n bytes, 16-byte aligned: a wave's 64 rows are one
    // contiguous span, read in consecutive 16-byte chunks with the per-row
    // maxima gathered in LDS
    const bool vec = columns == 2 && (n & 7) == 0;
    __shared__ uint32_t rmax[kWavesPerBlock][64];
    const uint32_t wid = threadIdx.x >> 6;
    unsigned long long best = ~0ull;
    const uint32_t nbatch = (gv.C + 63) / 64;
    for (uint32_t b = blockIdx.x * kWavesPerBlock + wid; b < nbatch; b += gridDim.x * kWavesPerBlock) {
        const uint32_t row = b * 64 + lane;
        bool uni = row < gv.C;
#pragma unroll
        for (int q = 0; q < W; ++q) uni = uni && a.won0[static_cast<uint64_t>(row) * W + q] == full[q];
        if (__ballot(uni) == 0) continue;  // uniform per wave
        uint32_t top = 0;  // largest cop-turn byte of the lane's row
        if (vec) {
            rmax[wid][lane] = 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const uint32_t rows_here = min(64u, gv.C - b * 64), cpr = static_cast<uint32_t>(n) / 8;
            const uint4* v = reinterpret_cast<const uint4*>(a.table + static_cast<uint64_t>(b) * 64 * n * 2);
            for (uint32_t i = lane; i < rows_here * cpr; i += 64) {
                const uint4 x = v[i];
                const uint32_t w[4] = {x.x, x.y, x.z, x.w};
                uint32_t m = 0;
#pragma unroll
                for (int h = 0; h < 4; ++h) m = max(m, max(w[h] & 0xFFu, (w[h] >> 16) & 0xFFu));
                atomicMax(&rmax[wid][i / cpr], m);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            top = rmax[wid][lane];
        } else if (uni) {
            const uint8_t* t = a.table + static_cast<uint64_t>(row) * n * columns;
            for (int r = 0; r < n; ++r) top = max(top, static_cast<uint32_t>(t[static_cast<uint64_t>(r) * columns]));
        }
        if (uni) best = min(best, (static_cast<unsigned long long>(top >> 1) << 32) | row);
    }
    for (int off = 32; off > 0; off >>= 1) best = min(best, static_cast<unsigned long long>(__shfl_xor(best, off)));
    if (lane == 0 && best != ~0ull) atomicMin(&a.ctl->verdict, best);
}

// The won marks of column 0 (the cop-turn column) of a table read from a
// cache blob, as the verdict's universal-row test reads them: lanes = rows.
__global__ __launch_bounds__(kBlock) void k_marks_from_table(const uint8_t* table, uint64_t C, int32_t n,
                                                             int32_t columns, int32_t W, uint64_t* marks) {
    for (uint64_t row = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; row < C;
         row += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint8_t* t = table + row * static_cast<uint64_t>(n) * columns;
        for (int q = 0; q < W; ++q) {
            uint64_t m = 0;
            for (int r = 64 * q; r < min(n, 64 * q + 64); ++r)
                m |= static_cast<uint64_t>(t[static_cast<uint64_t>(r) * columns] & 1u) << (r & 63);
            marks[row * W + q] = m;
        }
    }
}

// ---------------------------------------------------------------------------
// retro-edge count (metric only): |pred(C)| by an LDS hash set per wave
// ---------------------------------------------------------------------------
// mode 0: retro-edge count of a solved table over predecessor sets (in-lists);
// mode 1: total forward team moves sum |T(C)| (out-lists), the reference's
//         transitions.size() (AuxGraph.h:258-274);
// mode 2: |pred(C)| of every configuration into rowcnt[C];
// mode 3: the distinct predecessors of C into csr[heads[C] ..) (any order).
struct MoveArgs {
    uint32_t* rowcnt;
    const unsigned long long* heads;
    uint32_t* csr;
};

template <int K>
__global__ __launch_bounds__(kWave) void k_count_retro(GraphView gv, LevelArgs a, uint32_t slots,
                                                       uint32_t shift, int32_t mode, MoveArgs mv,
                                                       unsigned long long* ghs) {
    __shared__ uint32_t fill_pos;
    __shared__ unsigned long long hist[2][128];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i >> 7][i & 127] = 0;
    extern __shared__ __align__(16) uint8_t lds[];
    load_image(lds, gv.image, gv.image_bytes);
    // slot = (row + 1) << 32 | rank: a slot tagged with another row counts as
    // empty. The set is in LDS, or (ghs) in a per-workgroup slice of device
    // memory where the product of the move counts makes it too large for LDS.
    unsigned long long* hs = ghs ? ghs + static_cast<uint64_t>(blockIdx.x) * slots
                                 : reinterpret_cast<unsigned long long*>(lds + ((gv.image_bytes + 15) & ~15u));
    const uint32_t* rank_tab = reinterpret_cast<const uint32_t*>(lds);
    const uint32_t* binom = reinterpret_cast<const uint32_t*>(lds + gv.off_binom);
    const uint16_t* in_off = reinterpret_cast<const uint16_t*>(lds + gv.off_in_off);
    const uint16_t* mv_off = reinterpret_cast<const uint16_t*>(lds + (mode == 1 ? gv.off_out_off : gv.off_in_off));
    const uint8_t* mv_nbr = lds + (mode == 1 ? gv.off_out_nbr : gv.off_in_nbr);
    const int lane = threadIdx.x;
    for (uint32_t i = lane; i < slots; i += 64) hs[i] = 0;
    __syncthreads();
    unsigned long long e_cop = 0, e_rob = 0, won = 0;
    for (uint32_t row = blockIdx.x; row < gv.C; row += gridDim.x) {
        uint32_t cops[K], deg[K], base[K];
        unrank<K>(row, gv, binom, cops);
        uint32_t P = 1;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            base[i] = mv_off[cops[i]];
            deg[i] = mv_off[cops[i] + 1] - base[i];
            P *= deg[i];
        }
        const unsigned long long tag = static_cast<unsigned long long>(row + 1) << 32;
        uint32_t distinct = 0;
        if (lane == 0) fill_pos = 0;
        __syncthreads();
        uint32_t* out = mode == 3 ? mv.csr + mv.heads[row] : nullptr;
        for (uint32_t t = lane; t < P; t += 64) {
            uint32_t tup[K];
            if (!product_tuple<K>(t, cops, deg, base, mv_nbr, tup)) continue;
            const uint32_t pr = rank_sorted<K>(tup, rank_tab, gv.n, gv.ctot_m1);
            const unsigned long long key = tag | pr;
            uint32_t h = ((pr + 1u) * 2654435761u) >> shift;
            for (;;) {
                unsigned long long cur = hs[h];
                if (cur == key) break;
                if ((cur >> 32) != (tag >> 32)) {
                    const unsigned long long seen = atomicCAS(&hs[h], cur, key);
                    if (seen == cur) {
                        ++distinct;
                        if (mode == 3) out[atomicAdd(&fill_pos, 1u)] = pr;
                        break;
                    }
                    continue;  // lost the race: re-examine this slot
                }
                h = (h + 1) & (slots - 1);
            }
        }
        for (int off = 32; off > 0; off >>= 1) distinct += __shfl_xor(distinct, off);
        if (mode == 2 && lane == 0) mv.rowcnt[row] = distinct;
        if (mode) {
            e_cop += distinct;
            __syncthreads();  // fill_pos is reset for the next row
            continue;
        }
        __syncthreads();
        uint32_t rob = 0, cop_e = 0, w = 0;
        for (int r = lane; r < gv.n; r += 64) {
            const uint8_t b0 = a.table[(static_cast<uint64_t>(row) * gv.n + r) * 2];
            const uint8_t b1 = a.table[(static_cast<uint64_t>(row) * gv.n + r) * 2 + 1];
            if (b0 & 1) {
                ++w;
                if ((b0 >> 1) <= kMaxRound - 2) {
                    const uint32_t e = in_off[r + 1] - in_off[r];
                    cop_e += e;
                    atomicAdd(&hist[1][b0 >> 1], static_cast<unsigned long long>(e));
                }
            }
            if (b1 & 1) {
                ++w;
                if ((b1 >> 1) <= kMaxRound - 2) {
                    ++rob;
                    atomicAdd(&hist[0][b1 >> 1], static_cast<unsigned long long>(distinct));
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) rob += __shfl_xor(rob, off);
        e_cop += static_cast<unsigned long long>(rob) * distinct;
        e_rob += cop_e;
        won += w;
    }
    for (int off = 32; off > 0; off >>= 1) {
        e_rob += __shfl_xor(e_rob, off);
        won += __shfl_xor(won, off);
    }
    if (lane == 0) {
        atomicAdd(&a.ctl->retro[0], e_cop);
        atomicAdd(&a.ctl->retro[1], e_rob);
        atomicAdd(&a.ctl->retro[2], won);
    }
    __syncthreads();
    if (mode == 0)
        for (int i = threadIdx.x; i < 256; i += blockDim.x)
            if (hist[i >> 7][i & 127]) atomicAdd(&a.ctl->hist[i >> 7][i & 127], hist[i >> 7][i & 127]);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct KTime {
    double ms = 0;
    int64_t launches = 0;
};

enum Stage { kEmpty = 0, kLoaded, kBuilt, kInitialized, kSolved };

}  // namespace

struct cr_solver {
    int device = 0;
    hipStream_t stream = nullptr;
    Stage stage = kEmpty;
    cr::Adjacency g;
    cr::Adjacency gc;              // cop moves: g, plus each vertex itself when cops may stay
    std::vector<uint64_t> rmask;   // [n][4] robber moves: out-edges, plus r itself when the robber may stay
    int32_t cop_stay = 0, robber_stay = 0;  // AuxGraph::setSelfEdges (AuxGraph.h:81-91)
    std::string path;
    int32_t k = 0, p = 1, variant = CR_VARIANT_COLUMNS;
    uint64_t C = 0, S = 0;
    int32_t W = 0, columns = 2;
    // device memory
    uint8_t* d_table = nullptr;
    uint64_t* d_won0 = nullptr;
    uint64_t* d_won1 = nullptr;
    uint64_t* d_fr1[2] = {nullptr, nullptr};  // robber-turn frontier by depth parity
    uint64_t* d_fr0 = nullptr;    // per configuration: the cop-turn states its last depth added
    uint64_t* d_nx0 = nullptr;
    uint32_t* d_list[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [turn][parity] dense row lists
    // depth log (k_write_lists, k_apply_log)
    uint32_t* d_log_rows = nullptr;
    uint64_t* d_log_bits = nullptr;
    uint64_t log_cap = 0;          // entries; 0: the lists write their depth bytes directly
    uint32_t log_factor = 40;      // tuning: log capacity in units of C entries (0: no log)
    uint64_t log_min_configs = 1ull << 24;  // tuning: no log for fewer configurations
    LogList* d_lists = nullptr;    // descriptors of the depths k_apply_log applies
    uint32_t* d_starts = nullptr;  // depth log: per list and chunk, first entry and one past the last (k_log_lists)
    LogList* h_lists = nullptr;    // pinned staging of the descriptors
    uint32_t log_first = 0;        // first depth whose lists are not in the table yet
    uint64_t* d_flags[2] = {nullptr, nullptr};                          // producer flag words per turn
    uint32_t* d_gcnt[2] = {nullptr, nullptr};                           // per-group counts / offsets
    unsigned long long* d_gstates[2] = {nullptr, nullptr};              // per-group state counts
    unsigned long long* d_gcaps = nullptr;
    unsigned long long* d_heads = nullptr;  // predecessor CSR (C+1 heads, u32 ranks)
    uint32_t* d_csr = nullptr;
    uint64_t csr_entries = 0;
    double csr_build_ms = 0;
    Ctl* d_ctl = nullptr;
    uint8_t* d_image = nullptr;
    uint64_t* d_outmask = nullptr;
    uint64_t* d_inmask = nullptr;  // [n][4]: robber positions with a move onto each vertex
    Ctl* h_ctl = nullptr;  // pinned
    uint64_t device_bytes = 0;
    GraphView gv{};
    // tuning knobs (cr_set_tuning)
    int32_t filter = 1;
    int32_t use_csr = 0;  // precompute predecessor lists (off by default: no gain measured)
    int64_t pull_scale = -1;      // pull threshold scale (permille of the estimated break-even); -1: by size
    uint64_t* d_sep[2] = {nullptr, nullptr};  // factorised-pull intermediates H_i
    uint64_t ms_count[CR_MAX_K + 1] = {};     // C(n+j-1, j): multisets of size j
    bool pull_tiles = true;                   // k_pull_tile (else k_sep_stage) runs the pull stages
    uint32_t tile_rows = 4;                   // k_pull_tile: rows of each input block per tile
    uint32_t tile_blocks = 8192;              // k_pull_tile: workgroups
    uint32_t tile_chunk = 16;                 // k_pull_tile: consecutive tiles a workgroup takes at a time
    uint32_t tile1_rows = 9;                  // k_pull_tile1 (stage 1 of a tiled multi-cop pull): B per tile; 0 = off
    uint32_t open_factor = 2;                 // robber step: open-state path where open <= factor * new
    uint32_t scan_tile = kScanTileMax;        // groups per tile of the two-level level scan
    uint32_t tile_stride = 0;
    uint32_t* d_tile_cnt = nullptr;
    uint8_t* d_dirty = nullptr;               // per group: marked by the push for k_cop_finalize
    unsigned long long* d_lb_status = nullptr;  // k_end_level: published tile sums, [2][lb_stride]
    bool from_blob = false;                   // the table was read by cr_load_aux_graph (no solve state)
    uint64_t* d_marks = nullptr;              // won marks of column 0 of a table read from a blob
    uint32_t lb_stride = 0;
    bool end_pass = true;                     // tuning: one k_end_level instead of k_advance + k_write_lists
    bool push_dirty = true;
    unsigned long long* d_tile_states = nullptr;
    uint32_t pull_gate = 4;                   // a chunk may pull if its first frontier * pull_gate > pull_rows
    int32_t pull_pair = 1;                    // tuning: the last two pull stages as one kernel (k_pull_pair): 0 never, 1 k = 3, 2 k >= 3
    uint32_t pair_rows = 4;                   // tuning: A-multisets per k_pull_pair tile
    uint32_t pair_blocks = 1024;              // tuning: k_pull_pair workgroups
    bool pair_ready = false;                  // k_pull_pair runs this main loop (tile tables below built)
    uint32_t pair_r = 0, pair_nR = 0;         // tile rows and A-ranges of the built tables
    uint32_t* d_pair = nullptr;               // [pair_nR + 1] first tile of each A-range, then [pair_nR] its first c
    uint64_t pair_cap = 0;                    // words of d_pair
    bool fuse_finalize = true;                // tuning: with the depth log, pulled levels finalize in the last stage
    bool fuse_ok = false;                     // the last pull stage runs as LDS tiles (image and one row per vertex fit)
    bool stage2_cleared = false;              // the pull being enqueued clears the frontier in stage 2
    uint32_t pull_rows = 0xffffffffu;         // frontier rows above which a level pulls
    uint32_t push_blocks = 2048;
    uint32_t row_blocks = 32768;
    uint32_t list_blocks = 4096;
    int32_t levels_per_sync = 8;
    // limited-visibility variant (macro.hip)
    crd::MacroPlan* macro = nullptr;
    uint64_t* d_occ = nullptr;  // cop mask per configuration
    crd::MacroProfile mprof;
    // column solver for p >= 2 (columns.hip)
    crd::ColumnsPlan* colplan = nullptr;
    std::vector<uint64_t> reach;  // [n][4]: vertices the robber reaches from r in exactly p moves

    uint32_t nbatch() const { return static_cast<uint32_t>((C + 63) / 64); }
    uint64_t ngroups() const { return (nbatch() + kGroupWords - 1) / kGroupWords; }
    // grid of the row kernels (one workgroup per group of flag words, grid-stride)
    uint32_t row_grid() const {
        return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(ngroups(), row_blocks)));
    }
    uint32_t writer_grid() const {  // k_write_lists: a thread per flag word of both lists
        return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((2 * nbatch() + 255) / 256 + 1, 4096)));
    }
    uint32_t log_grid() const {  // k_log_lists: a thread per list entry (up to 2C), grid-stride
        return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((2 * C + 255) / 256, list_blocks)));
    }
    // results
    std::vector<uint64_t> pass_counts;
    cr_stats stats{};
    bool timing = false;
    std::vector<std::unique_ptr<Timer>>* timers = nullptr;  // events of the chunk in flight
    std::map<std::string, KTime> ktimes;
    std::vector<std::string> korder;
    std::vector<hipEvent_t> events;  // reusable timing events
    std::string time_only;           // when set, only launches of this kernel are timed
    const Timer* launch_timer = nullptr;  // events the next single-kernel launch records itself

    ~cr_solver() { release(); }

    void release() {
        hipFree(d_table);
        hipFree(d_won0);
        hipFree(d_won1);
        if (d_fr1[1] != d_fr1[0]) hipFree(d_fr1[1]);
        hipFree(d_fr1[0]);
        hipFree(d_nx0);
        for (auto& t : d_list)
            for (auto& b : t) { hipFree(b); b = nullptr; }
        for (int i = 0; i < 2; ++i) {
            hipFree(d_flags[i]);
            hipFree(d_gcnt[i]);
            hipFree(d_gstates[i]);
            d_flags[i] = nullptr;
            d_gcnt[i] = nullptr;
            d_gstates[i] = nullptr;
        }
        hipFree(d_gcaps);
        hipFree(d_tile_cnt);
        hipFree(d_dirty);
        d_dirty = nullptr;
        hipFree(d_lb_status);
        d_lb_status = nullptr;
        hipFree(d_marks);
        d_marks = nullptr;
        from_blob = false;
        if (d_pair) hipFree(d_pair);
        d_pair = nullptr;
        pair_cap = 0;
        pair_ready = false;
        pair_r = 0;
        hipFree(d_tile_states);
        d_tile_cnt = nullptr;
        d_tile_states = nullptr;
        d_gcaps = nullptr;
        hipFree(d_heads);
        hipFree(d_csr);
        d_heads = nullptr;
        d_csr = nullptr;
        hipFree(d_sep[0]);
        hipFree(d_sep[1]);
        d_sep[0] = d_sep[1] = nullptr;
        hipFree(d_log_rows);
        hipFree(d_log_bits);
        hipFree(d_lists);
        hipFree(d_starts);
        d_starts = nullptr;
        d_log_rows = nullptr;
        d_log_bits = nullptr;
        d_lists = nullptr;
        log_cap = 0;
        csr_entries = 0;
        crd::macro_free(macro, device_bytes);
        macro = nullptr;
        crd::columns_free(colplan, device_bytes);
        colplan = nullptr;
        hipFree(d_occ);
        d_occ = nullptr;
        hipFree(d_ctl);
        hipFree(d_image);
        hipFree(d_outmask);
        hipFree(d_inmask);
        hipFree(d_fr0);
        d_inmask = nullptr;
        d_fr0 = nullptr;
        d_table = nullptr;
        d_won0 = d_won1 = d_nx0 = nullptr;
        d_fr1[0] = d_fr1[1] = nullptr;
        d_ctl = nullptr;
        d_image = nullptr;
        d_outmask = nullptr;
        device_bytes = 0;
    }
};

namespace {

// Timing events are created once and reused: creating a pair per launch costs
// more host time than the short level kernels take on the device.
hipEvent_t acquire_event(cr_solver* s) {
    if (s->events.empty()) {
        hipEvent_t e = nullptr;
        hipEventCreate(&e);
        return e;
    }
    hipEvent_t e = s->events.back();
    s->events.pop_back();
    return e;
}

int check_solver(const cr_solver* s) {
    if (s == nullptr) return fail(CR_ERR_ARG, "null solver");
    return 0;
}

// Dispatch helpers over the compile-time cop count K and row width W.
template <template <int> class F, typename... Args>
int dispatch_k(int k, Args&&... args) {
    switch (k) {
        case 1: return F<1>::run(args...);
        case 2: return F<2>::run(args...);
        case 3: return F<3>::run(args...);
        case 4: return F<4>::run(args...);
        case 5: return F<5>::run(args...);
        case 6: return F<6>::run(args...);
        case 7: return F<7>::run(args...);
        case 8: return F<8>::run(args...);
        default: return fail(CR_ERR_K_RANGE, "k=%d outside 1..%d", k, CR_MAX_K);
    }
}

template <template <int> class F, typename... Args>
int dispatch_w(int w, Args&&... args) {
    switch (w) {
        case 1: return F<1>::run(args...);
        case 2: return F<2>::run(args...);
        case 3: return F<3>::run(args...);
        case 4: return F<4>::run(args...);
        default: return fail(CR_ERR_ARG, "row width %d", w);
    }
}

// Blocks for `waves` wavefronts of work, capped (kernels loop grid-stride).
// Kernels that copy the graph image into LDS use a smaller cap so the copy is
// amortised over many rows.
uint32_t grid_for(uint64_t waves, uint64_t cap = 8192) {
    const uint64_t blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(blocks, cap)));
}

struct Timer {
    cr_solver* s;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    bool on = false;
    bool in_launch = false;  // a and b are set by the one kernel launch timed (hipExtLaunchKernelGGL)
    Timer(cr_solver* s_, const char* n, bool one_launch = false) : s(s_), name(n), in_launch(one_launch) {
        on = s->timing && (s->time_only.empty() || s->time_only == n);
        if (on) {
            a = acquire_event(s);
            b = acquire_event(s);
            if (!in_launch) hipEventRecord(a, s->stream);
        }
    }
    void stop() {
        if (!on || in_launch) return;
        hipEventRecord(b, s->stream);
    }
    hipEvent_t start_event() const { return on ? a : nullptr; }
    hipEvent_t stop_event() const { return on ? b : nullptr; }
    // called after the stream has been synchronised
    void collect() {
        if (!on || a == nullptr) return;
        float ms = 0;
        hipEventElapsedTime(&ms, a, b);
        auto it = s->ktimes.find(name);
        if (it == s->ktimes.end()) {
            s->korder.push_back(name);
            it = s->ktimes.emplace(name, KTime{}).first;
        }
        it->second.ms += ms;
        it->second.launches += 1;
        s->events.push_back(a);
        s->events.push_back(b);
        a = b = nullptr;
    }
};

// A timer for the launch about to be enqueued, kept until the chunk syncs.
// one_launch: the timed region is a single kernel launch that records the
// timer's events itself (no event packets between launches)
Timer* start_timer(cr_solver* s, const char* name, bool one_launch = false) {
    s->timers->emplace_back(new Timer(s, name, one_launch));
    return s->timers->back().get();
}

template <int K, int W>
int launch_init(cr_solver* s, LevelArgs a) {
    if (int rc = raise_lds_limit(reinterpret_cast<const void*>(k_init<K, W>), s->gv.image_bytes)) return rc;
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(s->ngroups(), 2048));
    hipLaunchKernelGGL((k_init<K, W>), dim3(std::max<uint32_t>(grid, 1)), dim3(kBlock), s->gv.image_bytes, s->stream,
                       s->gv, a);
    hipLaunchKernelGGL(k_settle_init, dim3(1), dim3(kScanThreads), 0, s->stream, a);
    hipLaunchKernelGGL(k_write_lists, dim3(s->writer_grid()), dim3(kBlock), 0, s->stream, a, 1);
    return 0;
}

template <int K>
struct LaunchInit {
    static int run(cr_solver* s, LevelArgs a) {
        switch (s->W) {
            case 1: return launch_init<K, 1>(s, a);
            case 2: return launch_init<K, 2>(s, a);
            case 3: return launch_init<K, 3>(s, a);
            case 4: return launch_init<K, 4>(s, a);
            default: return fail(CR_ERR_ARG, "row width %d", s->W);
        }
    }
};

// Level kernels read their level and list lengths from the device, so their
// grids are sized by the tuning knobs, not by host-side counts.
template <int K>
struct LaunchPush {
    static int run(cr_solver* s, LevelArgs a) {
        const Timer* t = s->launch_timer;
        hipEvent_t e0 = t ? t->start_event() : nullptr, e1 = t ? t->stop_event() : nullptr;
        if (s->d_csr) {
            hipExtLaunchKernelGGL(k_cop_push_csr, dim3(s->push_blocks), dim3(kBlock), 0, s->stream, e0, e1, 0u,
                                  s->gv, a, static_cast<const unsigned long long*>(s->d_heads),
                                  static_cast<const uint32_t*>(s->d_csr));
            return 0;
        }
        if (int rc = raise_lds_limit(reinterpret_cast<const void*>(k_cop_push<K>), s->gv.image_bytes)) return rc;
        hipExtLaunchKernelGGL(k_cop_push<K>, dim3(s->push_blocks), dim3(kBlock), s->gv.image_bytes, s->stream, e0, e1,
                              0u, s->gv, a);
        return 0;
    }
};

// k_pull_tile1 exists for J >= 1; the J = 0 branch is never launched.
template <int J, int W>
auto pull_tile1_kernel() {
    if constexpr (J > 0)
        return k_pull_tile1<J, W>;
    else
        return k_pull_tile<1, 0, W>;
}

// Stages K-1 and K of one pull as one k_pull_pair launch (prepare_pair built
// its tile tables for this loop); `in` is H_{K-2}.
template <int K, int W>
int launch_pull_pair(cr_solver* s, LevelArgs a, const uint64_t* in) {
    if constexpr (K < 3) {
        return fail(CR_ERR_DEVICE, "k_pull_pair needs k >= 3");
    } else {
        SepStage st{};
        st.r = s->pair_r;
        st.fuse = a.fuse;
        st.clear = K == 3;  // this launch is stage 2, after stage 1 has read the frontier
        if (st.clear) s->stage2_cleared = true;
        if (st.fuse) hipLaunchKernelGGL(k_zero_groups, dim3(1024), dim3(kBlock), 0, s->stream, a);
        st.stage = K;
        st.in = in;
        st.out = st.fuse ? s->d_fr0 : s->d_nx0;
        st.cnt_a_in = static_cast<uint32_t>(s->ms_count[K - 2]);
        st.cnt_a = static_cast<uint32_t>(s->ms_count[K]);
        st.rows = s->ms_count[K];
        st.nR = s->pair_nR;
        st.tstart = s->d_pair;
        st.cstart = s->d_pair + s->pair_nR + 1;
        auto kern = k_pull_pair<K - 2, W>;
        const void* fn = reinterpret_cast<const void*>(kern);
        const uint32_t lds = ((s->gv.image_bytes + 15) & ~15u) + s->g.n * st.r * W * 8;
        if (int rc = raise_lds_limit(fn, lds)) return rc;
        const uint32_t blocks = std::max<uint32_t>(kXcds, s->pair_blocks / kXcds * kXcds);
        Timer* t = start_timer(s, "sep_stage");
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kPairBlock), lds, s->stream, s->gv, a, st);
        t->stop();
        return 0;
    }
}

// Stages I..K of one pull; stage I reads `in` (H_{I-1}) and writes H_I.
// The grid is a multiple of the XCD count (k_sep_stage splits its windows
// per XCD).
template <int K, int W, int I>
int launch_pull_stage(cr_solver* s, LevelArgs a, const uint64_t* in) {
    if constexpr (I > K) {
        return 0;
    } else {
        if constexpr (I == K - 1 && K >= 3) {
            if (s->pair_ready) return launch_pull_pair<K, W>(s, a, in);
        }
        SepStage st{};
        // k_pull_tile1 takes stage 1 of a multi-cop pull (r consecutive B per tile)
        const bool tile1 = I == 1 && K > 1 && s->tile1_rows > 0;
        st.r = tile1 ? s->tile1_rows : I == 1 ? 1u : s->tile_rows;
        // LDS of a tile: image and n * r rows; fewer rows per tile (or the row
        // kernel) where a large graph image leaves too little room
        auto tile_lds = [&](uint32_t r) { return ((s->gv.image_bytes + 15) & ~15u) + s->g.n * r * W * 8; };
        auto fits = [&](uint32_t r) {
            if (tile_lds(r) > kLdsBytes) return false;
            if (tile1)  // each thread loads kTileLoads rows v of one B_t
                return r * W <= static_cast<uint32_t>(kTile1Block) &&
                       s->g.n <= static_cast<uint32_t>(kTileLoads) * (kTile1Block / (r * W));
            if (r * W > static_cast<uint32_t>(kTileBlock)) return false;
            return s->g.n * r * W <= static_cast<uint32_t>(kTileLoads * kTileBlock);
        };
        while (st.r > 1 && !fits(st.r)) --st.r;
        const bool tiles = s->pull_tiles && fits(st.r);
        st.fuse = I == K && tiles && a.fuse;
        st.clear = I == 2 && tiles;
        if (st.clear) s->stage2_cleared = true;
        if (I == K && a.fuse && !st.fuse) return fail(CR_ERR_DEVICE, "fused finalize without the tiled last stage");
        if (st.fuse)
            hipLaunchKernelGGL(k_zero_groups, dim3(1024), dim3(kBlock), 0, s->stream, a);
        auto kern = !tiles ? k_sep_stage<I, K - I, W> : tile1 ? pull_tile1_kernel<K - I, W>() : k_pull_tile<I, K - I, W>;
        const void* fn = reinterpret_cast<const void*>(kern);
        st.stage = I;
        st.in = in;
        st.in_odd = I == 1 ? s->d_fr1[1] : nullptr;
        st.out = I < K ? s->d_sep[(I - 1) & 1] : st.fuse ? s->d_fr0 : s->d_nx0;
        st.cnt_b = static_cast<uint32_t>(s->ms_count[K - I]);
        st.cnt_b_in = static_cast<uint32_t>(s->ms_count[K - I + 1]);
        st.cnt_a = static_cast<uint32_t>(s->ms_count[I]);
        st.cnt_a_in = static_cast<uint32_t>(s->ms_count[I - 1]);
        st.rows = s->ms_count[I] * s->ms_count[K - I];
        uint32_t lds = s->gv.image_bytes;
        uint64_t blocks = 0;
        if (tiles) {
            lds = tile_lds(st.r);
            const uint64_t ntiles = tile1 ? (st.cnt_b + st.r - 1) / st.r
                                          : static_cast<uint64_t>(st.cnt_b) * ((st.cnt_a_in + st.r - 1) / st.r);
            // chunks of up to tile_chunk tiles, but at least ~2048 chunks, so that
            // small stages still spread over every CU (ER k = 3: 2-5 tiles)
            st.chunk = static_cast<uint32_t>(std::min<uint64_t>(s->tile_chunk, std::max<uint64_t>(1, (ntiles + 2047) / 2048)));
            blocks = std::min<uint64_t>((ntiles + st.chunk - 1) / st.chunk, s->tile_blocks);
        } else {
            blocks = std::min<uint64_t>(((st.rows + 63) / 64 + kWavesPerBlock - 1) / kWavesPerBlock, 16384);
        }
        if (int rc = raise_lds_limit(fn, lds)) return rc;
        blocks = std::max<uint64_t>(kXcds, (blocks + kXcds - 1) / kXcds * kXcds);
        Timer* t = start_timer(s, "sep_stage");
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(!tiles ? kBlock : tile1 ? kTile1Block : kTileBlock), lds, s->stream,
                           s->gv, a, st);
        t->stop();
        return launch_pull_stage<K, W, I + 1>(s, a, st.out);
    }
}

template <int K, int W>
int launch_pull(cr_solver* s, LevelArgs a) {
    s->stage2_cleared = false;
    if (int rc = launch_pull_stage<K, W, 1>(s, a, s->d_fr1[0])) return rc;
    if (s->stage2_cleared) return 0;  // the tiled stage 2 cleared the frontier
    Timer* t = start_timer(s, "clear_frontier");
    hipLaunchKernelGGL(k_clear_frontier, dim3(1024), dim3(kBlock), 0, s->stream, a, s->W);
    t->stop();
    return 0;
}

template <int K>
struct LaunchPull {
    static int run(cr_solver* s, LevelArgs a) {
        switch (s->W) {
            case 1: return launch_pull<K, 1>(s, a);
            case 2: return launch_pull<K, 2>(s, a);
            case 3: return launch_pull<K, 3>(s, a);
            case 4: return launch_pull<K, 4>(s, a);
            default: return fail(CR_ERR_ARG, "row width %d", s->W);
        }
    }
};

template <int W>
struct LaunchRobber {
    static int run(cr_solver* s, LevelArgs a) {
        const uint32_t lds = (2 * s->g.n + 16 * ((s->g.n + 3) / 4)) * W * 8;  // k_robber_step's LDS
        const uint32_t grid = (s->row_grid() + kRobberGroups - 1) / kRobberGroups;
        hipLaunchKernelGGL(k_robber_step<W>, dim3(grid), dim3(kRobberBlock), lds, s->stream, s->gv, a);
        return 0;
    }
};

template <int W>
struct LaunchFinalize {
    static int run(cr_solver* s, LevelArgs a) {
        hipLaunchKernelGGL(k_cop_finalize<W>, dim3(s->row_grid()), dim3(kBlock), 0, s->stream, s->gv, a);
        return 0;
    }
};

template <int W>
struct LaunchVerdict {
    static int run(cr_solver* s, LevelArgs a) {
        hipLaunchKernelGGL(k_verdict<W>, dim3(grid_for((s->C + 63) / 64)), dim3(kBlock), 0, s->stream, s->gv,
                           a, s->columns);
        return 0;
    }
};

template <int K>
struct LaunchRetro {
    static int run(cr_solver* s, LevelArgs a, uint32_t slots, int32_t mode, MoveArgs mv, unsigned long long* ghs,
                   uint32_t blocks) {
        const uint32_t lds = ((s->gv.image_bytes + 15) & ~15u) + (ghs ? 0u : slots * 8);
        uint32_t shift = 32;
        for (uint32_t v = slots; v > 1; v >>= 1) --shift;
        hipFuncSetAttribute(reinterpret_cast<const void*>(k_count_retro<K>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        hipLaunchKernelGGL(k_count_retro<K>, dim3(blocks), dim3(kWave), lds, s->stream, s->gv, a, slots, shift, mode,
                           mv, ghs);
        return 0;
    }
};

uint32_t log_chunk_rows(const cr_solver* s);

LevelArgs level_args(cr_solver* s) {
    LevelArgs a{};
    a.won0 = s->d_won0;
    a.won1 = s->d_won1;
    a.fr1s[0] = s->d_fr1[0];
    a.fr1s[1] = s->d_fr1[1];
    a.nx0 = s->d_nx0;
    a.table = s->d_table;
    a.log_rows = s->d_log_rows;
    a.log_bits = s->d_log_bits;
    if (s->d_starts) {
        a.log_R = log_chunk_rows(s);
        a.log_chunks = static_cast<uint32_t>((s->C + a.log_R - 1) / a.log_R);
        a.log_beg = s->d_starts;
        a.log_end = s->d_starts + static_cast<uint64_t>(a.log_chunks) * kMaxLogLists;
    }
    a.n = s->g.n;
    a.W = s->W;
    // fused finalize: the depth log, LDS tiles, and two robber-frontier
    // buffers (the robber step runs before the cop step)
    a.fuse = s->d_log_rows != nullptr && s->pull_tiles && s->fuse_finalize && s->fuse_ok && s->d_fr1[1] != s->d_fr1[0];
    for (int b = 0; b < 2; ++b) {
        a.list0[b] = s->d_list[0][b];
        a.list1[b] = s->d_list[1][b];
    }
    a.c0 = Compact{s->d_flags[0], s->d_gcnt[0], s->d_gstates[0]};
    a.c1 = Compact{s->d_flags[1], s->d_gcnt[1], s->d_gstates[1]};
    a.gcaps = s->d_gcaps;
    a.nbatch = s->nbatch();
    a.pull_rows = s->pull_rows;
    a.ctl = s->d_ctl;
    a.out_mask = s->d_outmask;
    a.in_mask = s->d_inmask;
    a.fr0 = s->d_fr0;
    a.filter = s->filter;
    a.open_factor = s->open_factor;
    // two-level level scan where the groups span more than one tile
    a.scan_tile = s->ngroups() > s->scan_tile ? s->scan_tile : 0u;
    a.tile_stride = s->tile_stride;
    a.tile_cnt = s->d_tile_cnt;
    a.tile_states = s->d_tile_states;
    a.dirty = s->push_dirty ? s->d_dirty : nullptr;
    a.lb_status = s->d_lb_status;
    a.lb_stride = s->lb_stride;
    return a;
}

// Builds the LDS image and uploads the graph tables.
int upload_graph(cr_solver* s) {
    const int n = s->g.n, K = s->k, M = n + K - 1;
    std::vector<uint8_t> img;
    auto align4 = [&] { while (img.size() % 4) img.push_back(0); };
    auto put32 = [&](uint32_t v) { const uint8_t* b = reinterpret_cast<const uint8_t*>(&v); img.insert(img.end(), b, b + 4); };
    for (int i = 0; i < K; ++i)
        for (int x = 0; x < n; ++x)
            put32(static_cast<uint32_t>(std::min<uint64_t>(cr::binom(M - 1 - (x + i), K - i), 0xffffffffu)));
    const uint32_t off_binom = static_cast<uint32_t>(img.size());
    for (int j = 0; j <= K; ++j)
        for (int x = 0; x <= M; ++x)
            put32(static_cast<uint32_t>(std::min<uint64_t>(cr::binom(x, j), 0xffffffffu)));
    const uint32_t off_in_off = static_cast<uint32_t>(img.size());
    uint32_t acc = 0;
    for (int v = 0; v <= n; ++v) {
        const uint16_t o = static_cast<uint16_t>(acc);
        img.push_back(static_cast<uint8_t>(o & 0xff));
        img.push_back(static_cast<uint8_t>(o >> 8));
        if (v < n) acc += static_cast<uint32_t>(s->gc.in[v].size());
    }
    align4();
    const uint32_t off_in_nbr = static_cast<uint32_t>(img.size());
    for (int v = 0; v < n; ++v) img.insert(img.end(), s->gc.in[v].begin(), s->gc.in[v].end());
    align4();
    const uint32_t off_out_off = static_cast<uint32_t>(img.size());
    acc = 0;
    for (int v = 0; v <= n; ++v) {
        const uint16_t o = static_cast<uint16_t>(acc);
        img.push_back(static_cast<uint8_t>(o & 0xff));
        img.push_back(static_cast<uint8_t>(o >> 8));
        if (v < n) acc += static_cast<uint32_t>(s->gc.out[v].size());
    }
    align4();
    const uint32_t off_out_nbr = static_cast<uint32_t>(img.size());
    for (int v = 0; v < n; ++v) img.insert(img.end(), s->gc.out[v].begin(), s->gc.out[v].end());
    align4();
    if (img.size() > 150 * 1024) return fail(CR_ERR_UNSUPPORTED, "graph image of %zu bytes exceeds LDS", img.size());
    int rc = 0;
    s->d_image = dev_alloc<uint8_t>(img.size(), s->device_bytes, rc);
    if (!s->d_image) return rc;
    HIP_TRY(hipMemcpy(s->d_image, img.data(), img.size(), hipMemcpyHostToDevice));
    const std::vector<uint64_t>& om = s->rmask;
    s->d_outmask = dev_alloc<uint64_t>(om.size(), s->device_bytes, rc);
    if (!s->d_outmask) return rc;
    HIP_TRY(hipMemcpy(s->d_outmask, om.data(), om.size() * 8, hipMemcpyHostToDevice));
    std::vector<uint64_t> im(om.size(), 0);  // transpose of the robber moves
    for (int r = 0; r < n; ++r)
        for (int v = 0; v < n; ++v)
            if ((om[r * 4 + v / 64] >> (v % 64)) & 1) im[v * 4 + r / 64] |= 1ull << (r % 64);
    s->d_inmask = dev_alloc<uint64_t>(im.size(), s->device_bytes, rc);
    if (!s->d_inmask) return rc;
    HIP_TRY(hipMemcpy(s->d_inmask, im.data(), im.size() * 8, hipMemcpyHostToDevice));
    s->gv.image = s->d_image;
    s->gv.image_bytes = static_cast<uint32_t>(img.size());
    s->gv.off_binom = off_binom;
    s->gv.off_in_off = off_in_off;
    s->gv.off_in_nbr = off_in_nbr;
    s->gv.off_out_off = off_out_off;
    s->gv.off_out_nbr = off_out_nbr;
    s->gv.n = n;
    s->gv.W = s->W;
    s->gv.M = M;
    s->gv.C = static_cast<uint32_t>(s->C);
    s->gv.ctot_m1 = static_cast<uint32_t>(s->C - 1);
    return 0;
}

int do_load(cr_solver* s, int32_t n, const uint8_t* adj, int32_t k, int32_t p) {
    s->release();
    s->stage = kEmpty;
    if (n <= 0) return fail(CR_ERR_EMPTY_GRAPH, "Graph is empty or failed to load.");
    if (n > CR_MAX_N) return fail(CR_ERR_GRAPH_LIMIT, "graph has %d vertices; the format limit is %d", n, CR_MAX_N);
    if (k <= 0 || k > CR_MAX_K) return fail(CR_ERR_K_RANGE, "k=%d outside 1..%d", k, CR_MAX_K);
    if (p < 1) return fail(CR_ERR_P_RANGE, "p=%d must be >= 1", p);
    s->g.build(n, adj);
    if (!s->cop_stay)  // with stay moves every cop has a move
        for (int v = 0; v < n; ++v)
            if (s->g.out[v].empty())
                return fail(CR_ERR_ISOLATED_VERTEX, "vertex %d has no move; the reference's move generation is undefined there", v);
    s->k = k;
    s->p = p;
    s->stage = kLoaded;
    return 0;
}

int build_csr(cr_solver* s);

// Break-even of the two cop-turn directions: a push costs about |pred| random
// relaxations per frontier row, the factorised pull about sum_i |H_i| * deg
// row reads per level regardless of the frontier.
void set_pull_threshold(cr_solver* s) {
    if (!s->d_sep[0]) {
        s->pull_rows = 0xffffffffu;
        return;
    }
    const int n = s->g.n, K = s->k;
    uint64_t edges = 0;
    for (int v = 0; v < n; ++v) edges += s->gc.out[v].size();
    const double deg = static_cast<double>(edges) / n;
    double pull_reads = 0;
    for (int i = 1; i <= K; ++i) pull_reads += static_cast<double>(s->ms_count[i]) * s->ms_count[K - i] * deg;
    const double pred = std::pow(std::max(1.0, static_cast<double>(edges) / n), K);
    // Measured on MI355X: a push costs ~4-6 ns per frontier row while the
    // won0/nx0 bitsets fit in the 256 MiB last-level cache and ~10 ns beyond,
    // a pull ~0.007 ns per row read. The defaults are the measured optima of
    // the scale on ER(200, 0.03) k = 3 (cached: 150-160) and planar(200) k = 4
    // (uncached: 150-190); pushes get cheaper per row as won0 fills, which
    // this row count does not model.
    const bool cached = 2ull * s->C * s->W * 8 <= (256ull << 20);
    const int64_t scale = s->pull_scale >= 0 ? s->pull_scale : (cached ? 150 : 190);
    const double rows = pull_reads / pred * static_cast<double>(scale) / 1000.0;
    s->pull_rows = rows >= 4.0e9 ? 0xffffffffu : static_cast<uint32_t>(rows);
}

// Intermediates of the factorised pull: H_1..H_{K-1} alternate between two
// buffers sized for the largest of each parity. Skipped (the push then runs
// every level) when they do not fit next to the state space.
int alloc_pull(cr_solver* s) {
    const int n = s->g.n, K = s->k;
    for (int j = 0; j <= K; ++j) s->ms_count[j] = cr::binom(n + j - 1, j);
    uint64_t need[2] = {0, 0};
    for (int i = 1; i < K; ++i) {
        const unsigned __int128 rows = static_cast<unsigned __int128>(s->ms_count[i]) * s->ms_count[K - i];
        if (rows >= (static_cast<unsigned __int128>(1) << 32)) return 0;
        need[(i - 1) & 1] = std::max<uint64_t>(need[(i - 1) & 1], static_cast<uint64_t>(rows) * s->W * 8);
    }
    for (int j = 0; j <= K; ++j)
        if (s->ms_count[j] >= 0xffffffffull) return 0;
    size_t freeb = 0, totalb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    if (need[0] + need[1] + (1ull << 30) > freeb) return 0;
    int rc = 0;
    if (!(s->d_sep[0] = dev_alloc<uint64_t>(std::max<uint64_t>(need[0] / 8, 1), s->device_bytes, rc))) return rc;
    if (!(s->d_sep[1] = dev_alloc<uint64_t>(std::max<uint64_t>(need[1] / 8, 1), s->device_bytes, rc))) return rc;
    set_pull_threshold(s);
    return 0;
}

// Tile tables of k_pull_pair for the current graph and pair_rows: A-ranges
// R of r consecutive (K-2)-multisets, each with tiles c = (least max A over
// R) .. n - 1, numbered R-major. Built on the host before a main loop; the
// kernel is used where the tiled stages are and its LDS tile fits.
int prepare_pair(cr_solver* s) {
    const int n = s->g.n, K = s->k;
    const uint32_t img = (s->gv.image_bytes + 15) & ~15u;
    uint32_t r = std::min<uint32_t>(std::max<uint32_t>(s->pair_rows, 1), kPairMaxRows);
    const uint32_t ni = s->W % 2 == 0 ? s->W / 2 : s->W;  // phase-1 items per (A, b)
    while (r > 1 && (img + static_cast<uint64_t>(n) * r * s->W * 8 > kLdsBytes ||
                     static_cast<uint64_t>(n) * r * ni > static_cast<uint64_t>(kPairItems) * kPairBlock))
        --r;
    const bool usable = (s->pull_pair == 2 || (s->pull_pair == 1 && K == 3)) && s->pull_tiles && s->d_sep[0] != nullptr && K >= 3 &&
                        img + static_cast<uint64_t>(n) * r * s->W * 8 <= kLdsBytes &&
                        static_cast<uint64_t>(n) * r * ni <= static_cast<uint64_t>(kPairItems) * kPairBlock &&
                        s->ms_count[K - 2] <= (1u << 24);
    s->pair_ready = false;
    if (!usable) return 0;
    if (s->pair_r == r) {
        s->pair_ready = true;
        return 0;
    }
    const uint32_t cnt = static_cast<uint32_t>(s->ms_count[K - 2]);
    const uint32_t nR = (cnt + r - 1) / r;
    std::vector<uint32_t> h(2ull * nR + 1);
    std::vector<uint32_t> A(K - 2, 0u);
    uint64_t total = 0;
    for (uint32_t R = 0; R < nR; ++R) {
        uint32_t cmin = static_cast<uint32_t>(n);
        for (uint32_t t = 0; t < r && R * r + t < cnt; ++t) {
            cmin = std::min<uint32_t>(cmin, A[K - 3]);
            // lexicographic successor of the non-decreasing tuple A
            int i = K - 3;
            while (i >= 0 && A[i] + 1 >= static_cast<uint32_t>(n)) --i;
            if (i >= 0) {
                const uint32_t v = A[i] + 1;
                for (int j = i; j < K - 2; ++j) A[j] = v;
            }
        }
        h[R] = static_cast<uint32_t>(total);
        h[nR + 1 + R] = cmin;
        total += static_cast<uint32_t>(n) - cmin;
    }
    if (total >= 0xffffffffull) return 0;  // stage by stage instead
    h[nR] = static_cast<uint32_t>(total);
    if (s->pair_cap < h.size()) {
        if (s->d_pair) hipFree(s->d_pair);
        s->d_pair = nullptr;
        s->pair_cap = 0;
        int rc = 0;
        if (!(s->d_pair = dev_alloc<uint32_t>(h.size(), s->device_bytes, rc))) return rc;
        s->pair_cap = h.size();
    }
    HIP_TRY(hipMemcpyAsync(s->d_pair, h.data(), h.size() * 4, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->pair_r = r;
    s->pair_nR = nR;
    s->pair_ready = true;
    return 0;
}

// Depth log: up to log_factor * C entries of (configuration, W robber
// words), within half of the free device memory (none below 2 * C entries,
// and none for fewer than log_min_configs configurations: small tables take
// their few bytes per depth directly). A depth whose lists do not fit in the
// remaining room writes its bytes directly (k_log_lists).

int alloc_log(cr_solver* s) {
    if (!s->log_factor || s->C < s->log_min_configs) return 0;
    const uint64_t entry = 4 + 8ull * s->W;
    int rc = 0;
    // the second robber-frontier buffer lets the robber step run before the
    // cop step, which the fused finalize of logged solves needs
    if (!(s->d_fr1[1] = dev_alloc<uint64_t>(s->C * s->W, s->device_bytes, rc))) return rc;
    size_t freeb = 0, totalb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    const uint64_t cap = std::min<uint64_t>(static_cast<uint64_t>(s->log_factor) * s->C, freeb / 2 / entry);
    if (cap < 2 * s->C) return 0;
    if (!(s->d_log_rows = dev_alloc<uint32_t>(cap, s->device_bytes, rc))) return rc;
    if (!(s->d_log_bits = dev_alloc<uint64_t>(cap * s->W, s->device_bytes, rc))) return rc;
    if (!(s->d_lists = dev_alloc<LogList>(kMaxLogLists, s->device_bytes, rc))) return rc;
    const uint64_t nchunks = (s->C + log_chunk_rows(s) - 1) / log_chunk_rows(s);
    if (!(s->d_starts = dev_alloc<uint32_t>(2 * nchunks * kMaxLogLists, s->device_bytes, rc))) return rc;
    if (!s->h_lists)
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&s->h_lists), kMaxLogLists * sizeof(LogList), hipHostMallocDefault));
    s->log_cap = cap;
    s->fuse_ok = ((s->gv.image_bytes + 15) & ~15u) + static_cast<uint64_t>(s->g.n) * s->W * 8 <= kLdsBytes &&
                 static_cast<uint32_t>(s->g.n * s->W) <= static_cast<uint32_t>(kTileLoads * kTileBlock);
    return 0;
}

// Rows per chunk of k_apply_log (a multiple of 8, 24 KiB of LDS).
uint32_t log_chunk_rows(const cr_solver* s) {
    return std::max<uint32_t>(8, std::min<uint32_t>(8192, 24 * 1024 / (2 * s->g.n))) & ~7u;
}

template <int W>
struct LaunchApplyLog {
    static int run(cr_solver* s, int32_t nl, uint64_t total) {
        const uint32_t n = s->g.n, R = log_chunk_rows(s);
        const uint32_t lds = R * 2 * n;
        const uint32_t nchunks = static_cast<uint32_t>((s->C + R - 1) / R);
        const uint64_t cells = static_cast<uint64_t>(nchunks) * nl;
        uint32_t* beg = s->d_starts;
        uint32_t* end = s->d_starts + static_cast<uint64_t>(nchunks) * kMaxLogLists;
        const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(nchunks, 65536));
        if (int rc = raise_lds_limit(reinterpret_cast<const void*>(k_apply_log<W>), lds)) return rc;
        hipLaunchKernelGGL(k_apply_log<W>, dim3(grid), dim3(kBlock), lds, s->stream, s->d_lists, nl, s->d_log_rows,
                           s->d_log_bits, beg, end, s->d_table, n, static_cast<uint32_t>(s->C), R);
        // cleared for the lists of the next segment (k_log_lists sets only chunks with entries)
        hipLaunchKernelGGL(k_fill_u32, dim3(static_cast<uint32_t>(std::min<uint64_t>((cells + kBlock - 1) / kBlock, 4096))),
                           dim3(kBlock), 0, s->stream, end, cells, 0u);
        return 0;
    }
};

template <int W>
struct LaunchLogLists {
    static int run(cr_solver* s, LevelArgs a) {
        hipLaunchKernelGGL(k_log_lists<W>, dim3(s->log_grid()), dim3(kBlock), 0, s->stream, a);
        return 0;
    }
};

// Applies the logged depths [log_first, level] to the table and empties the
// log, when the loop has ended (force) or the log may not hold the next
// levels_per_sync depths of the size seen so far. h_ctl is current.
int flush_log(cr_solver* s, bool force) {
    if (!s->log_cap) return 0;
    const Ctl* h = s->h_ctl;
    const uint32_t last = std::min<uint32_t>(h->level, 127);
    uint64_t peak = 0;
    for (uint32_t d = s->log_first; d <= last; ++d)
        peak = std::max<uint64_t>(peak, static_cast<uint64_t>(h->log_len[d][0]) + h->log_len[d][1]);
    if (!force && h->log_used + std::max<uint64_t>(peak, 1) * s->levels_per_sync <= s->log_cap) return 0;
    // list L = 2 (d - log_first) + turn, as k_log_lists indexed its chunk ranges
    // (lists written straight to the table are empty here)
    const int32_t nl = static_cast<int32_t>(2 * (last - s->log_first + 1));
    if (nl > kMaxLogLists) return fail(CR_ERR_DEVICE, "depth log: %d lists", nl);
    uint64_t total = 0;
    for (uint32_t d = s->log_first; d <= last; ++d)
        for (int t = 0; t < 2; ++t) {
            s->h_lists[2 * (d - s->log_first) + t] =
                LogList{h->log_off[d][t], h->log_len[d][t], static_cast<uint32_t>(t << 8) | static_cast<uint32_t>(1u | (d << 1))};
            total += h->log_len[d][t];
        }
    if (total) {
        HIP_TRY(hipMemcpyAsync(s->d_lists, s->h_lists, nl * sizeof(LogList), hipMemcpyHostToDevice, s->stream));
        Timer* t = start_timer(s, "apply_log");
        if (int rc = dispatch_w<LaunchApplyLog>(s->W, s, nl, total)) return rc;
        t->stop();
    }
    static const unsigned long long zero = 0;
    s->log_first = last + 1;
    HIP_TRY(hipMemcpyAsync(&s->d_ctl->log_used, &zero, sizeof zero, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(&s->d_ctl->log_first, &s->log_first, sizeof s->log_first, hipMemcpyHostToDevice, s->stream));
    // the next segment's lists start from cleared chunk ends
    HIP_TRY(hipStreamSynchronize(s->stream));  // the copies read host memory that changes after return
    return 0;
}

// Limited-visibility variant: one column of states, the marked plane (won0)
// and the cop masks; the trajectory DAG is built by the main loop.
int build_macro(cr_solver* s) {
    const int n = s->g.n;
    if (s->cop_stay || s->robber_stay)
        return fail(CR_ERR_UNSUPPORTED, "stay moves are not part of the limited-visibility rules (the robber must move, k_cops_visibility_v2.cpp:226)");
    s->columns = 1;
    s->S = s->C * static_cast<uint64_t>(n);
    s->W = (n + 63) / 64;
    const uint64_t plane = s->C * s->W;
    size_t freeb = 0, totalb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    const uint64_t need = s->S + 2 * plane * 8 + (4 << 20);
    if (need > freeb) return fail(CR_ERR_TOO_LARGE, "state space needs %llu bytes; %zu free", static_cast<unsigned long long>(need), freeb);
    int rc = 0;
    if (!(s->d_table = dev_alloc<uint8_t>(s->S, s->device_bytes, rc))) return rc;
    if (!(s->d_won0 = dev_alloc<uint64_t>(plane, s->device_bytes, rc))) return rc;
    if (!(s->d_occ = dev_alloc<uint64_t>(plane, s->device_bytes, rc))) return rc;
    if (!(s->d_ctl = dev_alloc<Ctl>(1, s->device_bytes, rc))) return rc;
    if (!s->h_ctl) HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&s->h_ctl), sizeof(Ctl), hipHostMallocDefault));
    if ((rc = upload_graph(s)) != 0) return rc;
    s->stage = kBuilt;
    return 0;
}

// Column solver with p >= 2: the 2p-column table and the p-step reach masks;
// the planes are allocated by initializeCaptures.
int build_columns(cr_solver* s) {
    const int n = s->g.n;
    if (2 * s->p > 126) return fail(CR_ERR_UNSUPPORTED, "p=%d: the device column solver keeps at most 126 columns", s->p);
    s->columns = 2 * s->p;
    s->S = s->C * static_cast<uint64_t>(n) * s->columns;
    s->W = (n + 63) / 64;
    size_t freeb = 0, totalb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    if (s->S + (4 << 20) > freeb) return fail(CR_ERR_TOO_LARGE, "state space needs %llu bytes; %zu free", static_cast<unsigned long long>(s->S), freeb);
    int rc = 0;
    if (!(s->d_table = dev_alloc<uint8_t>(s->S, s->device_bytes, rc))) return rc;
    if (!(s->d_ctl = dev_alloc<Ctl>(1, s->device_bytes, rc))) return rc;
    if (!s->h_ctl) HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&s->h_ctl), sizeof(Ctl), hipHostMallocDefault));
    if ((rc = upload_graph(s)) != 0) return rc;
    // exactly-p-step reach sets (k_cops_visibility.cpp:235-288): robber moves along out-edges
    s->reach.assign(static_cast<size_t>(n) * 4, 0);
    for (int r = 0; r < n; ++r) {
        uint64_t cur[4] = {0, 0, 0, 0};
        cur[r >> 6] = 1ull << (r & 63);
        for (int step = 0; step < s->p; ++step) {
            uint64_t nxt[4] = {0, 0, 0, 0};
            for (int v = 0; v < n; ++v)
                if ((cur[v >> 6] >> (v & 63)) & 1)
                    for (int w = 0; w < 4; ++w) nxt[w] |= s->rmask[v * 4 + w];  // robber moves (+ stay)
            std::memcpy(cur, nxt, sizeof cur);
        }
        std::memcpy(&s->reach[static_cast<size_t>(r) * 4], cur, sizeof cur);
    }
    s->stage = kBuilt;
    return 0;
}

// Move sets under the solver's rules: cop moves add the cop's own vertex when
// cops may stay (AuxGraph.h:203-205), robber moves add r when the robber may
// stay (k_cops_visibility.cpp:253-255).
int derive_moves(cr_solver* s) {
    const int n = s->g.n;
    std::vector<uint8_t> m = s->g.matrix;
    if (s->cop_stay)
        for (int v = 0; v < n; ++v) m[static_cast<size_t>(v) * n + v] = 1;
    s->gc.build(n, m.data());
    for (int v = 0; v < n; ++v)
        if (s->gc.out[v].empty())
            return fail(CR_ERR_ISOLATED_VERTEX, "vertex %d has no move; the reference's move generation is undefined there", v);
    s->rmask.assign(static_cast<size_t>(n) * 4, 0);
    for (int u = 0; u < n; ++u) {
        for (uint8_t v : s->g.out[u]) s->rmask[u * 4 + (v >> 6)] |= 1ull << (v & 63);
        if (s->robber_stay) s->rmask[u * 4 + (u >> 6)] |= 1ull << (u & 63);
    }
    return 0;
}

int do_build(cr_solver* s) {
    if (s->stage < kLoaded) return fail(CR_ERR_ORDER, "buildAuxGraph before loadGraphFile");
    s->release();
    const int n = s->g.n;
    if (int rc = derive_moves(s)) return rc;
    s->C = cr::config_count(n, s->k);
    if (s->C == 0) return fail(CR_ERR_K_RANGE, "no configuration for k=%d", s->k);
    if (s->C >= 0xffffffffull) return fail(CR_ERR_TOO_LARGE, "%llu configurations exceed 32-bit ids", static_cast<unsigned long long>(s->C));
    uint64_t prod = 1, prod_out = 1;
    for (int i = 0; i < s->k; ++i) {
        prod *= static_cast<uint64_t>(std::max(1, s->gc.max_in));
        prod_out *= static_cast<uint64_t>(std::max(1, s->gc.max_out));
    }
    if (prod >= (1ull << 32) || prod_out >= (1ull << 32))
        return fail(CR_ERR_TOO_LARGE, "team-move product %llu too large", static_cast<unsigned long long>(std::max(prod, prod_out)));
    if (s->variant == CR_VARIANT_MACRO) return build_macro(s);
    if (s->p > 1) return build_columns(s);
    s->columns = 2 * s->p;
    s->S = s->C * static_cast<uint64_t>(n) * s->columns;
    s->W = (n + 63) / 64;
    const uint64_t plane = s->C * s->W;
    const uint64_t need = s->S + 5 * plane * 8 + 4 * s->C * 4 + 2 * s->C / 4 + (4 << 20);
    size_t freeb = 0, totalb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    if (need > freeb) return fail(CR_ERR_TOO_LARGE, "state space needs %llu bytes; %zu free", static_cast<unsigned long long>(need), freeb);
    int rc = 0;
    if (!(s->d_table = dev_alloc<uint8_t>(s->S, s->device_bytes, rc))) return rc;
    if (!(s->d_won0 = dev_alloc<uint64_t>(plane, s->device_bytes, rc))) return rc;
    if (!(s->d_won1 = dev_alloc<uint64_t>(plane, s->device_bytes, rc))) return rc;
    if (!(s->d_fr1[0] = dev_alloc<uint64_t>(plane, s->device_bytes, rc))) return rc;
    s->d_fr1[1] = s->d_fr1[0];  // a second buffer only with the depth log (alloc_log)
    if (!(s->d_fr0 = dev_alloc<uint64_t>(plane, s->device_bytes, rc))) return rc;
    if (!(s->d_nx0 = dev_alloc<uint64_t>(plane, s->device_bytes, rc))) return rc;
    for (auto& t : s->d_list)
        for (auto& b : t)
            if (!(b = dev_alloc<uint32_t>(s->C, s->device_bytes, rc))) return rc;
    const uint64_t ngroups = (s->nbatch() + kGroupWords - 1) / kGroupWords;
    for (int i = 0; i < 2; ++i) {
        if (!(s->d_flags[i] = dev_alloc<uint64_t>(s->nbatch(), s->device_bytes, rc))) return rc;
        if (!(s->d_gcnt[i] = dev_alloc<uint32_t>(ngroups, s->device_bytes, rc))) return rc;
        if (!(s->d_gstates[i] = dev_alloc<unsigned long long>(ngroups, s->device_bytes, rc))) return rc;
    }
    if (!(s->d_gcaps = dev_alloc<unsigned long long>(ngroups, s->device_bytes, rc))) return rc;
    s->tile_stride = static_cast<uint32_t>(ngroups / kScanTileMin + 1);  // tiles at the smallest tile size
    if (!(s->d_tile_cnt = dev_alloc<uint32_t>(2ull * s->tile_stride, s->device_bytes, rc))) return rc;
    if (!(s->d_dirty = dev_alloc<uint8_t>(ngroups, s->device_bytes, rc))) return rc;
    s->lb_stride = static_cast<uint32_t>(ngroups / kEndTileGroups + 1);
    if (!(s->d_lb_status = dev_alloc<unsigned long long>(2ull * s->lb_stride, s->device_bytes, rc))) return rc;
    if (!(s->d_tile_states = dev_alloc<unsigned long long>(2ull * s->tile_stride, s->device_bytes, rc))) return rc;
    if (!(s->d_ctl = dev_alloc<Ctl>(1, s->device_bytes, rc))) return rc;
    if (!s->h_ctl) HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&s->h_ctl), sizeof(Ctl), hipHostMallocDefault));
    if ((rc = upload_graph(s)) != 0) return rc;
    if (s->use_csr && (rc = build_csr(s)) != 0) return rc;
    if ((rc = alloc_pull(s)) != 0) return rc;
    if ((rc = alloc_log(s)) != 0) return rc;
    s->stage = kBuilt;
    return 0;
}

// zeroes the control block with a kernel-free host copy (never captured)
int reset_ctl(cr_solver* s) {
    std::memset(s->h_ctl, 0, sizeof(Ctl));
    s->h_ctl->verdict = ~0ull;
    s->h_ctl->log_cap = s->log_cap;
    s->log_first = 0;
    HIP_TRY(hipMemcpyAsync(s->d_ctl, s->h_ctl, sizeof(Ctl), hipMemcpyHostToDevice, s->stream));
    if (s->d_lb_status)  // tile sums of the last solve carry the same level epochs
        HIP_TRY(hipMemsetAsync(s->d_lb_status, 0, 2ull * s->lb_stride * sizeof(unsigned long long), s->stream));
    if (s->d_starts) {  // chunk ends of the depth log's lists (a solve that stopped early may leave some)
        const LevelArgs a = level_args(s);
        const uint64_t cells = static_cast<uint64_t>(a.log_chunks) * kMaxLogLists;
        hipLaunchKernelGGL(k_fill_u32, dim3(static_cast<uint32_t>(std::min<uint64_t>((cells + kBlock - 1) / kBlock, 4096))),
                           dim3(kBlock), 0, s->stream, a.log_end, cells, 0u);
    }
    return 0;
}

void add_ktime(void* ctx, const char* name, double ms) {
    cr_solver* s = static_cast<cr_solver*>(ctx);
    auto it = s->ktimes.find(name);
    if (it == s->ktimes.end()) {
        s->korder.push_back(name);
        it = s->ktimes.emplace(name, KTime{}).first;
    }
    it->second.ms += ms;
    it->second.launches += 1;
}

crd::TimeSink time_sink(cr_solver* s) {
    crd::TimeSink t;
    if (s->timing) {
        t.ctx = s;
        t.add = add_ktime;
    }
    return t;
}

int init_columns(cr_solver* s, uint64_t* base) {
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s->stream));
    uint64_t caps = 0;
    const int rc = crd::columns_init(&s->colplan, s->gv, s->reach, s->k, s->p, s->stream, s->d_table, s->device_bytes, &caps);
    HIP_TRY(hipEventRecord(e1, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    s->stats = cr_stats{};
    s->stats.init_ms = ms;
    s->stats.base_captures = caps;
    if (base) *base = caps;
    s->stage = kInitialized;
    return 0;
}

int loop_columns(cr_solver* s) {
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s->stream));
    uint64_t won = 0;
    int32_t levels = 0;
    const int rc = crd::columns_solve(s->colplan, time_sink(s), s->pass_counts, &won, &levels);
    HIP_TRY(hipEventRecord(e1, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    s->stats.loop_ms = ms;
    s->stats.won_states = won;
    s->stats.n_levels = levels;
    s->stats.n_passes = static_cast<int32_t>(s->pass_counts.size());
    s->stage = kSolved;
    return 0;
}

int init_macro(cr_solver* s, uint64_t* base) {
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s->stream));
    uint64_t caps = 0;
    int rc = crd::macro_init(s->gv, s->k, s->stream, s->d_table, s->d_occ, &caps);
    HIP_TRY(hipEventRecord(e1, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    s->stats = cr_stats{};
    s->stats.init_ms = ms;
    s->stats.base_captures = caps;
    if (base) *base = caps;
    s->stage = kInitialized;
    return 0;
}

int do_init(cr_solver* s, uint64_t* base) {
    if (s->stage < kBuilt) return fail(CR_ERR_ORDER, "initializeCaptures before buildAuxGraph");
    s->from_blob = false;
    int rc = reset_ctl(s);
    if (rc) return rc;
    if (s->variant == CR_VARIANT_MACRO) return init_macro(s, base);
    if (s->p > 1) return init_columns(s, base);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s->stream));
    // init resolves depths 0 and 1; the depth-1 lists go to parity 1, which
    // level d = 1 reads
    LevelArgs a = level_args(s);
    Timer t(s, "init");
    if ((rc = dispatch_k<LaunchInit>(s->k, s, a)) != 0) return rc;
    t.stop();
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e1, s->stream));
    HIP_TRY(hipMemcpyAsync(s->h_ctl, s->d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    t.collect();
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    s->stats = cr_stats{};
    s->stats.init_ms = ms;
    s->stats.pred_entries = s->csr_entries;
    s->stats.build_ms = s->csr_build_ms;
    s->stats.base_captures = s->h_ctl->lvl[0][0];
    if (base) *base = s->stats.base_captures;
    s->stage = kInitialized;
    return 0;
}

void passes_from_levels(cr_solver* s) {
    const auto& lv = s->h_ctl->lvl;
    s->pass_counts.clear();
    for (int i = 1;; ++i) {
        uint64_t c = 0;
        const int c0 = 2 * i - 2, c1 = 2 * i - 1, r0 = 2 * i - 1, r1 = 2 * i;
        if (c0 > 0 && c0 < 127) c += lv[c0][0];
        if (c1 < 127) c += lv[c1][0];
        if (r0 < 127) c += lv[r0][1];
        if (r1 < 127) c += lv[r1][1];
        s->pass_counts.push_back(c);
        if (c == 0) break;
    }
}

int loop_macro(cr_solver* s) {
    int rc = crd::macro_solve(&s->macro, s->gv, s->d_outmask, s->k, s->p, s->stream, s->d_table, s->d_occ, s->d_won0,
                              s->device_bytes, time_sink(s), s->pass_counts, s->mprof);
    if (rc) return rc;
    s->stats.loop_ms = s->mprof.build_ms + s->mprof.passes_ms;
    s->stats.n_passes = static_cast<int32_t>(s->pass_counts.size());
    s->stats.n_levels = s->stats.n_passes;
    uint64_t won = s->stats.base_captures;
    for (uint64_t c : s->pass_counts) won += c;
    s->stats.won_states = won;
    s->stage = kSolved;
    return 0;
}

int do_loop(cr_solver* s) {
    if (s->stage < kInitialized || s->from_blob) return fail(CR_ERR_ORDER, "mainLoop before initializeCaptures");
    s->ktimes.clear();
    s->korder.clear();
    if (s->variant == CR_VARIANT_MACRO) return loop_macro(s);
    if (s->p > 1) return loop_columns(s);
    if (int rc = prepare_pair(s)) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s->stream));
    LevelArgs a = level_args(s);
    int rc = 0;
    bool done = s->h_ctl->done != 0;
    // Levels are enqueued in chunks without host synchronisation: every level
    // kernel reads its depth and list lengths from the control block, and
    // k_advance ends the level (or flags completion, after which the queued
    // kernels return at once).
    while (!done) {
        std::vector<std::unique_ptr<Timer>> timers;
        s->timers = &timers;
        // A chunk of levels that starts with a robber frontier far below the
        // pull threshold pushes every level and enqueues no pull launches
        // (on push levels they only return, at several microseconds each):
        // its k_advance sees an unreachable threshold. Any level already set
        // to pull keeps its pull.
        LevelArgs la = a;
        bool pulls = s->d_sep[0] != nullptr;
        if (pulls && s->pull_gate > 0) {
            const Ctl* h = s->h_ctl;
            const uint64_t frontier = h->cnt1[h->level & 1];
            pulls = h->pull != 0 || frontier * static_cast<uint64_t>(s->pull_gate) > s->pull_rows;
        }
        if (!pulls) la.pull_rows = 0xffffffffu;
        const uint32_t first_level = s->h_ctl->level;  // each enqueued level advances one (or returns after the last)
        for (int l = 0; l < s->levels_per_sync; ++l) {
            // The robber step (depth d + 1 robber-turn states from cop-turn
            // states of depth <= d) and the cop step (depth d + 1 cop-turn
            // states from the depth-d robber frontier) read only depths <= d.
            // Where the last pull stage writes fr0 itself (fused finalize),
            // the robber step goes first; otherwise the cop step does, right
            // after the lists that made its frontier (still in the caches on
            // small tables), and k_cop_finalize settles after both.
            Timer* t = nullptr;
            if (la.fuse) {
                t = start_timer(s, "robber_step");
                if ((rc = dispatch_w<LaunchRobber>(s->W, s, la)) != 0) return rc;
                t->stop();
            }
            t = start_timer(s, "cop_push", true);
            s->launch_timer = t;
            rc = dispatch_k<LaunchPush>(s->k, s, la);
            s->launch_timer = nullptr;
            if (rc) return rc;
            if (pulls && (rc = dispatch_k<LaunchPull>(s->k, s, la)) != 0) return rc;
            if (!la.fuse) {
                t = start_timer(s, "robber_step");
                if ((rc = dispatch_w<LaunchRobber>(s->W, s, la)) != 0) return rc;
                t->stop();
            }
            t = start_timer(s, "cop_finalize");
            if ((rc = dispatch_w<LaunchFinalize>(s->W, s, la)) != 0) return rc;
            t->stop();
            if (!la.scan_tile && s->end_pass) {  // one pass: scan, level end and lists
                const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(2ull * s->lb_stride, 1024));
                hipLaunchKernelGGL(k_end_level, dim3(grid), dim3(kBlock), 0, s->stream, la, first_level + l);
            } else {
                if (la.scan_tile) {
                    const uint32_t T = static_cast<uint32_t>((s->ngroups() + la.scan_tile - 1) / la.scan_tile);
                    hipLaunchKernelGGL(k_scan_tiles, dim3(2 * T), dim3(kScanThreads), 0, s->stream, la, T);
                }
                hipLaunchKernelGGL(k_advance, dim3(1), dim3(kScanThreads), 0, s->stream, la);
                hipLaunchKernelGGL(k_write_lists, dim3(s->writer_grid()), dim3(kBlock), 0, s->stream, la, 0);
            }
            if (s->log_cap && (rc = dispatch_w<LaunchLogLists>(s->W, s, la)) != 0) return rc;
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(s->h_ctl, s->d_ctl, offsetof(Ctl, lvl), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        done = s->h_ctl->done != 0;
        if ((rc = flush_log(s, done)) != 0) return rc;
        HIP_TRY(hipGetLastError());
        if (s->timing) HIP_TRY(hipStreamSynchronize(s->stream));
        for (auto& t : timers) t->collect();
        s->timers = nullptr;
    }
    HIP_TRY(hipEventRecord(e1, s->stream));
    HIP_TRY(hipMemcpyAsync(s->h_ctl, s->d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    s->stats.loop_ms = ms;
    s->stats.n_levels = static_cast<int32_t>(s->h_ctl->level);
    uint64_t won = 0;
    for (int d = 0; d < 127; ++d) won += s->h_ctl->lvl[d][0] + s->h_ctl->lvl[d][1];
    s->stats.won_states = won;
    passes_from_levels(s);
    s->stats.n_passes = static_cast<int32_t>(s->pass_counts.size());
    s->stage = kSolved;
    return 0;
}

// A temporary device allocation, freed when it goes out of scope (hipFree
// waits for the work that uses it).
struct DeviceBuffer {
    void* p = nullptr;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() {
        if (p) hipFree(p);
    }
    int alloc(uint64_t bytes) {
        if (hipMalloc(&p, bytes) != hipSuccess) {
            p = nullptr;
            return fail(CR_ERR_TOO_LARGE, "%llu bytes of device scratch", static_cast<unsigned long long>(bytes));
        }
        return 0;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

int count_moves(cr_solver* s, int32_t mode, unsigned long long* r, MoveArgs mv = MoveArgs{}) {
    HIP_TRY(hipSetDevice(s->device));
    uint64_t prod = 1;
    const int deg = std::max(1, mode == 1 ? s->gc.max_out : s->gc.max_in);
    for (int i = 0; i < s->k; ++i) prod *= static_cast<uint64_t>(deg);
    uint32_t slots = 64;
    while (slots < 2 * prod) slots <<= 1;
    const uint64_t lds = ((s->gv.image_bytes + 15) & ~15u) + static_cast<uint64_t>(slots) * 8;
    uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>(s->C, 8192));
    DeviceBuffer ghs;  // the dedup sets in device memory when they do not fit in LDS (counts only)
    if (lds > 160 * 1024) {
        // the CSR build (modes 2, 3) is an optimisation and goes without
        if (mode >= 2 || slots > (1u << 24))
            return fail(CR_ERR_UNSUPPORTED, "move count needs %llu bytes of LDS", static_cast<unsigned long long>(lds));
        blocks = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(blocks, (1ull << 30) / (slots * 8ull))));
        if (int rc = ghs.alloc(static_cast<uint64_t>(blocks) * slots * 8)) return rc;
    }
    static const unsigned long long zero[256] = {};
    HIP_TRY(hipMemcpyAsync(s->d_ctl->retro, zero, 3 * sizeof(unsigned long long), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_ctl->hist, zero, sizeof zero, hipMemcpyHostToDevice, s->stream));
    LevelArgs a = level_args(s);
    if (int rc = dispatch_k<LaunchRetro>(s->k, s, a, slots, mode, mv, ghs.as<unsigned long long>(), blocks)) return rc;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(r, s->d_ctl->retro, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return 0;
}

// Predecessor lists of every configuration, deduplicated, as a CSR of u32
// ranks (the device form of AuxGraph::createTransitions, AuxGraph.h:179-275,
// over in-edges). Skipped -- the push then enumerates team moves on the fly --
// when it does not fit in device memory or the dedup table exceeds LDS.
int build_csr(cr_solver* s) {
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s->stream));
    uint64_t tmp_bytes = 0;
    int rc = 0;
    uint32_t* cnt = dev_alloc<uint32_t>(s->C, tmp_bytes, rc);
    if (!cnt) return rc;
    unsigned long long r[3];
    rc = count_moves(s, 2, r, MoveArgs{cnt, nullptr, nullptr});
    if (rc == CR_ERR_UNSUPPORTED) {  // dedup table does not fit in LDS: no CSR
        hipFree(cnt);
        g_err[0] = '\0';
        return 0;
    }
    if (rc) {
        hipFree(cnt);
        return rc;
    }
    const uint64_t total = r[0];
    size_t freeb = 0, totalb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    const uint64_t ntiles = (s->C + kScanTile - 1) / kScanTile;
    const uint64_t need = total * 4 + (s->C + 1) * 8 + ntiles * 8 + (256ull << 20);
    if (need > freeb) {
        hipFree(cnt);
        return 0;
    }
    unsigned long long* tiles = dev_alloc<unsigned long long>(ntiles, tmp_bytes, rc);
    if (!tiles) { hipFree(cnt); return rc; }
    if (!(s->d_heads = dev_alloc<unsigned long long>(s->C + 1, s->device_bytes, rc))) {
        hipFree(cnt);
        hipFree(tiles);
        return rc;
    }
    hipLaunchKernelGGL(k_tile_sums, dim3(ntiles), dim3(kScanThreads), 0, s->stream, cnt, s->C, tiles);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, s->stream, tiles, static_cast<uint32_t>(ntiles),
                       s->d_heads, s->C);
    hipLaunchKernelGGL(k_tile_apply, dim3(ntiles), dim3(kScanThreads), 0, s->stream, cnt, s->C, tiles, s->d_heads);
    HIP_TRY(hipGetLastError());
    if (!(s->d_csr = dev_alloc<uint32_t>(total, s->device_bytes, rc))) {
        hipFree(cnt);
        hipFree(tiles);
        return rc;
    }
    rc = count_moves(s, 3, r, MoveArgs{nullptr, s->d_heads, s->d_csr});
    HIP_TRY(hipEventRecord(e1, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    hipFree(cnt);
    hipFree(tiles);
    if (rc) return rc;
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    s->csr_entries = total;
    s->csr_build_ms = ms;
    return 0;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* cr_last_error(void) { return g_err; }

int cr_solver_create(int device, cr_solver** out) {
    if (out == nullptr) return fail(CR_ERR_ARG, "null output pointer");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(CR_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count) return fail(CR_ERR_ARG, "device %d of %d", device, count);
    HIP_TRY(hipSetDevice(device));
    auto s = std::make_unique<cr_solver>();
    s->device = device;
    HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    *out = s.release();
    g_err[0] = '\0';
    return CR_OK;
}

void cr_solver_destroy(cr_solver* s) {
    if (!s) return;
    hipSetDevice(s->device);
    if (s->stream) hipStreamSynchronize(s->stream);
    s->release();
    if (s->h_ctl) hipHostFree(s->h_ctl);
    if (s->h_lists) hipHostFree(s->h_lists);
    for (hipEvent_t e : s->events) hipEventDestroy(e);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
}

int32_t cr_parse_matrix(const uint8_t* text, uint64_t len, uint8_t* adj, int32_t cap) {
    return cr::parse_matrix(text, len, adj, cap);
}

int cr_load_graph_file(cr_solver* s, const char* filename, int32_t k, int32_t p) {
    if (int rc = check_solver(s)) return rc;
    if (filename == nullptr) return fail(CR_ERR_ARG, "null filename");
    std::FILE* f = std::fopen(filename, "rb");
    if (!f) return fail(CR_ERR_EMPTY_GRAPH, "Graph is empty or failed to load: %s", filename);
    std::vector<uint8_t> text;
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.insert(text.end(), buf, buf + got);
    std::fclose(f);
    const int32_t n = cr::parse_matrix(text.data(), text.size(), nullptr, 0);
    if (n <= 0) return fail(CR_ERR_EMPTY_GRAPH, "Graph is empty or failed to load: %s", filename);
    if (n > CR_MAX_N) return fail(CR_ERR_GRAPH_LIMIT, "graph has %d vertices; the format limit is %d", n, CR_MAX_N);
    std::vector<uint8_t> adj(static_cast<size_t>(n) * n);
    cr::parse_matrix(text.data(), text.size(), adj.data(), n);
    s->path = filename;
    return do_load(s, n, adj.data(), k, p);
}

int32_t cr_parse_edge_list(const uint8_t* text, uint64_t len, uint8_t* adj, int32_t cap) {
    return cr::parse_edge_list(text, len, adj, cap);
}

int cr_load_edge_file(cr_solver* s, const char* filename, int32_t k, int32_t p) {
    if (int rc = check_solver(s)) return rc;
    if (filename == nullptr) return fail(CR_ERR_ARG, "null filename");
    std::FILE* f = std::fopen(filename, "rb");
    if (!f) return fail(CR_ERR_EMPTY_GRAPH, "Error: The file '%s' was not found.", filename);
    std::vector<uint8_t> text;
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.insert(text.end(), buf, buf + got);
    std::fclose(f);
    const int32_t n = cr::parse_edge_list(text.data(), text.size(), nullptr, 0);
    if (n <= 0) return fail(CR_ERR_EMPTY_GRAPH, "Error: No valid edges found in file %s.", filename);
    if (n > CR_MAX_N) return fail(CR_ERR_GRAPH_LIMIT, "graph has %d vertices; the format limit is %d", n, CR_MAX_N);
    std::vector<uint8_t> adj(static_cast<size_t>(n) * n);
    cr::parse_edge_list(text.data(), text.size(), adj.data(), n);
    s->path = filename;
    return do_load(s, n, adj.data(), k, p);
}

int cr_load_graph_matrix(cr_solver* s, int32_t n, const uint8_t* adj, int32_t k, int32_t p) {
    if (int rc = check_solver(s)) return rc;
    if (adj == nullptr && n > 0) return fail(CR_ERR_ARG, "null matrix");
    s->path = "matrix";
    return do_load(s, n, adj, k, p);
}

int cr_set_rules(cr_solver* s, int32_t cop_stay, int32_t robber_stay) {
    if (int rc = check_solver(s)) return rc;
    s->cop_stay = cop_stay ? 1 : 0;
    s->robber_stay = robber_stay ? 1 : 0;
    if (s->stage > kLoaded) s->stage = kLoaded;
    return 0;
}

int cr_set_variant(cr_solver* s, int32_t variant) {
    if (int rc = check_solver(s)) return rc;
    if (variant != CR_VARIANT_COLUMNS && variant != CR_VARIANT_MACRO) return fail(CR_ERR_ARG, "variant %d", variant);
    s->variant = variant;
    if (s->stage > kLoaded) s->stage = kLoaded;
    return 0;
}

int cr_build_aux_graph(cr_solver* s) {
    if (int rc = check_solver(s)) return rc;
    HIP_TRY(hipSetDevice(s->device));
    return do_build(s);
}

int cr_initialize_captures(cr_solver* s, uint64_t* base_captures) {
    if (int rc = check_solver(s)) return rc;
    HIP_TRY(hipSetDevice(s->device));
    return do_init(s, base_captures);
}

int cr_main_loop(cr_solver* s, uint64_t* pass_counts, int32_t max_passes, int32_t* n_passes) {
    if (int rc = check_solver(s)) return rc;
    HIP_TRY(hipSetDevice(s->device));
    if (int rc = do_loop(s)) return rc;
    if (pass_counts)
        for (int32_t i = 0; i < max_passes && i < static_cast<int32_t>(s->pass_counts.size()); ++i)
            pass_counts[i] = s->pass_counts[i];
    if (n_passes) *n_passes = static_cast<int32_t>(s->pass_counts.size());
    return 0;
}

int cr_find_final_result(cr_solver* s, cr_verdict* out) {
    if (int rc = check_solver(s)) return rc;
    if (out == nullptr) return fail(CR_ERR_ARG, "null verdict");
    if (s->stage < kSolved) return fail(CR_ERR_ORDER, "findFinalResult before mainLoop");
    HIP_TRY(hipSetDevice(s->device));
    const unsigned long long init = ~0ull;
    HIP_TRY(hipMemcpyAsync(&s->d_ctl->verdict, &init, sizeof init, hipMemcpyHostToDevice, s->stream));
    LevelArgs a = level_args(s);
    if (s->from_blob)
        a.won0 = s->d_marks;  // derived from the table read from the blob
    else if (s->colplan)
        a.won0 = crd::columns_won0(s->colplan);  // column 0 of the p >= 2 solver
    if (int rc = dispatch_w<LaunchVerdict>(s->W, s, a)) return rc;
    HIP_TRY(hipGetLastError());
    unsigned long long v = 0;
    HIP_TRY(hipMemcpyAsync(&v, &s->d_ctl->verdict, sizeof v, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::memset(out, 0, sizeof *out);
    const uint32_t worst = static_cast<uint32_t>(v >> 32);
    if (v != ~0ull && worst < kMaxRound) {
        out->win = 1;
        out->plys = static_cast<int32_t>(worst);
        out->best_cid = v & 0xffffffffull;
        cr::unrank_tuple(out->best_cid, s->g.n, s->k, out->positions);
    } else {
        out->win = 0;
        out->plys = -1;
    }
    return 0;
}

int cr_save_aux_graph(cr_solver* s, const char* path) {
    if (int rc = check_solver(s)) return rc;
    if (path == nullptr) return fail(CR_ERR_ARG, "null path");
    if (s->stage < kBuilt) return fail(CR_ERR_ORDER, "outputData before buildAuxGraph");
    HIP_TRY(hipSetDevice(s->device));
    auto fetch = [s](uint64_t off, uint64_t cnt, uint8_t* dst) -> int {
        if (hipMemcpyAsync(dst, s->d_table + off, cnt, hipMemcpyDeviceToHost, s->stream) != hipSuccess) return -1;
        return hipStreamSynchronize(s->stream) == hipSuccess ? 0 : -1;
    };
    const int rc = cr::write_aux_blob(path, s->g, s->k, s->columns, s->cop_stay != 0, s->robber_stay != 0, s->S, fetch);
    if (rc != 0) return fail(rc, "could not write %s", path);
    return 0;
}

int cr_load_aux_graph(cr_solver* s, const char* path) {
    if (int rc = check_solver(s)) return rc;
    if (path == nullptr) return fail(CR_ERR_ARG, "null path");
    if (s->stage < kLoaded) return fail(CR_ERR_ORDER, "loadAuxGraph before loadGraphFile");
    HIP_TRY(hipSetDevice(s->device));
    uint64_t header[12];
    std::string why;
    if (int rc = cr::read_aux_header(path, header, nullptr, why)) return fail(rc, "%s", why.c_str());
    if (header[6] != (s->cop_stay ? 1u : 0u) || header[7] != (s->robber_stay ? 1u : 0u))  // before any allocation
        return fail(CR_ERR_CACHE_RULES, "Cache rules mismatch: Cached SelfEdges [%llu%llu] vs Requested [%d%d]",
                    static_cast<unsigned long long>(header[6]), static_cast<unsigned long long>(header[7]),
                    s->cop_stay ? 1 : 0, s->robber_stay ? 1 : 0);
    if (int rc = do_build(s)) return rc;
    auto put = [s](uint64_t off, uint64_t cnt, const uint8_t* src) -> int {
        if (hipMemcpyAsync(s->d_table + off, src, cnt, hipMemcpyHostToDevice, s->stream) != hipSuccess) return CR_ERR_DEVICE;
        return hipStreamSynchronize(s->stream) == hipSuccess ? 0 : CR_ERR_DEVICE;
    };
    if (int rc = cr::read_aux_blob(path, s->g, s->k, s->columns, s->cop_stay != 0, s->robber_stay != 0, s->S, why, put))
        return fail(rc, "%s: %s", path, why.c_str());
    int rc = 0;
    if (!(s->d_marks = dev_alloc<uint64_t>(s->C * s->W, s->device_bytes, rc))) return rc;
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((s->C + kBlock - 1) / kBlock, 4096));
    hipLaunchKernelGGL(k_marks_from_table, dim3(std::max<uint32_t>(grid, 1)), dim3(kBlock), 0, s->stream, s->d_table,
                       static_cast<uint64_t>(s->C), s->g.n, s->columns, static_cast<int32_t>(s->W), s->d_marks);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->from_blob = true;
    s->stage = kSolved;
    return 0;
}

int cr_read_aux_header(const char* path, uint64_t* header12) {
    if (path == nullptr || header12 == nullptr) return fail(CR_ERR_ARG, "null argument");
    std::string why;
    if (int rc = cr::read_aux_header(path, header12, nullptr, why)) return fail(rc, "%s", why.c_str());
    return 0;
}

int cr_get_info(const cr_solver* s, cr_info* out) {
    if (int rc = check_solver(s)) return rc;
    if (out == nullptr) return fail(CR_ERR_ARG, "null info");
    std::memset(out, 0, sizeof *out);
    out->n = s->g.n;
    out->k = s->k;
    out->p = s->p;
    out->columns = s->variant == CR_VARIANT_MACRO ? 1 : 2 * s->p;
    out->variant = s->variant;
    out->max_degree = s->g.max_out;
    out->config_count = s->stage >= kLoaded ? cr::config_count(s->g.n, s->k) : 0;
    out->num_states = out->config_count * static_cast<uint64_t>(s->g.n) * out->columns;
    out->device_bytes = s->device_bytes;
    return 0;
}

int cr_get_stats(const cr_solver* s, cr_stats* out) {
    if (int rc = check_solver(s)) return rc;
    if (out == nullptr) return fail(CR_ERR_ARG, "null stats");
    *out = s->stats;
    return 0;
}

int cr_copy_states(cr_solver* s, uint8_t* dst, uint64_t offset, uint64_t count) {
    if (int rc = check_solver(s)) return rc;
    if (s->stage < kBuilt) return fail(CR_ERR_ORDER, "no state space");
    if (dst == nullptr || offset > s->S || count > s->S - offset) return fail(CR_ERR_ARG, "range outside the %llu states", static_cast<unsigned long long>(s->S));
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMemcpyAsync(dst, s->d_table + offset, count, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return 0;
}

int cr_copy_configs(cr_solver* s, uint8_t* dst, uint64_t first, uint64_t count) {
    if (int rc = check_solver(s)) return rc;
    if (s->stage < kLoaded) return fail(CR_ERR_ORDER, "no graph");
    const uint64_t C = cr::config_count(s->g.n, s->k);
    if (dst == nullptr || first > C || count > C - first) return fail(CR_ERR_ARG, "config range");
    for (uint64_t i = 0; i < count; ++i) cr::unrank_tuple(first + i, s->g.n, s->k, dst + i * s->k);
    return 0;
}

int cr_level_counts(cr_solver* s, uint64_t* out) {
    if (int rc = check_solver(s)) return rc;
    if (s->variant == CR_VARIANT_MACRO || s->p != 1)
        return fail(CR_ERR_UNSUPPORTED, "depth levels and retro-edges are kept for the p = 1 column solver");
    if (out == nullptr) return fail(CR_ERR_ARG, "null output");
    if (s->stage < kSolved) return fail(CR_ERR_ORDER, "not solved");
    if (s->from_blob) return fail(CR_ERR_ORDER, "a table read from a blob has no level data: solve it first");
    for (int d = 0; d < 127; ++d) {
        out[2 * d] = s->h_ctl->lvl[d][0];
        out[2 * d + 1] = s->h_ctl->lvl[d][1];
    }
    return 0;
}

int cr_count_retro_edges(cr_solver* s, uint64_t* e_cop, uint64_t* e_rob, uint64_t* won) {
    if (int rc = check_solver(s)) return rc;
    if (s->cop_stay != s->robber_stay)
        return fail(CR_ERR_UNSUPPORTED, "retro-edges are counted for equal cop and robber move rules");
    if (s->variant == CR_VARIANT_MACRO || s->p != 1)
        return fail(CR_ERR_UNSUPPORTED, "depth levels and retro-edges are kept for the p = 1 column solver");
    if (s->stage < kSolved) return fail(CR_ERR_ORDER, "not solved");
    if (s->from_blob) return fail(CR_ERR_ORDER, "a table read from a blob has no level data: solve it first");
    unsigned long long r[3];
    if (int rc = count_moves(s, 0, r)) return rc;
    if (e_cop) *e_cop = r[0];
    if (e_rob) *e_rob = r[1];
    if (won) *won = r[2];
    return 0;
}

int cr_work_profile(cr_solver* s, uint64_t* e_cop, uint64_t* e_rob, uint8_t* pulled) {
    if (int rc = check_solver(s)) return rc;
    if (s->cop_stay != s->robber_stay)
        return fail(CR_ERR_UNSUPPORTED, "retro-edges are counted for equal cop and robber move rules");
    if (s->variant == CR_VARIANT_MACRO || s->p != 1)
        return fail(CR_ERR_UNSUPPORTED, "depth levels and retro-edges are kept for the p = 1 column solver");
    if (s->stage < kSolved) return fail(CR_ERR_ORDER, "not solved");
    if (s->from_blob) return fail(CR_ERR_ORDER, "a table read from a blob has no level data: solve it first");
    unsigned long long r[3];
    if (int rc = count_moves(s, 0, r)) return rc;
    Ctl* h = s->h_ctl;
    HIP_TRY(hipMemcpyAsync(h, s->d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int d = 0; d < 127; ++d) {
        if (e_cop) e_cop[d] = h->hist[0][d];
        if (e_rob) e_rob[d] = h->hist[1][d];
        if (pulled) pulled[d] = static_cast<uint8_t>(h->dir[d] != 0);
    }
    return 0;
}

int cr_level_rows(cr_solver* s, uint64_t* rows) {
    if (int rc = check_solver(s)) return rc;
    if (rows == nullptr) return fail(CR_ERR_ARG, "null output");
    if (s->variant == CR_VARIANT_MACRO || s->p != 1)
        return fail(CR_ERR_UNSUPPORTED, "depth levels are kept for the p = 1 column solver");
    if (s->stage < kSolved) return fail(CR_ERR_ORDER, "not solved");
    if (s->from_blob) return fail(CR_ERR_ORDER, "a table read from a blob has no level data: solve it first");
    for (int d = 0; d < 127; ++d) {
        rows[2 * d] = s->h_ctl->rows[d][0];
        rows[2 * d + 1] = s->h_ctl->rows[d][1];
    }
    return 0;
}

int cr_get_macro_stats(cr_solver* s, cr_macro_stats* out, uint64_t* nodes_per_step, uint64_t* edges_per_step,
                       int32_t cap) {
    if (int rc = check_solver(s)) return rc;
    if (out == nullptr) return fail(CR_ERR_ARG, "null output");
    if (s->variant != CR_VARIANT_MACRO) return fail(CR_ERR_UNSUPPORTED, "no trajectory DAG in the column solver");
    if (s->stage < kSolved) return fail(CR_ERR_ORDER, "not solved");
    if (s->from_blob) return fail(CR_ERR_ORDER, "a table read from a blob has no level data: solve it first");
    const crd::MacroProfile& m = s->mprof;
    std::memset(out, 0, sizeof *out);
    out->steps = static_cast<int32_t>(m.nodes.size());
    for (uint64_t x : m.nodes) out->dag_nodes += x;
    for (uint64_t x : m.edges) out->dag_edges += x;
    out->build_moves = m.build_moves;
    out->leaf_moves = m.leaf_moves;
    out->up_edges = m.up_edges;
    out->build_ms = m.build_ms;
    out->passes_ms = m.passes_ms;
    for (int32_t d = 0; d < cap && d < out->steps; ++d) {
        if (nodes_per_step) nodes_per_step[d] = m.nodes[d];
        if (edges_per_step) edges_per_step[d] = d < static_cast<int32_t>(m.edges.size()) ? m.edges[d] : 0;
    }
    return 0;
}

namespace {
int check_round_table(cr_solver* s) {
    if (int rc = check_solver(s)) return rc;
    if (s->stage < kSolved) return fail(CR_ERR_ORDER, "not solved");
    if (s->variant != CR_VARIANT_COLUMNS || s->p != 1)
        return fail(CR_ERR_UNSUPPORTED, "rounds are defined on the two-column table (column variant, p = 1)");
    return 0;
}
}  // namespace

int cr_write_dp_table(cr_solver* s, const char* path) {
    if (int rc = check_round_table(s)) return rc;
    if (path == nullptr) return fail(CR_ERR_ARG, "null path");
    HIP_TRY(hipSetDevice(s->device));
    auto fetch = [s](uint64_t off, uint64_t cnt, uint8_t* dst) -> int {
        if (hipMemcpyAsync(dst, s->d_table + off, cnt, hipMemcpyDeviceToHost, s->stream) != hipSuccess) return -1;
        return hipStreamSynchronize(s->stream) == hipSuccess ? 0 : -1;
    };
    const int rc = cr::write_dp_table(path, s->g.n, s->k, s->columns, s->C, fetch);
    if (rc != 0) return fail(rc, "could not write %s", path);
    return 0;
}

int cr_extract_game(cr_solver* s, int32_t order, cr_game_step* out, int32_t cap, int32_t* length) {
    if (int rc = check_round_table(s)) return rc;
    if (cap > 0 && out == nullptr) return fail(CR_ERR_ARG, "null output");
    if (order != CR_GAME_ORDER_PY && order != CR_GAME_ORDER_CPP) return fail(CR_ERR_ARG, "game order %d", order);
    HIP_TRY(hipSetDevice(s->device));
    std::vector<cr::GameStep> steps;
    if (int rc = crd::game_extract(s->gv, s->k, s->columns, s->d_table, s->g.matrix, s->cop_stay != 0,
                                   s->robber_stay != 0, order == CR_GAME_ORDER_CPP, s->stream, steps))
        return rc;
    for (int32_t i = 0; i < cap && i < static_cast<int32_t>(steps.size()); ++i) {
        std::memset(&out[i], 0, sizeof out[i]);
        std::memcpy(out[i].cops, steps[i].cops, CR_MAX_K);
        out[i].robber = steps[i].robber;
        out[i].turn = steps[i].turn;
    }
    if (length) *length = static_cast<int32_t>(steps.size());
    return 0;
}

int cr_write_perfect_game(cr_solver* s, int32_t order, const char* path) {
    if (int rc = check_round_table(s)) return rc;
    if (path == nullptr) return fail(CR_ERR_ARG, "null path");
    if (order != CR_GAME_ORDER_PY && order != CR_GAME_ORDER_CPP) return fail(CR_ERR_ARG, "game order %d", order);
    HIP_TRY(hipSetDevice(s->device));
    std::vector<cr::GameStep> steps;
    if (int rc = crd::game_extract(s->gv, s->k, s->columns, s->d_table, s->g.matrix, s->cop_stay != 0,
                                   s->robber_stay != 0, order == CR_GAME_ORDER_CPP, s->stream, steps))
        return rc;
    if (int rc = cr::write_game_json(path, s->k, steps)) return fail(rc, "could not write %s", path);
    return 0;
}

int cr_cop_number(cr_solver* s, int32_t k_max, int32_t* cop_number, cr_verdict* verdict) {
    if (int rc = check_solver(s)) return rc;
    if (s->stage < kLoaded) return fail(CR_ERR_ORDER, "cop number before loadGraphFile");
    if (k_max < 1 || k_max > CR_MAX_K) return fail(CR_ERR_K_RANGE, "k_max=%d outside 1..%d", k_max, CR_MAX_K);
    HIP_TRY(hipSetDevice(s->device));
    const std::vector<uint8_t> matrix = s->g.matrix;
    const int32_t n = s->g.n, p = s->p;
    if (cop_number) *cop_number = 0;
    for (int32_t k = 1; k <= k_max; ++k) {
        int rc = do_load(s, n, matrix.data(), k, p);
        if (!rc) rc = do_build(s);
        if (!rc) rc = do_init(s, nullptr);
        if (!rc) rc = do_loop(s);
        cr_verdict v{};
        if (!rc) rc = cr_find_final_result(s, &v);
        if (rc) return rc;
        if (verdict) *verdict = v;
        if (v.win) {
            if (cop_number) *cop_number = k;
            return 0;
        }
    }
    return 0;
}

int cr_count_transitions(cr_solver* s, uint64_t* total) {
    if (int rc = check_solver(s)) return rc;
    if (total == nullptr) return fail(CR_ERR_ARG, "null output");
    if (s->stage < kBuilt) return fail(CR_ERR_ORDER, "count before buildAuxGraph");
    unsigned long long r[3];
    if (int rc = count_moves(s, 1, r)) return rc;
    *total = r[0];
    return 0;
}

int cr_set_tuning(cr_solver* s, int32_t key, int64_t value) {
    if (int rc = check_solver(s)) return rc;
    switch (key) {
        case CR_TUNE_PUSH_FILTER: s->filter = value != 0; return 0;
        case CR_TUNE_PUSH_BLOCKS:
            if (value < 1 || value > (1 << 20)) return fail(CR_ERR_ARG, "push blocks %lld", static_cast<long long>(value));
            s->push_blocks = static_cast<uint32_t>((value + kXcds - 1) / kXcds * kXcds);  // whole XCD rounds
            return 0;
        case CR_TUNE_ROW_BLOCKS:
            if (value < 1 || value > (1 << 20)) return fail(CR_ERR_ARG, "row blocks %lld", static_cast<long long>(value));
            s->row_blocks = static_cast<uint32_t>(value);
            return 0;
        case CR_TUNE_LOG_BLOCKS:
            if (value < 1 || value > (1 << 20)) return fail(CR_ERR_ARG, "log blocks %lld", static_cast<long long>(value));
            s->list_blocks = static_cast<uint32_t>(value);
            return 0;
        case CR_TUNE_PULL_SCALE:
            if (value < 0) value = -1;  // the size-dependent default
            s->pull_scale = value;
            set_pull_threshold(s);
            return 0;
        case CR_TUNE_PULL_TILES: s->pull_tiles = value != 0; return 0;
        case CR_TUNE_DEPTH_LOG:
            if (value < 0 || value > 1024) return fail(CR_ERR_ARG, "depth log %lld", static_cast<long long>(value));
            s->log_factor = static_cast<uint32_t>(value);
            if (s->stage > kLoaded) s->stage = kLoaded;
            return 0;
        case CR_TUNE_DEPTH_LOG_MIN:
            if (value < 0) return fail(CR_ERR_ARG, "depth log minimum %lld", static_cast<long long>(value));
            s->log_min_configs = static_cast<uint64_t>(value);
            if (s->stage > kLoaded) s->stage = kLoaded;
            return 0;
        case CR_TUNE_PULL_TILE_BLOCKS:
            if (value < 1 || value > (1 << 20)) return fail(CR_ERR_ARG, "tile blocks %lld", static_cast<long long>(value));
            s->tile_blocks = static_cast<uint32_t>(value);
            return 0;
        case CR_TUNE_FUSE_FINALIZE: s->fuse_finalize = value != 0; return 0;
        case CR_TUNE_PUSH_DIRTY: s->push_dirty = value != 0; return 0;
        case CR_TUNE_END_PASS: s->end_pass = value != 0; return 0;
        case CR_TUNE_PULL_PAIR:
            if (value < 0 || value > 2) return fail(CR_ERR_ARG, "pull pair %lld", static_cast<long long>(value));
            s->pull_pair = static_cast<int32_t>(value);
            return 0;
        case CR_TUNE_PAIR_ROWS:
            if (value < 1 || value > kPairMaxRows) return fail(CR_ERR_ARG, "pair tile rows %lld", static_cast<long long>(value));
            s->pair_rows = static_cast<uint32_t>(value);
            return 0;
        case CR_TUNE_PAIR_BLOCKS:
            if (value < 1 || value > (1 << 20)) return fail(CR_ERR_ARG, "pair blocks %lld", static_cast<long long>(value));
            s->pair_blocks = static_cast<uint32_t>(value);
            return 0;
        case CR_TUNE_SCAN_TILE:
            if (value < kScanTileMin || value > kScanTileMax) return fail(CR_ERR_ARG, "scan tile %lld", static_cast<long long>(value));
            s->scan_tile = static_cast<uint32_t>(value);
            return 0;
        case CR_TUNE_PULL_TILE1:
            if (value < 0 || value > kTileMaxRows) return fail(CR_ERR_ARG, "stage-1 tile rows %lld", static_cast<long long>(value));
            s->tile1_rows = static_cast<uint32_t>(value);
            return 0;
        case CR_TUNE_ROBBER_OPEN:
            if (value < 0 || value > 1000) return fail(CR_ERR_ARG, "robber open factor %lld", static_cast<long long>(value));
            s->open_factor = static_cast<uint32_t>(value);
            return 0;
        case CR_TUNE_PULL_GATE:
            if (value < 0 || value > (1 << 20)) return fail(CR_ERR_ARG, "pull gate %lld", static_cast<long long>(value));
            s->pull_gate = static_cast<uint32_t>(value);
            return 0;
        case CR_TUNE_PULL_TILE_CHUNK:
            if (value < 1 || value > (1 << 20)) return fail(CR_ERR_ARG, "tile chunk %lld", static_cast<long long>(value));
            s->tile_chunk = static_cast<uint32_t>(value);
            return 0;
        case CR_TUNE_PULL_TILE_ROWS:
            if (value < 1 || value > kTileMaxRows) return fail(CR_ERR_ARG, "tile rows %lld", static_cast<long long>(value));
            s->tile_rows = static_cast<uint32_t>(value);
            return 0;
        case CR_TUNE_PRED_CSR: s->use_csr = value != 0; if (s->stage > kLoaded) s->stage = kLoaded; return 0;
        case CR_TUNE_LEVELS_PER_SYNC:
            if (value < 1 || value > 128) return fail(CR_ERR_ARG, "levels per sync %lld", static_cast<long long>(value));
            s->levels_per_sync = static_cast<int32_t>(value);
            return 0;
        default: return fail(CR_ERR_ARG, "unknown tuning key %d", key);
    }
}

int cr_set_kernel_timing_filter(cr_solver* s, const char* name) {
    if (int rc = check_solver(s)) return rc;
    s->time_only = name ? name : "";
    return 0;
}

int cr_set_kernel_timing(cr_solver* s, int32_t enabled) {
    if (int rc = check_solver(s)) return rc;
    s->timing = enabled != 0;
    return 0;
}

int cr_kernel_times(cr_solver* s, char* names_out, int32_t names_cap, double* ms, int64_t* launches,
                    int32_t cap, int32_t* count) {
    if (int rc = check_solver(s)) return rc;
    int32_t i = 0;
    std::string names;
    for (const auto& nm : s->korder) {
        if (i >= cap) break;
        const KTime& t = s->ktimes[nm];
        if (ms) ms[i] = t.ms;
        if (launches) launches[i] = t.launches;
        names += nm;
        names.push_back('\0');
        ++i;
    }
    if (names_out && names_cap > 0) {
        const size_t m = std::min<size_t>(names.size(), static_cast<size_t>(names_cap));
        std::memcpy(names_out, names.data(), m);
    }
    if (count) *count = i;
    return 0;
}

int cr_cache_key(const char* graph_path, int32_t k, int32_t p, int32_t variant, char* out, int32_t cap) {
    if (graph_path == nullptr || out == nullptr || cap <= 0) return fail(CR_ERR_ARG, "bad arguments");
    const std::string algo = variant == CR_VARIANT_MACRO ? "k_cops_v2_" + std::to_string(p) + "vis"
                                                         : "k_cops_" + std::to_string(p) + "vis";
    const std::string key = cr::cache_key(algo, graph_path, k, p, false, false);
    if (static_cast<int32_t>(key.size()) >= cap) return fail(CR_ERR_ARG, "buffer too small");
    std::memcpy(out, key.c_str(), key.size() + 1);
    return static_cast<int>(key.size());
}

}  // extern "C"
