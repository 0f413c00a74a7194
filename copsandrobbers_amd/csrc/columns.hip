// columns.hip -- the column solver (solvers/k_cops_visibility.cpp) for robber
// visibility p >= 2 on the device.
//
// The reference keeps 2p columns per (C, r) (AuxGraph::columns,
// k_cops_visibility.cpp:59-83) and sweeps them until nothing changes
// (:129-299): column c reads column c+1, the last column reads column 0.
//   even c       cop turn: won if some team move C -> C' reaches a won state
//                (C', r, c+1); depth = min + 1 (:160-197);
//   odd c < 2p-1 robber turn out of sight: won if (C, r, c+1) is; depth + 1
//                (:209-228);
//   c = 2p-1     the robber is seen again: won if (C, v, 0) is won for every
//                v reachable from r in exactly p robber moves; depth = max + 1
//                (:235-288).
// Captures are depth 0 in columns 0 and 1 only (:85-127); marks stop below
// the 127 cap (:189, :220, :280).
//
// Device form: level-synchronous retrograde over that column graph. Level d
// turns the states of depth d of every column into those of depth d+1:
//   k_cl_stage     for every even column c, the factorised pull of the
//                  column-(c+1) frontier (H_i[A][B] = OR_{v in N(max A)}
//                  H_{i-1}[A - max A][B + v], H_0 = frontier, H_K = candidates);
//   k_cl_finalize  per configuration row: the last column's AND over the
//                  p-step reach masks against the column-0 plane (before this
//                  level touches it), the interior copies, the pulled
//                  candidates; table bytes, won planes, frontiers, counts.
// The reference's pass in which a state of column c and depth t is marked is
// t - floor((c + t) / 2p) + [(c + t) mod 2p == 0]: its chain of depth-1
// successors visits columns c, c+1, ... and ends on a capture in column 0 or
// 1; each step costs one pass except a wrap from the last column to column 0,
// which the same sweep sees, unless it is the final step onto a capture,
// which no pass before the first can use.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "cr_columns.h"
#include "cr_host.h"

namespace {

constexpr int kMaxCols = 126;  // p <= 63

struct ClArgs {
    uint64_t* won;       // [cols][C][W]
    uint64_t* fr;        // [cols][C][W]: states of the current depth
    uint64_t* nx;        // [p][C][W]: pulled candidates of the even columns
    const uint64_t* reach;  // [n][4]
    uint8_t* table;
    unsigned long long* hist;  // [cols][128]
    uint32_t C;
    int32_t n, cols;
};

template <int K>
__global__ __launch_bounds__(kBlock) void k_cl_stage(GraphView gv, const uint64_t* in, uint64_t* out,
                                                     const uint64_t* won, uint64_t rows, uint32_t cnt_b,
                                                     uint32_t cnt_b_in, int32_t stage) {
    extern __shared__ __align__(16) uint8_t lds[];
    load_image(lds, gv.image, gv.image_bytes);
    const uint32_t* binom = reinterpret_cast<const uint32_t*>(lds + gv.off_binom);
    const uint16_t* out_off = reinterpret_cast<const uint16_t*>(lds + gv.off_out_off);
    const uint8_t* out_nbr = lds + gv.off_out_nbr;
    const int W = gv.W, n = gv.n, i = stage, jb = K - i;
    uint64_t full[4];
    for (int q = 0; q < 4; ++q) {
        const int bits = min(64, max(0, n - 64 * q));
        full[q] = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    }
    for (uint64_t o = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; o < rows;
         o += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint32_t o32 = static_cast<uint32_t>(o);  // rows < 2^32 (checked on the host)
        const uint32_t ra = o32 / cnt_b, rb = o32 - ra * cnt_b;
        uint32_t A[K], B[K + 1];
        unrank_j(ra, i, n, binom, gv.M, A);
        unrank_j(rb, jb, n, binom, gv.M, B);
        const uint32_t ra_in = rank_j(A, i - 1, n, binom, gv.M);
        const uint32_t c = A[i - 1];
        uint64_t acc[4] = {0, 0, 0, 0};
        for (int e = out_off[c]; e < out_off[c + 1]; ++e) {
            const uint32_t v = out_nbr[e];
            uint32_t M2[K + 1];
            int pos = 0;
            bool placed = false;
            for (int t = 0; t < jb; ++t) {
                if (!placed && v <= B[t]) {
                    M2[pos++] = v;
                    placed = true;
                }
                M2[pos++] = B[t];
            }
            if (!placed) M2[pos++] = v;
            const uint64_t r_in = static_cast<uint64_t>(ra_in) * cnt_b_in + rank_j(M2, jb + 1, n, binom, gv.M);
            for (int q = 0; q < W; ++q) acc[q] |= in[r_in * W + q];
        }
        if (i < K) {
            for (int q = 0; q < W; ++q) out[o * W + q] = acc[q];
        } else {
            for (int q = 0; q < W; ++q) out[o * W + q] = acc[q] & ~won[o * W + q] & full[q];
        }
    }
}

// Seeds the captures: columns 0 and 1 at depth 0, every other state unmarked.
template <int K>
__global__ __launch_bounds__(kBlock) void k_cl_init(GraphView gv, ClArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    load_image(lds, gv.image, gv.image_bytes);
    const uint32_t* binom = reinterpret_cast<const uint32_t*>(lds + gv.off_binom);
    const int W = gv.W, n = gv.n, cols = a.cols;
    const uint64_t plane = static_cast<uint64_t>(a.C) * W;
    unsigned long long caps = 0, unused = 0;
    for (uint64_t c = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; c < a.C;
         c += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        uint32_t cops[K];
        unrank<K>(static_cast<uint32_t>(c), gv, binom, cops);
        uint64_t occ[4] = {0, 0, 0, 0};
        for (int j = 0; j < K; ++j) occ[cops[j] >> 6] |= 1ull << (cops[j] & 63);
        for (int col = 0; col < cols; ++col)
            for (int q = 0; q < W; ++q) {
                const uint64_t m = col < 2 ? occ[q] : 0ull;
                a.won[col * plane + c * W + q] = m;
                a.fr[col * plane + c * W + q] = m;
            }
        uint8_t* row = a.table + c * static_cast<uint64_t>(n) * cols;
        for (int r = 0; r < n; ++r) {
            const bool cap = (occ[r >> 6] >> (r & 63)) & 1;
            uint8_t* s = row + static_cast<uint64_t>(r) * cols;
            for (int col = 0; col < cols; ++col) s[col] = (cap && col < 2) ? won_byte(0) : kUnmarked;
            caps += cap;
        }
    }
    caps = wave_sum(caps);
    block_sum2(caps, unused);
    if (threadIdx.x == 0 && caps) {
        atomicAdd(&a.hist[0], caps);
        atomicAdd(&a.hist[128], caps);
    }
}

template <int W>
__device__ __forceinline__ void apply_new(const ClArgs& a, uint64_t c, int col, const uint64_t* fresh, uint8_t byte,
                                          unsigned long long* cnt) {
    const uint64_t plane = static_cast<uint64_t>(a.C) * W;
    uint8_t* row = a.table + c * static_cast<uint64_t>(a.n) * a.cols;
    unsigned long long k = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
        a.won[col * plane + c * W + q] |= fresh[q];
        a.fr[col * plane + c * W + q] = fresh[q];
        uint64_t m = fresh[q];
        k += __builtin_popcountll(m);
        while (m) {
            const int r = q * 64 + __builtin_ctzll(m);
            m &= m - 1;
            row[static_cast<uint64_t>(r) * a.cols + col] = byte;
        }
    }
    if (k) atomicAdd(&cnt[col], k);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_cl_finalize(ClArgs a, uint32_t d, uint64_t pulled_mask_lo,
                                                        uint64_t pulled_mask_hi) {
    __shared__ unsigned long long cnt[kMaxCols];
    for (int i = threadIdx.x; i < a.cols; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    const uint64_t plane = static_cast<uint64_t>(a.C) * W;
    const int last = a.cols - 1, n = a.n;
    uint64_t full[W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const int bits = min(64, max(0, n - 64 * q));
        full[q] = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    }
    const uint8_t byte = won_byte(d + 1);
    for (uint64_t c = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; c < a.C;
         c += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        // last column first: it reads column 0 as it was before this level
        uint64_t w0[W], lastnew[W];
#pragma unroll
        for (int q = 0; q < W; ++q) {
            w0[q] = a.won[c * W + q];
            lastnew[q] = 0;
        }
#pragma unroll
        for (int q = 0; q < W; ++q) {
            uint64_t cand = full[q] & ~a.won[last * plane + c * W + q];
            while (cand) {
                const int r = q * 64 + __builtin_ctzll(cand);
                cand &= cand - 1;
                bool ok = true;
#pragma unroll
                for (int w = 0; w < W; ++w) ok = ok && (a.reach[r * 4 + w] & ~w0[w]) == 0;
                if (ok) lastnew[q] |= 1ull << (r & 63);
            }
        }
        for (int col = 0; col < last; ++col) {
            uint64_t fresh[W];
            if ((col & 1) == 0) {
                const int h = col >> 1;
                const bool pulled = ((h < 64 ? pulled_mask_lo >> h : pulled_mask_hi >> (h - 64)) & 1) != 0;
#pragma unroll
                for (int q = 0; q < W; ++q) fresh[q] = pulled ? a.nx[h * plane + c * W + q] : 0ull;
            } else {  // interior robber column: the next column's frontier (not yet overwritten)
#pragma unroll
                for (int q = 0; q < W; ++q)
                    fresh[q] = a.fr[(col + 1) * plane + c * W + q] & ~a.won[col * plane + c * W + q];
            }
            apply_new<W>(a, c, col, fresh, byte, cnt);
        }
        apply_new<W>(a, c, last, lastnew, byte, cnt);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.cols; i += blockDim.x)
        if (cnt[i]) atomicAdd(&a.hist[i * 128 + d + 1], cnt[i]);
}

template <typename F>
int with_k(int k, F&& f) {
    switch (k) {
        case 1: return f(std::integral_constant<int, 1>{});
        case 2: return f(std::integral_constant<int, 2>{});
        case 3: return f(std::integral_constant<int, 3>{});
        case 4: return f(std::integral_constant<int, 4>{});
        case 5: return f(std::integral_constant<int, 5>{});
        case 6: return f(std::integral_constant<int, 6>{});
        case 7: return f(std::integral_constant<int, 7>{});
        case 8: return f(std::integral_constant<int, 8>{});
        default: return fail(CR_ERR_K_RANGE, "k=%d outside 1..%d", k, CR_MAX_K);
    }
}

template <typename F>
int with_w(int w, F&& f) {
    switch (w) {
        case 1: return f(std::integral_constant<int, 1>{});
        case 2: return f(std::integral_constant<int, 2>{});
        case 3: return f(std::integral_constant<int, 3>{});
        case 4: return f(std::integral_constant<int, 4>{});
        default: return fail(CR_ERR_ARG, "row width %d", w);
    }
}

uint32_t grid_threads(uint64_t threads, uint64_t cap) {
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((threads + kBlock - 1) / kBlock, cap)));
}

}  // namespace

namespace crd {

struct ColumnsPlan {
    GraphView gv{};
    hipStream_t st = nullptr;
    int k = 0, p = 0, cols = 0;
    uint64_t C = 0;
    uint64_t* won = nullptr;
    uint64_t* fr = nullptr;
    uint64_t* nx = nullptr;
    uint64_t* h[2] = {nullptr, nullptr};
    uint64_t* reach = nullptr;
    unsigned long long* hist = nullptr;  // [cols][128]
    uint8_t* table = nullptr;
    std::vector<uint64_t> ms;             // C(n+j-1, j), j = 0..k
    std::vector<unsigned long long> h_hist;
    uint64_t bytes = 0;

    ClArgs args() const {
        return ClArgs{won, fr, nx, reach, table, hist, static_cast<uint32_t>(C), gv.n, cols};
    }
};

uint64_t* columns_won0(ColumnsPlan* plan) { return plan ? plan->won : nullptr; }

void columns_free(ColumnsPlan* plan, uint64_t& device_bytes) {
    if (!plan) return;
    for (void* ptr : {static_cast<void*>(plan->won), static_cast<void*>(plan->fr), static_cast<void*>(plan->nx),
                      static_cast<void*>(plan->h[0]), static_cast<void*>(plan->h[1]), static_cast<void*>(plan->reach),
                      static_cast<void*>(plan->hist)})
        if (ptr) hipFree(ptr);
    device_bytes -= plan->bytes;
    delete plan;
}

int columns_init(ColumnsPlan** planp, const GraphView& gv, const std::vector<uint64_t>& reach, int k, int p,
                 hipStream_t st, uint8_t* table, uint64_t& device_bytes, uint64_t* base) {
    if (2 * p > kMaxCols) return fail(CR_ERR_UNSUPPORTED, "p=%d: the device column solver keeps at most %d columns", p, kMaxCols);
    ColumnsPlan* pl = *planp;
    if (pl == nullptr || pl->k != k || pl->p != p || pl->C != gv.C || pl->gv.n != gv.n) {
        if (pl) columns_free(pl, device_bytes);
        *planp = nullptr;
        pl = new ColumnsPlan();
        pl->k = k;
        pl->p = p;
        pl->cols = 2 * p;
        pl->C = gv.C;
        const int n = gv.n, W = gv.W;
        for (int j = 0; j <= k; ++j) pl->ms.push_back(cr::binom(n + j - 1, j));
        uint64_t hmax = 1;
        for (int i = 1; i < k; ++i) {
            const unsigned __int128 rows = static_cast<unsigned __int128>(pl->ms[i]) * pl->ms[k - i];
            if (rows >= (static_cast<unsigned __int128>(1) << 32)) {
                delete pl;
                return fail(CR_ERR_TOO_LARGE, "pull intermediates exceed 32-bit rows");
            }
            hmax = std::max<uint64_t>(hmax, static_cast<uint64_t>(rows));
        }
        const uint64_t plane = pl->C * W * 8;
        const uint64_t need = plane * (2ull * pl->cols + p) + 2 * hmax * W * 8 + static_cast<uint64_t>(n) * 32 +
                              static_cast<uint64_t>(pl->cols) * 128 * 8;
        size_t freeb = 0, totalb = 0;
        if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) {
            delete pl;
            return fail(CR_ERR_DEVICE, "hipMemGetInfo failed");
        }
        if (need + (256ull << 20) > freeb) {
            delete pl;
            return fail(CR_ERR_TOO_LARGE, "column planes need %llu bytes; %zu free", static_cast<unsigned long long>(need), freeb);
        }
        auto alloc = [&](void** ptr, uint64_t b) -> int {
            if (hipMalloc(ptr, std::max<uint64_t>(b, 8)) != hipSuccess) return fail(CR_ERR_DEVICE, "hipMalloc(%llu)", static_cast<unsigned long long>(b));
            pl->bytes += std::max<uint64_t>(b, 8);
            return 0;
        };
        int rc = 0;
        if ((rc = alloc(reinterpret_cast<void**>(&pl->won), plane * pl->cols)) ||
            (rc = alloc(reinterpret_cast<void**>(&pl->fr), plane * pl->cols)) ||
            (rc = alloc(reinterpret_cast<void**>(&pl->nx), plane * p)) ||
            (rc = alloc(reinterpret_cast<void**>(&pl->h[0]), hmax * W * 8)) ||
            (rc = alloc(reinterpret_cast<void**>(&pl->h[1]), hmax * W * 8)) ||
            (rc = alloc(reinterpret_cast<void**>(&pl->reach), static_cast<uint64_t>(n) * 32)) ||
            (rc = alloc(reinterpret_cast<void**>(&pl->hist), static_cast<uint64_t>(pl->cols) * 128 * 8))) {
            device_bytes += pl->bytes;
            columns_free(pl, device_bytes);
            return rc;
        }
        device_bytes += pl->bytes;
        *planp = pl;
    }
    pl->gv = gv;
    pl->st = st;
    pl->table = table;
    HIP_TRY(hipMemcpyAsync(pl->reach, reach.data(), static_cast<uint64_t>(gv.n) * 32, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(pl->hist, 0, static_cast<uint64_t>(pl->cols) * 128 * 8, st));
    const ClArgs a = pl->args();
    const int rc = with_k(k, [&](auto kc) -> int {
        constexpr int K = decltype(kc)::value;
        if (int r = raise_lds_limit(reinterpret_cast<const void*>(k_cl_init<K>), gv.image_bytes)) return r;
        hipLaunchKernelGGL(k_cl_init<K>, dim3(grid_threads(pl->C, 4096)), dim3(kBlock), gv.image_bytes, st, gv, a);
        return 0;
    });
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    unsigned long long b = 0;
    HIP_TRY(hipMemcpyAsync(&b, pl->hist, sizeof b, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (base) *base = b;
    return 0;
}

int columns_solve(ColumnsPlan* pl, TimeSink sink, std::vector<uint64_t>& passes, uint64_t* won_states,
                  int32_t* levels) {
    if (!pl) return fail(CR_ERR_ORDER, "mainLoop before initializeCaptures");
    const GraphView& gv = pl->gv;
    const int k = pl->k, cols = pl->cols, W = gv.W;
    const uint64_t plane_words = pl->C * W;
    const ClArgs a = pl->args();
    pl->h_hist.assign(static_cast<size_t>(cols) * 128, 0);
    // counts of the depth-d states per column (the frontier sizes)
    std::vector<unsigned long long> cur(cols, 0);
    HIP_TRY(hipMemcpyAsync(pl->h_hist.data(), pl->hist, pl->h_hist.size() * 8, hipMemcpyDeviceToHost, pl->st));
    HIP_TRY(hipStreamSynchronize(pl->st));
    for (int c = 0; c < cols; ++c) cur[c] = pl->h_hist[c * 128 + 0];
    hipEvent_t ev[2];
    HIP_TRY(hipEventCreate(&ev[0]));
    HIP_TRY(hipEventCreate(&ev[1]));
    int32_t lv = 0;
    for (uint32_t d = 0; d + 1 < CR_MAX_ROUND; ++d) {
        uint64_t pulled_lo = 0, pulled_hi = 0;
        if (sink.add) HIP_TRY(hipEventRecord(ev[0], pl->st));
        for (int c = 0; c + 1 < cols; c += 2) {
            if (cur[c + 1] == 0) continue;  // no frontier in column c+1: nothing to pull
            const int h = c >> 1;
            if (h < 64) pulled_lo |= 1ull << h;
            else pulled_hi |= 1ull << (h - 64);
            const int rc = with_k(k, [&](auto kc) -> int {
                constexpr int K = decltype(kc)::value;
                if (int r = raise_lds_limit(reinterpret_cast<const void*>(k_cl_stage<K>), gv.image_bytes)) return r;
                const uint64_t* in = pl->fr + (c + 1) * plane_words;
                for (int i = 1; i <= K; ++i) {
                    uint64_t* out = i == K ? pl->nx + h * plane_words : pl->h[(i - 1) & 1];
                    const uint64_t rows = pl->ms[i] * pl->ms[K - i];
                    hipLaunchKernelGGL(k_cl_stage<K>, dim3(grid_threads(rows, 16384)), dim3(kBlock), gv.image_bytes,
                                       pl->st, gv, in, out, static_cast<const uint64_t*>(pl->won + c * plane_words),
                                       rows, static_cast<uint32_t>(pl->ms[K - i]),
                                       static_cast<uint32_t>(pl->ms[K - i + 1]), i);
                    in = out;
                }
                return 0;
            });
            if (rc) return rc;
        }
        if (sink.add) HIP_TRY(hipEventRecord(ev[1], pl->st));
        const int rc = with_w(W, [&](auto wc) -> int {
            constexpr int WW = decltype(wc)::value;
            hipLaunchKernelGGL(k_cl_finalize<WW>, dim3(grid_threads(pl->C, 4096)), dim3(kBlock), 0, pl->st, a, d,
                               pulled_lo, pulled_hi);
            return 0;
        });
        if (rc) return rc;
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(pl->h_hist.data(), pl->hist, pl->h_hist.size() * 8, hipMemcpyDeviceToHost, pl->st));
        HIP_TRY(hipStreamSynchronize(pl->st));
        if (sink.add) {
            float ms = 0;
            hipEventElapsedTime(&ms, ev[0], ev[1]);
            sink.add(sink.ctx, "cl_pull", ms);
        }
        unsigned long long total = 0;
        for (int c = 0; c < cols; ++c) {
            cur[c] = pl->h_hist[c * 128 + d + 1];
            total += cur[c];
        }
        lv = static_cast<int32_t>(d + 1);
        if (total == 0) break;
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    // the reference's passes from the per-column depth histogram
    std::vector<uint64_t> per_pass;
    uint64_t won = 0;
    for (int c = 0; c < cols; ++c)
        for (int t = 0; t < CR_MAX_ROUND; ++t) {
            const uint64_t cnt = pl->h_hist[c * 128 + t];
            won += cnt;
            if (t == 0 || cnt == 0) continue;
            const int pass = t - (c + t) / cols + ((c + t) % cols == 0 ? 1 : 0);
            if (static_cast<int>(per_pass.size()) < pass) per_pass.resize(pass, 0);
            per_pass[pass - 1] += cnt;
        }
    per_pass.push_back(0);
    passes = per_pass;
    if (won_states) *won_states = won;
    if (levels) *levels = lv;
    return 0;
}

}  // namespace crd
