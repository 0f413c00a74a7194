// game.hip -- perfect-game extraction over a solved column table
// (legacy/python/k_cops.py:174-241; the C++ round solver's version is
// legacy/src/k_cops_rounds.cpp:307-397).
//
// Values are rounds: a cop-turn state won at depth t has (t + 1) / 2 rounds
// left (two plies per round, k_cops_rounds.cpp:272). The start is the
// universal configuration with the least worst case (lowest cId on ties) and
// the robber vertex with the most rounds (lowest r on ties). Then, until the
// robber is caught:
//   cop turn     every team move is scored by the robber's best reply: 0 if it
//                lands on the robber, else the most rounds over the robber's
//                options, invalid if one of them is not won; the first move
//                with the least score in the reference's option order wins
//                (itertools.product over [neighbours ascending, stay], last
//                cop fastest); no valid move leaves the cops where they are;
//   robber turn  the first option with the most rounds among won states
//                (strictly greater), else the robber stays.
// A cop step is one workgroup: threads score the team moves, a block-wide
// argmin over (score, product index) keeps the reference's tie order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "cr_game.h"

namespace {

constexpr int kWalkBlock = 256;

struct WalkArgs {
    const uint8_t* table;
    const uint16_t* cop_off;  // [n+1] into cop_opt: each cop's options in the reference's order
    const uint8_t* cop_opt;
    const uint16_t* rob_off;  // [n+1] into rob_opt: the robber's options
    const uint8_t* rob_opt;
    uint32_t start_cid;
    int32_t start_r;
    int32_t columns;
    int32_t max_iters;        // k_cops.py:192 hard limit: 5 * n loop iterations
    int32_t by_cid;           // 1: ties between equal moves go to the lower next cId (k_cops_rounds.cpp:331-368)
    crd::DevGameStep* out;    // capacity 2 * max_iters + 2
    int32_t* length;
};

__device__ __forceinline__ uint32_t rounds_of(uint8_t b) { return ((b >> 1) + 1u) >> 1; }

// One wavefront per configuration row: universal rows (every robber vertex won
// in column 0) give key = worst rounds << 32 | cId, minimised.
__global__ __launch_bounds__(kBlock) void k_game_start(const uint8_t* table, uint32_t C, int32_t n, int32_t columns,
                                                       unsigned long long* best) {
    const int lane = threadIdx.x & 63;
    for (uint64_t row = static_cast<uint64_t>(blockIdx.x) * kWavesPerBlock + wave_id(); row < C;
         row += static_cast<uint64_t>(gridDim.x) * kWavesPerBlock) {
        bool universal = true;
        uint32_t worst = 0;
        for (int r = lane; r < n; r += 64) {
            const uint8_t b = table[(row * n + r) * columns];
            universal = universal && (b & 1);
            worst = max(worst, rounds_of(b));
        }
        const bool all = __ballot(!universal) == 0;
        for (int off = 32; off > 0; off >>= 1) worst = max(worst, static_cast<uint32_t>(__shfl_xor(worst, off)));
        if (lane == 0 && all) atomicMin(best, (static_cast<unsigned long long>(worst) << 32) | row);
    }
}

template <int K>
__device__ __forceinline__ void emit(const WalkArgs& a, int32_t& len, const uint32_t* cops, int32_t r, int32_t turn) {
    if (threadIdx.x == 0) {
        crd::DevGameStep st{};
        for (int i = 0; i < K; ++i) st.cops[i] = static_cast<uint8_t>(cops[i]);
        st.robber = r;
        st.turn = turn;
        a.out[len] = st;
    }
    ++len;
}

template <int K>
__global__ __launch_bounds__(kWalkBlock) void k_game_walk(GraphView gv, WalkArgs a) {
    __shared__ unsigned long long red[kWalkBlock / 64];
    const uint32_t* rank_tab = reinterpret_cast<const uint32_t*>(gv.image);
    const uint32_t* binom = reinterpret_cast<const uint32_t*>(gv.image + gv.off_binom);
    const int n = gv.n, cols = a.columns;
    uint32_t cid = a.start_cid;
    int32_t r = a.start_r, len = 0;
    for (int32_t it = 0; it < a.max_iters; ++it) {
        uint32_t cops[K];
        unrank<K>(cid, gv, binom, cops);
        bool caught = false;
#pragma unroll
        for (int i = 0; i < K; ++i) caught = caught || static_cast<int32_t>(cops[i]) == r;
        if (caught) {  // the robber walked onto a cop: k_cops_rounds.cpp:325-327 records it, k_cops.py does not
            if (a.by_cid) emit<K>(a, len, cops, r, 2);
            break;
        }
        emit<K>(a, len, cops, r, 0);
        // cop turn: minimise the robber's best reply
        uint32_t base[K], deg[K], P = 1;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            base[i] = a.cop_off[cops[i]];
            deg[i] = a.cop_off[cops[i] + 1] - base[i];
            P *= deg[i];
        }
        unsigned long long best = ~0ull;
        for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) {
            uint32_t tup[K], q = t;
#pragma unroll
            for (int i = K - 1; i >= 0; --i) {  // the last cop varies fastest
                tup[i] = a.cop_opt[base[i] + q % deg[i]];
                q /= deg[i];
            }
            bool hit = false;
#pragma unroll
            for (int i = 0; i < K; ++i) hit = hit || static_cast<int32_t>(tup[i]) == r;
            sort_tuple<K>(tup);
            const uint32_t nc = rank_sorted<K>(tup, rank_tab, n, gv.ctot_m1);
            uint32_t score = 1;  // score = rounds + 1; a capture scores 0 + 1
            bool ok = true;
            if (!hit) {
                int32_t wr = -1;
                for (int e = a.rob_off[r]; e < a.rob_off[r + 1]; ++e) {
                    const uint8_t b = a.table[(static_cast<uint64_t>(nc) * n + a.rob_opt[e]) * cols];
                    if (!(b & 1)) {
                        ok = false;
                        break;
                    }
                    wr = max(wr, static_cast<int32_t>(rounds_of(b)));
                }
                score = static_cast<uint32_t>(wr + 1);
            }
            if (ok) {
                const unsigned long long key = (static_cast<unsigned long long>(score) << 32) | (a.by_cid ? nc : t);
                best = key < best ? key : best;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(best, off);
            best = o < best ? o : best;
        }
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
        __syncthreads();
        best = ~0ull;
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) best = red[w] < best ? red[w] : best;
        __syncthreads();  // red is rewritten next step
        if (best != ~0ull) {  // otherwise the cops stay where they are
            if (a.by_cid) {
                cid = static_cast<uint32_t>(best & 0xffffffffu);
            } else {
                uint32_t tup[K], q = static_cast<uint32_t>(best & 0xffffffffu);
#pragma unroll
                for (int i = K - 1; i >= 0; --i) {
                    tup[i] = a.cop_opt[base[i] + q % deg[i]];
                    q /= deg[i];
                }
                sort_tuple<K>(tup);
                cid = rank_sorted<K>(tup, rank_tab, n, gv.ctot_m1);
            }
        }
        unrank<K>(cid, gv, binom, cops);
        caught = false;
#pragma unroll
        for (int i = 0; i < K; ++i) caught = caught || static_cast<int32_t>(cops[i]) == r;
        if (caught) {
            emit<K>(a, len, cops, r, 2);
            break;
        }
        emit<K>(a, len, cops, r, 1);
        // robber turn: the first option with the most rounds among won states
        int32_t next = r, most = -1;
        for (int e = a.rob_off[r]; e < a.rob_off[r + 1]; ++e) {
            const uint8_t b = a.table[(static_cast<uint64_t>(cid) * n + a.rob_opt[e]) * cols];
            if ((b & 1) && static_cast<int32_t>(rounds_of(b)) > most) {
                most = static_cast<int32_t>(rounds_of(b));
                next = a.rob_opt[e];
            }
        }
        r = next;
    }
    if (threadIdx.x == 0) *a.length = len;
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

// Options of every vertex in the reference's order: neighbours ascending
// (matrix row), then the vertex itself when staying is allowed and the
// matrix has no self-loop (k_cops.py:33-36 appends the self-loop).
void option_lists(const std::vector<uint8_t>& matrix, int n, bool stay, std::vector<uint16_t>& off,
                  std::vector<uint8_t>& opt) {
    off.assign(static_cast<size_t>(n) + 1, 0);
    opt.clear();
    for (int u = 0; u < n; ++u) {
        off[u] = static_cast<uint16_t>(opt.size());
        for (int v = 0; v < n; ++v)
            if (matrix[static_cast<size_t>(u) * n + v]) opt.push_back(static_cast<uint8_t>(v));
        if (stay && !matrix[static_cast<size_t>(u) * n + u]) opt.push_back(static_cast<uint8_t>(u));
    }
    off[n] = static_cast<uint16_t>(opt.size());
}

}  // namespace

namespace crd {

int game_extract(const GraphView& gv, int k, int columns, const uint8_t* d_table, const std::vector<uint8_t>& matrix,
                 bool cop_stay, bool robber_stay, bool cpp_order, hipStream_t st, std::vector<cr::GameStep>& out) {
    out.clear();
    const int n = gv.n;
    std::vector<uint16_t> coff, roff;
    std::vector<uint8_t> copt, ropt;
    option_lists(matrix, n, cop_stay, coff, copt);
    option_lists(matrix, n, robber_stay && !cpp_order, roff, ropt);
    // k_cops.py stops after 5n rounds; k_cops_rounds.cpp has no limit, and an
    // optimal game has at most 64 rounds under the 127-ply cap
    const int32_t max_iters = cpp_order ? std::max(5 * n, 256) : 5 * n;
    const size_t cap = 2 * static_cast<size_t>(max_iters) + 2;
    // one device allocation: [best u64][length i32 pad][steps][cop_off][rob_off][cop_opt][rob_opt]
    const size_t o_steps = 16, o_coff = o_steps + cap * sizeof(DevGameStep);
    const size_t o_roff = o_coff + coff.size() * 2, o_copt = o_roff + roff.size() * 2;
    const size_t o_ropt = o_copt + copt.size(), total = o_ropt + ropt.size() + 8;
    uint8_t* d = nullptr;
    HIP_TRY(hipMalloc(&d, total));
    auto cleanup = [&] { hipFree(d); };
    const unsigned long long init = ~0ull;
    int rc = 0;
    do {
        if (hipMemcpyAsync(d, &init, 8, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(d + o_coff, coff.data(), coff.size() * 2, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(d + o_roff, roff.data(), roff.size() * 2, hipMemcpyHostToDevice, st) != hipSuccess ||
            (!copt.empty() && hipMemcpyAsync(d + o_copt, copt.data(), copt.size(), hipMemcpyHostToDevice, st) != hipSuccess) ||
            (!ropt.empty() && hipMemcpyAsync(d + o_ropt, ropt.data(), ropt.size(), hipMemcpyHostToDevice, st) != hipSuccess)) {
            rc = fail(CR_ERR_DEVICE, "game inputs: copy failed");
            break;
        }
        const uint64_t waves = gv.C;
        const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((waves + kWavesPerBlock - 1) / kWavesPerBlock, 8192)));
        hipLaunchKernelGGL(k_game_start, dim3(grid), dim3(kBlock), 0, st, d_table, gv.C, n, columns,
                           reinterpret_cast<unsigned long long*>(d));
        unsigned long long best = 0;
        if (hipGetLastError() != hipSuccess || hipMemcpyAsync(&best, d, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rc = fail(CR_ERR_DEVICE, "game start: device error");
            break;
        }
        if (best == ~0ull) break;  // no configuration wins: no game
        const uint32_t cid = static_cast<uint32_t>(best & 0xffffffffu);
        std::vector<uint8_t> row(static_cast<size_t>(n) * columns);
        if (hipMemcpyAsync(row.data(), d_table + static_cast<uint64_t>(cid) * n * columns, row.size(),
                           hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
            rc = fail(CR_ERR_DEVICE, "game start row: copy failed");
            break;
        }
        int32_t r0 = -1, most = -1;
        for (int r = 0; r < n; ++r) {
            const uint8_t b = row[static_cast<size_t>(r) * columns];
            if ((b & 1) && static_cast<int32_t>(((b >> 1) + 1) >> 1) > most) {
                most = ((b >> 1) + 1) >> 1;
                r0 = r;
            }
        }
        WalkArgs a{d_table,
                   reinterpret_cast<const uint16_t*>(d + o_coff),
                   d + o_copt,
                   reinterpret_cast<const uint16_t*>(d + o_roff),
                   d + o_ropt,
                   cid,
                   r0,
                   columns,
                   max_iters,
                   cpp_order ? 1 : 0,
                   reinterpret_cast<DevGameStep*>(d + o_steps),
                   reinterpret_cast<int32_t*>(d + 8)};
        rc = with_k(k, [&](auto kc) -> int {
            constexpr int K = decltype(kc)::value;
            hipLaunchKernelGGL(k_game_walk<K>, dim3(1), dim3(kWalkBlock), 0, st, gv, a);
            return 0;
        });
        if (rc) break;
        int32_t len = 0;
        if (hipGetLastError() != hipSuccess || hipMemcpyAsync(&len, d + 8, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rc = fail(CR_ERR_DEVICE, "game walk: device error");
            break;
        }
        len = std::max(0, std::min<int32_t>(len, static_cast<int32_t>(cap)));
        std::vector<DevGameStep> steps(len);
        if (len && (hipMemcpy(steps.data(), d + o_steps, len * sizeof(DevGameStep), hipMemcpyDeviceToHost) != hipSuccess)) {
            rc = fail(CR_ERR_DEVICE, "game walk: copy failed");
            break;
        }
        for (const DevGameStep& s : steps) {
            cr::GameStep g{};
            std::memcpy(g.cops, s.cops, sizeof s.cops);
            g.robber = s.robber;
            g.turn = s.turn;
            out.push_back(g);
        }
    } while (false);
    cleanup();
    return rc;
}

}  // namespace crd
