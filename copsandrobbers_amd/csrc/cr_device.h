// cr_device.h -- pieces shared by the device translation units of
// libcrsolve.so: launch constants, error reporting, the LDS graph image and the
// tuple rank/unrank/move helpers, the 64-bit exclusive scan and allocation.
// Internal header (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "crsolve.h"

namespace crd {
// Last error message of the calling thread (cr_last_error) and its setter;
// fail returns `code` so call sites can `return fail(...)`.
extern thread_local char g_err[512];
int fail(int code, const char* fmt, ...);

// Graph image copied into LDS by every block that ranks or unranks tuples:
//   rank_tab u32[K*n]   rank_tab[i*n + x] = C(M-1-(x+i), K-i), M = n+K-1
//   binom    u32[(K+1)*(M+1)]  binom[j*(M+1) + x] = C(x, j), saturated
//   in_off   u16[n+1]   in-neighbour offsets
//   in_nbr   u8[E]      in-neighbours (cop predecessors step back along them)
//   out_off  u16[n+1], out_nbr u8[E]: out-neighbours (forward team moves)
struct GraphView {
    const uint8_t* image;
    uint32_t image_bytes;
    uint32_t off_binom, off_in_off, off_in_nbr, off_out_off, off_out_nbr;
    int32_t n, W, M;
    uint32_t C;
    uint32_t ctot_m1;  // C - 1
};
}  // namespace crd

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr unsigned kMaxRound = CR_MAX_ROUND;
constexpr uint8_t kUnmarked = static_cast<uint8_t>(kMaxRound << 1);

using crd::fail;
using crd::g_err;

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(CR_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

using crd::GraphView;

// Copies the robber out-masks (n rows of W words) into LDS; returns the base.
template <int W>
__device__ __forceinline__ const uint64_t* stage_out_masks(uint64_t* lds_om, const uint64_t* out_mask, int n) {
    for (int i = threadIdx.x; i < n * W; i += blockDim.x) lds_om[i] = out_mask[(i / W) * 4 + (i % W)];
    __syncthreads();
    return lds_om;
}

__device__ __forceinline__ uint32_t wave_id() {
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6)));
}

__device__ __forceinline__ uint8_t won_byte(uint32_t d) { return static_cast<uint8_t>(1u | (d << 1)); }

__device__ __forceinline__ void load_image(uint8_t* lds, const uint8_t* src, uint32_t bytes) {
    const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
    uint32_t* d = reinterpret_cast<uint32_t*>(lds);
    for (uint32_t i = threadIdx.x; i < bytes / 4; i += blockDim.x) d[i] = s[i];
    __syncthreads();
}

// Combinatorial-number-system decode of a lexicographic rank.
template <int K>
__device__ __forceinline__ void unrank(uint32_t rank, const GraphView& gv, const uint32_t* binom,
                                       uint32_t* cops) {
    uint32_t s = gv.ctot_m1 - rank;
    int xmax = gv.M - 1;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const uint32_t* row = binom + (K - i) * (gv.M + 1);
        int lo = 0, hi = xmax;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (row[mid] <= s) lo = mid; else hi = mid - 1;
        }
        s -= row[lo];
        cops[i] = static_cast<uint32_t>(gv.M - 1 - lo - i);
        xmax = lo - 1;
    }
}

template <int K>
__device__ __forceinline__ void sort_tuple(uint32_t* t) {
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = 0; j + 1 < K - i; ++j) {
            const uint32_t a = t[j], b = t[j + 1];
            t[j] = min(a, b);
            t[j + 1] = max(a, b);
        }
}

template <int K>
__device__ __forceinline__ uint32_t rank_sorted(const uint32_t* t, const uint32_t* rank_tab, int n,
                                                uint32_t ctot_m1) {
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) s += rank_tab[i * n + t[i]];
    return ctot_m1 - s;
}

// Decodes product index t into one predecessor tuple (sorted) and returns
// false when the tuple is a duplicate produced by swapping two stacked cops.
template <int K>
__device__ __forceinline__ bool product_tuple(uint32_t t, const uint32_t* cops, const uint32_t* deg,
                                              const uint32_t* base, const uint8_t* nbr,
                                              uint32_t* tup) {
    uint32_t idx[K];
#pragma unroll
    for (int i = K - 1; i > 0; --i) {
        const uint32_t q = t / deg[i];
        idx[i] = t - q * deg[i];
        t = q;
    }
    idx[0] = t;
    bool keep = true;
#pragma unroll
    for (int i = 0; i + 1 < K; ++i)
        if (cops[i] == cops[i + 1] && idx[i] > idx[i + 1]) keep = false;
#pragma unroll
    for (int i = 0; i < K; ++i) tup[i] = nbr[base[i] + idx[i]];
    sort_tuple<K>(tup);
    return keep;
}


// rank of a sorted j-tuple over [0, n): C(M,j) - 1 - sum_i C(M-1-(t_i+i), j-i), M = n+j-1
__device__ __forceinline__ uint32_t rank_j(const uint32_t* t, int j, int n, const uint32_t* binom, int Mimg) {
    const int M = n + j - 1;
    uint32_t s = 0;
    for (int i = 0; i < j; ++i) s += binom[(j - i) * (Mimg + 1) + (M - 1 - static_cast<int>(t[i]) - i)];
    return binom[j * (Mimg + 1) + M] - 1 - s;
}

__device__ __forceinline__ void unrank_j(uint32_t rank, int j, int n, const uint32_t* binom, int Mimg, uint32_t* out) {
    const int M = n + j - 1;
    uint32_t s = binom[j * (Mimg + 1) + M] - 1 - rank;
    int xmax = M - 1;
    for (int i = 0; i < j; ++i) {
        const uint32_t* row = binom + (j - i) * (Mimg + 1);
        int lo = 0, hi = xmax;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (row[mid] <= s) lo = mid; else hi = mid - 1;
        }
        s -= row[lo];
        out[i] = static_cast<uint32_t>(M - 1 - lo - i);
        xmax = lo - 1;
    }
}

// Block-wide sums of wave-uniform per-wave values; thread 0 gets the totals.
// Every thread of the block must call it.
__device__ __forceinline__ void block_sum2(unsigned long long& x, unsigned long long& y) {
    __shared__ unsigned long long red[2][kWavesPerBlock];
    __syncthreads();  // previous use of red is finished
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = x;
        red[1][threadIdx.x >> 6] = y;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        x = y = 0;
        for (int i = 0; i < kWavesPerBlock; ++i) {
            x += red[0][i];
            y += red[1][i];
        }
    }
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

constexpr int kScanThreads = 1024;

// ---------------------------------------------------------------------------
// exclusive scan of per-configuration counts into 64-bit CSR heads
// ---------------------------------------------------------------------------
constexpr int kScanTile = kScanThreads * 8;

__device__ unsigned long long block_exclusive_scan(unsigned long long v, unsigned long long& total) {
    __shared__ unsigned long long sh[kScanThreads];
    const uint32_t t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
        const unsigned long long x = t >= off ? sh[t - off] : 0ull;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    total = sh[kScanThreads - 1];
    const unsigned long long r = sh[t] - v;
    __syncthreads();
    return r;
}

[[maybe_unused]] __global__ __launch_bounds__(kScanThreads) void k_tile_sums(const uint32_t* cnt, uint64_t n,
                                                            unsigned long long* tiles) {
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * 8ull;
    unsigned long long v = 0;
    for (int i = 0; i < 8; ++i)
        if (base + i < n) v += cnt[base + i];
    unsigned long long total;
    block_exclusive_scan(v, total);
    if (threadIdx.x == 0) tiles[blockIdx.x] = total;
}

// Scans the tile sums in place and stores the grand total as heads[n].
[[maybe_unused]] __global__ __launch_bounds__(kScanThreads) void k_scan_tiles(unsigned long long* tiles, uint32_t ntiles,
                                                             unsigned long long* heads, uint64_t n) {
    unsigned long long carry = 0;
    for (uint32_t b0 = 0; b0 < ntiles; b0 += kScanThreads) {
        const uint32_t i = b0 + threadIdx.x;
        const unsigned long long v = i < ntiles ? tiles[i] : 0ull;
        unsigned long long total;
        const unsigned long long ex = block_exclusive_scan(v, total);
        if (i < ntiles) tiles[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) heads[n] = carry;
}

[[maybe_unused]] __global__ __launch_bounds__(kScanThreads) void k_tile_apply(const uint32_t* cnt, uint64_t n,
                                                             const unsigned long long* tiles,
                                                             unsigned long long* heads) {
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * 8ull;
    unsigned long long v = 0;
    for (int i = 0; i < 8; ++i)
        if (base + i < n) v += cnt[base + i];
    unsigned long long total;
    unsigned long long run = tiles[blockIdx.x] + block_exclusive_scan(v, total);
    for (int i = 0; i < 8; ++i) {
        if (base + i < n) {
            heads[base + i] = run;
            run += cnt[base + i];
        }
    }
}

template <typename T>
T* dev_alloc(uint64_t count, uint64_t& total, int& rc) {
    void* p = nullptr;
    if (count == 0) count = 1;
    const hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
        rc = fail(CR_ERR_DEVICE, "hipMalloc(%llu bytes): %s",
                  static_cast<unsigned long long>(count * sizeof(T)), hipGetErrorString(e));
        return nullptr;
    }
    total += count * sizeof(T);
    return static_cast<T*>(p);
}

// Kernels whose dynamic LDS (the graph image) exceeds the 64 KiB default need
// the per-function limit raised before launch.
inline int raise_lds_limit(const void* fn, uint32_t bytes) {
    if (bytes <= 64 * 1024) return 0;
    HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
    return 0;
}

}  // namespace
