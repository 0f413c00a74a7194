// cr_host.cpp -- host-side pieces of the solver (see cr_host.h).
#include "cr_host.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <thread>

#include "crsolve.h"

namespace cr {

int32_t parse_matrix(const uint8_t* text, uint64_t len, uint8_t* adj, int32_t cap) {
    if (text == nullptr || len == 0) return 0;  // fileio.cpp:12 rejects empty files
    uint64_t width = 0;  // Graph.cpp:27-30: width of the first line
    while (width < len && text[width] != '\n' && text[width] != '\r' && text[width] != '-') ++width;
    if (width == 0 || width > 0x7fffffff) return 0;
    const int32_t n = static_cast<int32_t>(width);
    if (adj == nullptr || n > cap) return n;
    std::memset(adj, 0, static_cast<size_t>(n) * n);
    int32_t row = 0, col = 0;
    for (uint64_t i = 0; i < len; ++i) {
        const uint8_t ch = text[i];
        if (ch == '-') break;
        if (ch == '0' || ch == '1') {
            if (row < n && col < n) adj[static_cast<size_t>(row) * n + col] = (ch == '1');
            ++col;
        } else if (ch == '\n' && col > 0) {
            ++row;
            col = 0;
        }
    }
    return n;
}

void Adjacency::build(int32_t n_, const uint8_t* adj) {
    n = n_;
    matrix.assign(adj, adj + static_cast<size_t>(n) * n);
    out.assign(n, {});
    in.assign(n, {});
    for (int32_t u = 0; u < n; ++u)
        for (int32_t v = 0; v < n; ++v)
            if (adj[static_cast<size_t>(u) * n + v]) {
                out[u].push_back(static_cast<uint8_t>(v));
                in[v].push_back(static_cast<uint8_t>(u));
            }
    max_out = max_in = 0;
    for (int32_t v = 0; v < n; ++v) {
        max_out = std::max<int32_t>(max_out, static_cast<int32_t>(out[v].size()));
        max_in = std::max<int32_t>(max_in, static_cast<int32_t>(in[v].size()));
    }
}

uint64_t binom(int64_t a, int64_t b) {
    if (b < 0 || a < b) return 0;
    if (b > a - b) b = a - b;
    unsigned __int128 acc = 1;
    for (int64_t i = 1; i <= b; ++i) acc = acc * static_cast<uint64_t>(a - b + i) / static_cast<uint64_t>(i);
    return acc > UINT64_MAX ? UINT64_MAX : static_cast<uint64_t>(acc);
}

uint64_t config_count(int32_t n, int32_t k) {
    if (k <= 0 || k > 256 || n <= 0) return 0;
    return binom(static_cast<int64_t>(n) + k - 1, k);
}

// Non-decreasing tuples c_0 <= ... <= c_{k-1} over [0, n) are the strictly
// increasing subsets e_i = c_i + i of [0, M), M = n+k-1, in the same
// lexicographic order. The subsets after e number
// sum_i C(M-1-e_i, k-i) (combinatorial number system), so
// rank = C(M, k) - 1 - that sum.
uint64_t rank_tuple(const uint8_t* c, int32_t n, int32_t k) {
    const int64_t M = static_cast<int64_t>(n) + k - 1;
    uint64_t tail = 0;
    for (int32_t i = 0; i < k; ++i) tail += binom(M - 1 - (static_cast<int64_t>(c[i]) + i), k - i);
    return binom(M, k) - 1 - tail;
}

void unrank_tuple(uint64_t rank, int32_t n, int32_t k, uint8_t* out) {
    const int64_t M = static_cast<int64_t>(n) + k - 1;
    uint64_t s = binom(M, k) - 1 - rank;
    int64_t xmax = M - 1;
    for (int32_t i = 0; i < k; ++i) {
        const int32_t j = k - i;
        int64_t lo = 0, hi = xmax;  // largest x <= xmax with C(x, j) <= s
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) / 2;
            if (binom(mid, j) <= s) lo = mid; else hi = mid - 1;
        }
        s -= binom(lo, j);
        out[i] = static_cast<uint8_t>((M - 1 - lo) - i);
        xmax = lo - 1;
    }
}

void team_moves(const std::vector<std::vector<uint8_t>>& nbr, const uint8_t* cops, int32_t k,
                int32_t n, bool stay, std::vector<uint64_t>& out) {
    out.clear();
    std::vector<std::vector<uint8_t>> opt(k);
    for (int32_t i = 0; i < k; ++i) {
        if (stay) opt[i].push_back(cops[i]);
        opt[i].insert(opt[i].end(), nbr[cops[i]].begin(), nbr[cops[i]].end());
        if (opt[i].empty()) return;
    }
    std::vector<int32_t> digit(k, 0);
    uint8_t t[CR_MAX_K > 16 ? CR_MAX_K : 16];
    for (;;) {
        for (int32_t i = 0; i < k; ++i) t[i] = opt[i][digit[i]];
        std::sort(t, t + k);
        out.push_back(rank_tuple(t, n, k));
        int32_t j = k - 1;
        while (j >= 0 && ++digit[j] == static_cast<int32_t>(opt[j].size())) digit[j--] = 0;
        if (j < 0) break;
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}

std::string cache_key(const std::string& algo, const std::string& graph_path, int32_t k, int32_t p,
                      bool cop_stay, bool robber_stay) {
    const std::string stem = std::filesystem::path(graph_path).stem().string();
    const std::string flags = std::string(cop_stay ? "1" : "0") + (robber_stay ? "1" : "0");
    return "cache/" + algo + "__" + stem + "__" + std::to_string(k) + "-cops__" + std::to_string(p) +
           "-vis__" + flags + "__auxgraph.bin";
}

namespace {

unsigned worker_count() {
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 1;
    return std::min(h, 16u);
}

template <typename Fn>
void parallel_ranges(uint64_t total, Fn fn) {
    const unsigned nt = worker_count();
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        const uint64_t lo = total * t / nt, hi = total * (t + 1) / nt;
        th.emplace_back([=] { fn(lo, hi); });
    }
    for (auto& x : th) x.join();
}

bool put(std::FILE* f, const void* p, size_t bytes) {
    return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes;
}

}  // namespace

int write_aux_blob(const std::string& path, const Adjacency& g, int32_t k, int32_t columns,
                   bool cop_stay, bool robber_stay, uint64_t num_states,
                   const std::function<int(uint64_t, uint64_t, uint8_t*)>& fetch_states) {
    const uint64_t C = config_count(g.n, k);
    const uint64_t N = static_cast<uint64_t>(g.n);
    // pass 1: |T(C)| per configuration -> heads
    std::vector<uint64_t> heads(C + 1, 0);
    parallel_ranges(C, [&](uint64_t lo, uint64_t hi) {
        std::vector<uint64_t> mv;
        uint8_t cops[CR_MAX_K];
        for (uint64_t c = lo; c < hi; ++c) {
            unrank_tuple(c, g.n, k, cops);
            team_moves(g.out, cops, k, g.n, cop_stay, mv);
            heads[c + 1] = mv.size();
        }
    });
    for (uint64_t c = 0; c < C; ++c) heads[c + 1] += heads[c];
    const uint64_t ntrans = heads[C];
    const uint64_t header_bytes = 12 * sizeof(uint64_t);
    const uint64_t s1 = header_bytes;
    const uint64_t s2 = s1 + C * static_cast<uint64_t>(k);
    const uint64_t s3 = s2 + (C + 1) * sizeof(uint64_t);
    const uint64_t s4 = s3 + ntrans * sizeof(uint64_t);
    const uint64_t header[12] = {static_cast<uint64_t>(k), N, static_cast<uint64_t>(columns), C,
                                 num_states, 1, cop_stay ? 1u : 0u, robber_stay ? 1u : 0u,
                                 s1, s2, s3, s4};
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return CR_ERR_IO;
    bool ok = put(f, header, sizeof header);
    {  // configs, in chunks
        const uint64_t chunk = 1 << 20;
        std::vector<uint8_t> buf;
        for (uint64_t c0 = 0; ok && c0 < C; c0 += chunk) {
            const uint64_t c1 = std::min(C, c0 + chunk);
            buf.resize((c1 - c0) * k);
            parallel_ranges(c1 - c0, [&](uint64_t lo, uint64_t hi) {
                for (uint64_t c = lo; c < hi; ++c) unrank_tuple(c0 + c, g.n, k, &buf[c * k]);
            });
            ok = put(f, buf.data(), buf.size());
        }
    }
    ok = ok && put(f, heads.data(), heads.size() * sizeof(uint64_t));
    {  // transitions (nextId * N), in chunks of configurations
        const uint64_t chunk = 1 << 16;
        std::vector<uint64_t> buf;
        for (uint64_t c0 = 0; ok && c0 < C; c0 += chunk) {
            const uint64_t c1 = std::min(C, c0 + chunk);
            buf.assign(heads[c1] - heads[c0], 0);
            parallel_ranges(c1 - c0, [&](uint64_t lo, uint64_t hi) {
                std::vector<uint64_t> mv;
                uint8_t cops[CR_MAX_K];
                for (uint64_t c = c0 + lo; c < c0 + hi; ++c) {
                    unrank_tuple(c, g.n, k, cops);
                    team_moves(g.out, cops, k, g.n, cop_stay, mv);
                    uint64_t* dst = &buf[heads[c] - heads[c0]];
                    for (size_t i = 0; i < mv.size(); ++i) dst[i] = mv[i] * N;
                }
            });
            ok = put(f, buf.data(), buf.size() * sizeof(uint64_t));
        }
    }
    {  // states, fetched from the device in chunks
        const uint64_t chunk = 64ull << 20;
        std::vector<uint8_t> buf;
        for (uint64_t o = 0; ok && o < num_states; o += chunk) {
            const uint64_t cnt = std::min(chunk, num_states - o);
            buf.resize(cnt);
            if (fetch_states(o, cnt, buf.data()) != 0) ok = false;
            ok = ok && put(f, buf.data(), cnt);
        }
    }
    if (std::fclose(f) != 0) ok = false;
    return ok ? 0 : CR_ERR_IO;
}

namespace {

struct InFile {
    std::FILE* f;
    explicit InFile(const std::string& p) : f(std::fopen(p.c_str(), "rb")) {}
    ~InFile() {
        if (f) std::fclose(f);
    }
};

bool get(std::FILE* f, void* p, size_t bytes) {
    return bytes == 0 || std::fread(p, 1, bytes, f) == bytes;
}

bool seek_to(std::FILE* f, uint64_t off) {
    return fseeko(f, static_cast<off_t>(off), SEEK_SET) == 0;
}

}  // namespace

int read_aux_header(const std::string& path, uint64_t header[12], uint64_t* file_size, std::string& why) {
    InFile in(path);
    if (!in.f) {
        why = "cannot open " + path;
        return CR_ERR_IO;
    }
    if (fseeko(in.f, 0, SEEK_END) != 0) {
        why = "cannot seek in " + path;
        return CR_ERR_IO;
    }
    const off_t end = ftello(in.f);
    if (end < 0 || !seek_to(in.f, 0)) {
        why = "cannot seek in " + path;
        return CR_ERR_IO;
    }
    const uint64_t size = static_cast<uint64_t>(end);
    if (size < 12 * sizeof(uint64_t) || !get(in.f, header, 12 * sizeof(uint64_t))) {
        why = path + " is shorter than the 96-byte header";
        return CR_ERR_IO;
    }
    // header | configs | transition heads | transitions | states, in this order
    const uint64_t s1 = header[8], s2 = header[9], s3 = header[10], s4 = header[11];
    if (s1 != 12 * sizeof(uint64_t) || s2 < s1 || s3 < s2 || s4 < s3 || s4 > size || (s3 - s2) % 8 != 0 ||
        (s4 - s3) % 8 != 0) {
        why = "section offsets " + std::to_string(s1) + ", " + std::to_string(s2) + ", " + std::to_string(s3) + ", " +
              std::to_string(s4) + " do not fit a file of " + std::to_string(size) + " bytes";
        return CR_ERR_FORMAT;
    }
    if (header[5] == 0 || (size - s4) % header[5] != 0 || header[6] > 1 || header[7] > 1) {
        why = "item size " + std::to_string(header[5]) + " or self-edge flags " + std::to_string(header[6]) +
              std::to_string(header[7]) + " out of range";
        return CR_ERR_FORMAT;
    }
    if (file_size) *file_size = size;
    return 0;
}

int read_aux_blob(const std::string& path, const Adjacency& g, int32_t k, int32_t columns, bool cop_stay,
                  bool robber_stay, uint64_t num_states, std::string& why,
                  const std::function<int(uint64_t, uint64_t, const uint8_t*)>& put_states) {
    uint64_t h[12];
    uint64_t size = 0;
    if (int rc = read_aux_header(path, h, &size, why)) return rc;
    if (h[6] != (cop_stay ? 1u : 0u) || h[7] != (robber_stay ? 1u : 0u)) {
        why = "Cache rules mismatch: Cached SelfEdges [" + std::to_string(h[6]) + std::to_string(h[7]) +
              "] vs Requested [" + (cop_stay ? "1" : "0") + (robber_stay ? "1" : "0") + "]";
        return CR_ERR_CACHE_RULES;
    }
    const uint64_t C = config_count(g.n, k);
    const uint64_t want[6] = {static_cast<uint64_t>(k), static_cast<uint64_t>(g.n), static_cast<uint64_t>(columns),
                              C, num_states, 1};
    static const char* const names[6] = {"k", "N", "columns", "configCount", "numStates", "sizeof(DataItem)"};
    for (int i = 0; i < 6; ++i)
        if (h[i] != want[i]) {
            why = std::string("header field ") + names[i] + " is " + std::to_string(h[i]) + "; the loaded graph has " +
                  std::to_string(want[i]);
            return CR_ERR_FORMAT;
        }
    if (h[9] != h[8] + C * static_cast<uint64_t>(k) || h[10] != h[9] + (C + 1) * sizeof(uint64_t) ||
        size - h[11] != num_states) {
        why = "section sizes do not match " + std::to_string(C) + " configurations and " + std::to_string(num_states) +
              " states";
        return CR_ERR_FORMAT;
    }
    InFile in(path);
    if (!in.f || !seek_to(in.f, h[8])) {
        why = "cannot read " + path;
        return CR_ERR_IO;
    }
    {  // configs: the ranking's order (AuxGraph.h:155-176), in chunks
        const uint64_t chunk = 1 << 20;
        std::vector<uint8_t> buf;
        for (uint64_t c0 = 0; c0 < C; c0 += chunk) {
            const uint64_t c1 = std::min(C, c0 + chunk);
            buf.resize((c1 - c0) * k);
            if (!get(in.f, buf.data(), buf.size())) {
                why = "cannot read the configs section of " + path;
                return CR_ERR_IO;
            }
            std::atomic<uint64_t> bad{~0ull};
            parallel_ranges(c1 - c0, [&](uint64_t lo, uint64_t hi) {
                uint8_t t[CR_MAX_K];
                for (uint64_t c = lo; c < hi; ++c) {
                    unrank_tuple(c0 + c, g.n, k, t);
                    if (std::memcmp(t, &buf[c * k], k) != 0) {
                        uint64_t cur = bad.load();
                        while (c0 + c < cur && !bad.compare_exchange_weak(cur, c0 + c)) {
                        }
                        return;
                    }
                }
            });
            if (bad.load() != ~0ull) {
                why = "configuration " + std::to_string(bad.load()) + " differs from the ranking's";
                return CR_ERR_FORMAT;
            }
        }
    }
    uint64_t head0 = 1, headC = 0;
    if (!get(in.f, &head0, sizeof head0) || !seek_to(in.f, h[9] + C * sizeof(uint64_t)) ||
        !get(in.f, &headC, sizeof headC)) {
        why = "cannot read the transition heads of " + path;
        return CR_ERR_IO;
    }
    if (head0 != 0 || headC != (h[11] - h[10]) / sizeof(uint64_t)) {
        why = "transition heads (" + std::to_string(head0) + " .. " + std::to_string(headC) +
              ") do not span the transitions section";
        return CR_ERR_FORMAT;
    }
    if (!seek_to(in.f, h[11])) {
        why = "cannot seek to the states of " + path;
        return CR_ERR_IO;
    }
    const uint64_t chunk = 64ull << 20;
    std::vector<uint8_t> buf;
    for (uint64_t o = 0; o < num_states; o += chunk) {
        const uint64_t cnt = std::min(chunk, num_states - o);
        buf.resize(cnt);
        if (!get(in.f, buf.data(), cnt)) {
            why = "cannot read the states section of " + path;
            return CR_ERR_IO;
        }
        if (int rc = put_states(o, cnt, buf.data())) {
            why = "could not store the states";
            return rc;
        }
    }
    return 0;
}

int32_t parse_edge_list(const uint8_t* text, uint64_t len, uint8_t* adj, int32_t cap) {
    if (text == nullptr) return 0;
    std::vector<std::pair<uint64_t, uint64_t>> edges;
    uint64_t max_id = 0;
    auto is_space = [](uint8_t c) { return c == ' ' || c == '\t' || c == '\v' || c == '\f'; };
    uint64_t i = 0;
    while (i < len) {
        uint64_t end = i;
        while (end < len && text[end] != '\n' && text[end] != '\r') ++end;  // universal newlines
        uint64_t a = i, b = end;
        while (a < b && is_space(text[a])) ++a;
        while (b > a && is_space(text[b - 1])) --b;
        if (a < b && text[a] != '#') {
            // tokens split on whitespace and commas; only all-digit tokens count
            uint64_t nums[2];
            int got = 0;
            uint64_t t = a;
            while (t < b && got < 2) {
                while (t < b && (is_space(text[t]) || text[t] == ',')) ++t;
                uint64_t e = t;
                while (e < b && !is_space(text[e]) && text[e] != ',') ++e;
                if (e > t) {
                    bool digits = true;
                    uint64_t v = 0;
                    for (uint64_t q = t; q < e; ++q) {
                        if (text[q] < '0' || text[q] > '9') { digits = false; break; }
                        v = v > 100000000ull ? v : v * 10 + (text[q] - '0');  // saturates far above any graph size
                    }
                    if (digits) nums[got++] = v;
                }
                t = e;
            }
            if (got == 2) {
                edges.emplace_back(nums[0], nums[1]);
                max_id = std::max(max_id, std::max(nums[0], nums[1]));
            }
        }
        i = end + 1;
    }
    if (edges.empty() || max_id == 0) return 0;
    if (max_id > 0x7fffffffull) return 0;
    const int32_t n = static_cast<int32_t>(max_id);
    if (adj == nullptr || n > cap) return n;
    std::memset(adj, 0, static_cast<size_t>(n) * n);
    for (auto [u, v] : edges) {
        const uint64_t iu = u == 0 ? static_cast<uint64_t>(n - 1) : u - 1;  // Python index -1
        const uint64_t iv = v == 0 ? static_cast<uint64_t>(n - 1) : v - 1;
        adj[iu * n + iv] = 1;
        adj[iv * n + iu] = 1;
    }
    return n;
}

int write_dp_table(const std::string& path, int32_t n, int32_t k, int32_t columns, uint64_t configs,
                   const std::function<int(uint64_t, uint64_t, uint8_t*)>& fetch_states) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return CR_ERR_IO;
    std::vector<uint8_t> cur(static_cast<size_t>(k), 0);  // lexicographic odometer (AuxGraph.h:155-176)
    const uint64_t chunk = std::max<uint64_t>(1, (4ull << 20) / (static_cast<uint64_t>(n) * columns));
    std::vector<uint8_t> rows;
    std::string out;
    out.reserve(1 << 22);
    out.push_back('[');
    bool first = true;
    int rc = 0;
    char buf[64];
    for (uint64_t c0 = 0; c0 < configs && rc == 0; c0 += chunk) {
        const uint64_t cnt = std::min(chunk, configs - c0);
        rows.resize(cnt * n * columns);
        if (fetch_states(c0 * n * columns, rows.size(), rows.data()) != 0) {
            rc = CR_ERR_DEVICE;
            break;
        }
        for (uint64_t c = 0; c < cnt; ++c) {
            std::string cops = "[";
            for (int32_t i = 0; i < k; ++i) {
                std::snprintf(buf, sizeof buf, i ? ", %d" : "%d", cur[i]);
                cops += buf;
            }
            cops += "]";
            for (int32_t r = 0; r < n; ++r) {
                const uint8_t b = rows[(c * n + r) * columns];
                const bool win = b & 1;
                const int steps = win ? ((b >> 1) + 1) / 2 : -1;
                if (!first) out += ", ";
                first = false;
                out += "{\"cops\": ";
                out += cops;
                std::snprintf(buf, sizeof buf, ", \"robber\": %d, \"cop_wins\": %s, \"steps_to_win\": %d}", r,
                              win ? "true" : "false", steps);
                out += buf;
            }
            // next configuration
            int32_t q = k - 1;
            while (q >= 0 && cur[q] == n - 1) --q;
            if (q >= 0) {
                ++cur[q];
                for (int32_t i = q + 1; i < k; ++i) cur[i] = cur[q];
            }
        }
        if (out.size() > (1u << 21)) {
            if (std::fwrite(out.data(), 1, out.size(), f) != out.size()) rc = CR_ERR_IO;
            out.clear();
        }
    }
    out.push_back(']');
    if (rc == 0 && std::fwrite(out.data(), 1, out.size(), f) != out.size()) rc = CR_ERR_IO;
    if (std::fclose(f) != 0 && rc == 0) rc = CR_ERR_IO;
    return rc;
}

int write_game_json(const std::string& path, int32_t k, const std::vector<GameStep>& steps) {
    static const char* turns[] = {"Cop's Turn", "Robber's Turn", "Game Over - Captured!"};
    std::string out;
    if (steps.empty()) {
        out = "[]";
    } else {
        out = "[\n";
        char buf[64];
        for (size_t i = 0; i < steps.size(); ++i) {
            const GameStep& s = steps[i];
            out += "    {\n        \"cops\": [\n";
            for (int32_t j = 0; j < k; ++j) {
                std::snprintf(buf, sizeof buf, "            %d%s\n", s.cops[j], j + 1 < k ? "," : "");
                out += buf;
            }
            std::snprintf(buf, sizeof buf, "        ],\n        \"robber\": %d,\n", s.robber);
            out += buf;
            out += "        \"turn\": \"";
            out += turns[s.turn >= 0 && s.turn <= 2 ? s.turn : 0];
            out += i + 1 < steps.size() ? "\"\n    },\n" : "\"\n    }\n";
        }
        out += "]";
    }
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return CR_ERR_IO;
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0 && ok) ? 0 : CR_ERR_IO;
}

}  // namespace cr
