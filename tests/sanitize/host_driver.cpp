// Exercises the solver's host code (copsandrobbers_amd/csrc/cr_host.cpp)
// under sanitizers: matrix and edge-list parsing, neighbour lists, ranks,
// team moves, and the multithreaded cache-blob, DP-table and game writers.
// Usage: host_driver MATRIX K OUTDIR
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "cr_host.h"

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    std::ifstream f(argv[1], std::ios::binary);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    const int k = std::atoi(argv[2]);
    const std::string dir = argv[3];
    const auto* t = reinterpret_cast<const uint8_t*>(text.data());
    const int n = cr::parse_matrix(t, text.size(), nullptr, 0);
    std::vector<uint8_t> adj(static_cast<size_t>(n) * n);
    cr::parse_matrix(t, text.size(), adj.data(), n);
    cr::Adjacency g;
    g.build(n, adj.data());
    // edge list of the same graph, 1-based, parsed back
    std::string edges = "# edges\n";
    for (int u = 0; u < n; ++u)
        for (int v = u + 1; v < n; ++v)
            if (adj[u * n + v]) edges += std::to_string(u + 1) + ", " + std::to_string(v + 1) + "\n";
    std::vector<uint8_t> adj2(static_cast<size_t>(n) * n);
    const int n2 = cr::parse_edge_list(reinterpret_cast<const uint8_t*>(edges.data()), edges.size(), adj2.data(), n);
    if (n2 != n || adj2 != adj) { std::fprintf(stderr, "edge list differs\n"); return 1; }
    const uint64_t C = cr::config_count(n, k);
    std::vector<uint8_t> cops(k);
    std::vector<uint64_t> moves;
    for (uint64_t r = 0; r < C; ++r) {
        cr::unrank_tuple(r, n, k, cops.data());
        if (cr::rank_tuple(cops.data(), n, k) != r) { std::fprintf(stderr, "rank\n"); return 1; }
        cr::team_moves(g.out, cops.data(), k, n, false, moves);
    }
    const uint64_t S = C * static_cast<uint64_t>(n) * 2;
    auto fetch = [&](uint64_t off, uint64_t cnt, uint8_t* dst) {
        for (uint64_t i = 0; i < cnt; ++i) dst[i] = static_cast<uint8_t>((off + i) * 2654435761u >> 24) | 1u;
        return 0;
    };
    if (cr::write_aux_blob(dir + "/blob.bin", g, k, 2, false, false, S, fetch)) { std::fprintf(stderr, "blob\n"); return 1; }
    if (cr::write_dp_table(dir + "/dp.json", n, k, 2, C, fetch)) { std::fprintf(stderr, "dp\n"); return 1; }
    std::vector<cr::GameStep> game(3);
    for (int i = 0; i < 3; ++i) {
        std::memset(&game[i], 0, sizeof game[i]);
        game[i].robber = i;
        game[i].turn = i % 3;
    }
    if (cr::write_game_json(dir + "/game.json", k, game)) { std::fprintf(stderr, "game\n"); return 1; }
    std::printf("host driver ok: n=%d k=%d C=%llu key=%s\n", n, k, static_cast<unsigned long long>(C),
                cr::cache_key("k_cops_visibility", argv[1], k, 1, false, false).c_str());
    return 0;
}
