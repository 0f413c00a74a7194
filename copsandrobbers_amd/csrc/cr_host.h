// cr_host.h -- host-side pieces of the solver: matrix parsing, configuration
// ranking, team-move lists and the cache-blob writer. Plain C++17, no HIP.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace cr {

// Graph::Graph(const char*) parsing, lib/src/Graph.cpp:26-70.
int32_t parse_matrix(const uint8_t* text, uint64_t len, uint8_t* adj, int32_t cap);

// Out- and in-neighbour lists in ascending vertex order (AdjacencyList.cpp:33-72
// keeps row order; the in-lists serve predecessor enumeration).
struct Adjacency {
    int32_t n = 0;
    std::vector<uint8_t> matrix;           // n*n
    std::vector<std::vector<uint8_t>> out; // moves out of u
    std::vector<std::vector<uint8_t>> in;  // vertices that move into v
    int32_t max_out = 0, max_in = 0;
    void build(int32_t n_, const uint8_t* adj);
};

// C(a, b) in 64 bits (0 when b < 0 or a < b).
uint64_t binom(int64_t a, int64_t b);
// AuxGraph.h:124-138: C(N+k-1, k).
uint64_t config_count(int32_t n, int32_t k);
// Lexicographic rank of a non-decreasing k-tuple (AuxGraph.h:155-176 order).
uint64_t rank_tuple(const uint8_t* sorted, int32_t n, int32_t k);
// Inverse of rank_tuple.
void unrank_tuple(uint64_t rank, int32_t n, int32_t k, uint8_t* out);

// Sorted, de-duplicated successor ranks of one configuration over the given
// neighbour lists (AuxGraph.h:194-254). stay adds the cop's own vertex.
void team_moves(const std::vector<std::vector<uint8_t>>& nbr, const uint8_t* cops, int32_t k,
                int32_t n, bool stay, std::vector<uint64_t>& out);

// CacheManager::buildCacheKey, CacheManager.h:16-26 (without mkdir).
std::string cache_key(const std::string& algo, const std::string& graph_path, int32_t k,
                      int32_t p, bool cop_stay, bool robber_stay);

// CacheManager::saveAuxGraph, CacheManager.h:28-75, streamed: header, configs,
// heads, transitions (nextId*N) then `states` fetched in chunks by the
// callback (offset, count, dst). Returns 0 or a negative CR_ERR_* status.
int write_aux_blob(const std::string& path, const Adjacency& g, int32_t k, int32_t columns,
                   bool cop_stay, bool robber_stay, uint64_t num_states,
                   const std::function<int(uint64_t, uint64_t, uint8_t*)>& fetch_states);

// The 96-byte header of a cache blob (CacheManager.h:49-58) and the file's
// size, with the section offsets checked against each other and the size.
// Returns 0, CR_ERR_IO or CR_ERR_FORMAT (why says what failed).
int read_aux_header(const std::string& path, uint64_t header[12], uint64_t* file_size, std::string& why);

// CacheManager::loadAuxGraph, CacheManager.h:77-141, for a graph already
// loaded: the rules flags first (CR_ERR_CACHE_RULES, :89-95), then every
// header field against what write_aux_blob writes for (g, k, columns, rules,
// num_states), the configs section against the ranking, heads[0] and heads[C]
// against the transitions section (CR_ERR_FORMAT). The states section goes to
// put_states(offset, count, src) in chunks; the transitions are not read.
int read_aux_blob(const std::string& path, const Adjacency& g, int32_t k, int32_t columns, bool cop_stay,
                  bool robber_stay, uint64_t num_states, std::string& why,
                  const std::function<int(uint64_t, uint64_t, const uint8_t*)>& put_states);

// legacy/util/edge_parser.py:4-63: 1-based "u v" / "u, v" lines, '#'
// comments, undirected; N = the largest id (an id 0 lands on row N-1, as
// Python's negative index does there). Returns N, 0 when there is no edge.
int32_t parse_edge_list(const uint8_t* text, uint64_t len, uint8_t* adj, int32_t cap);

// The dp_tables JSON of the legacy round solver (k_cops_rounds.cpp:211-281,
// k_cops.py:86-131) from the column-0 bytes of a solved table: one object per
// (cId, r), json.dumps spacing. fetch_states(offset, count, dst) reads table
// bytes. Returns 0 or a negative CR_ERR_* status.
int write_dp_table(const std::string& path, int32_t n, int32_t k, int32_t columns, uint64_t configs,
                   const std::function<int(uint64_t, uint64_t, uint8_t*)>& fetch_states);

// One step of an extracted game: cops, robber, turn (0 cop, 1 robber, 2 captured).
struct GameStep {
    uint8_t cops[16];
    int32_t robber;
    int32_t turn;
};
// json.dump(game, indent=4) of k_cops.py:243-260 / export_helper.py:47-49.
int write_game_json(const std::string& path, int32_t k, const std::vector<GameStep>& steps);

}  // namespace cr
