// cr_columns.h -- interface of the column solver for p >= 2 in columns.hip
// to the C ABI in crsolve.hip. Internal header.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "cr_macro.h"

namespace crd {

struct ColumnsPlan;  // device planes of one (graph, k, p)

// Allocates the planes and seeds the captures (k_cops_visibility.cpp:85-127:
// columns 0 and 1 only); table holds 2p*C*n bytes. reach[r*4 + w] is the set
// of vertices the robber reaches from r in exactly p moves.
int columns_init(ColumnsPlan** plan, const GraphView& gv, const std::vector<uint64_t>& reach, int k, int p,
                 hipStream_t st, uint8_t* table, uint64_t& device_bytes, uint64_t* base);

// Main loop (k_cops_visibility.cpp:129-299 with 2p columns): depth levels
// until one wins nothing; passes receives the reference's per-pass counts.
int columns_solve(ColumnsPlan* plan, TimeSink sink, std::vector<uint64_t>& passes, uint64_t* won_states,
                  int32_t* levels);

// Won plane of column 0 (C rows of W words) for the verdict.
uint64_t* columns_won0(ColumnsPlan* plan);

void columns_free(ColumnsPlan* plan, uint64_t& device_bytes);

}  // namespace crd
