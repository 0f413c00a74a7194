// cr_game.h -- interface of the perfect-game walk in game.hip to the C ABI in
// crsolve.hip. Internal header.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "cr_device.h"
#include "cr_host.h"

namespace crd {

// One step as the walk kernel writes it.
struct DevGameStep {
    uint8_t cops[16];
    int32_t robber;
    int32_t turn;  // 0 cop's turn, 1 robber's turn, 2 captured
};

// Extracts the perfect game from a solved table (column 0 holds the cop-turn
// states, `columns` per (cId, r)); matrix is the n*n input graph, the stay
// flags select the options. cpp_order selects k_cops_rounds.cpp's choices
// (cop moves by ascending next cId, robber options without staying) instead
// of k_cops.py's. out is empty when no configuration wins.
int game_extract(const GraphView& gv, int k, int columns, const uint8_t* d_table, const std::vector<uint8_t>& matrix,
                 bool cop_stay, bool robber_stay, bool cpp_order, hipStream_t st, std::vector<cr::GameStep>& out);

}  // namespace crd
