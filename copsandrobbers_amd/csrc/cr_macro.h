// cr_macro.h -- interface of the limited-visibility (macro-turn) solver in
// macro.hip to the C ABI in crsolve.hip. Internal header.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "cr_device.h"

namespace crd {

// Per-launch kernel time callback (name, milliseconds); null when not timing.
struct TimeSink {
    void* ctx = nullptr;
    void (*add)(void* ctx, const char* name, double ms) = nullptr;
};

// Shape and cost of one solve: nodes of the trajectory DAG per cop step
// d = 1..p, candidate successor pairs generated while building it, and per
// pass the (node, team move) pairs evaluated at the last step plus the DAG
// edges read on the way up.
struct MacroProfile {
    std::vector<uint64_t> nodes;       // [p]
    std::vector<uint64_t> candidates;  // [p-1]: pairs generated from step d (before dedup)
    std::vector<uint64_t> edges;       // [p-1]: DAG edges out of step d
    uint64_t leaf_moves = 0;           // summed over passes
    uint64_t up_edges = 0;             // summed over passes
    uint64_t build_moves = 0;          // team moves enumerated while building
    double build_ms = 0, passes_ms = 0;
};

struct MacroPlan;  // device buffers of the DAG, kept between solves

// Captures (k_cops_visibility_v2.cpp:200-223): table[c*n + r] = r in C ? depth 0
// : unmarked; occ[c*W + q] = cop mask of configuration c. Returns the count.
int macro_init(const GraphView& gv, int k, hipStream_t st, uint8_t* table, uint64_t* occ, uint64_t* base);

// Main loop (k_cops_visibility_v2.cpp:343-565): builds the DAG, runs snapshot
// passes until one marks nothing, and leaves the marked-state plane in
// `marked` (C*W words) for the verdict.
int macro_solve(MacroPlan** plan, const GraphView& gv, const uint64_t* d_nbr_mask, int k, int p,
                hipStream_t st, uint8_t* table, const uint64_t* occ, uint64_t* marked,
                uint64_t& device_bytes, TimeSink sink, std::vector<uint64_t>& passes, MacroProfile& prof);

void macro_free(MacroPlan* plan, uint64_t& device_bytes);

}  // namespace crd
