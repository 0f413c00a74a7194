// plugin_solver.cpp -- the reference's six-step solver plug-in, backed by
// libcrsolve.so.
//
// Drop-in for solvers/k_cops_visibility.cpp (CR_PLUGIN_VARIANT == 1) or
// solvers/k_cops_visibility_v2.cpp (CR_PLUGIN_VARIANT == 2): same six free
// functions (app/main.cpp:11-16), same bool convention (true = failed step),
// same stdout lines and the same cache blob path and layout
// (CacheManager.h:16-75). The solve itself runs on the GPU through the C ABI.
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "crsolve.h"

#ifndef CR_PLUGIN_VARIANT
#define CR_PLUGIN_VARIANT 1
#endif

bool loadGraphFile(const char* filename, int k, int p);
bool buildAuxGraph();
bool initializeCaptures();
bool mainLoop();
bool findFinalResult();
bool outputData();

namespace {

const char* g_filename = nullptr;
int g_k = 0;
int g_p = 0;
cr_solver* g_solver = nullptr;

int columns() { return CR_PLUGIN_VARIANT == CR_VARIANT_MACRO ? 1 : 2 * g_p; }

void report_error(const char* step) {
    std::cerr << "Error (" << step << "): " << cr_last_error() << "\n";
}

std::string cache_path() {
    char buf[4096];
    if (cr_cache_key(g_filename, g_k, g_p, CR_PLUGIN_VARIANT, buf, sizeof buf) < 0) return "";
    return buf;
}

// Allocator::print (lib/src/Allocator.cpp:69-209) equivalent: the line the
// regression harness reads is "Total Footprint ---=> <bytes> B".
void print_memory(uint64_t device_bytes) {
    const double mb = device_bytes / (1024.0 * 1024.0), gb = mb / 1024.0;
    std::cout << "\n||>>>>>=====-----=====<<<<<     Memory Tracking Report     >>>>>=====-----=====<<<<<\n";
    std::cout << "||\n";
    std::cout << "||   Total Footprint -------------------------=> " << std::right << std::setw(12) << device_bytes
              << " B / " << std::fixed << std::setprecision(2) << std::setw(7) << mb << " MB / " << std::setw(5)
              << gb << " GB (100.00%)\n";
    std::cout << "||    -> Device (HBM) ---------------------=> " << std::setw(12) << device_bytes << " B / "
              << std::setw(7) << mb << " MB / " << std::setw(5) << gb << " GB (100.00%)\n";
    std::cout << "||>>>>>>>>>>>>>>>>================------------------================<<<<<<<<<<<<<<<<\n\n";
}

}  // namespace

// loadGraphFile, solvers/k_cops_visibility.cpp:38-57
bool loadGraphFile(const char* filename_param, int k_param, int p_param) {
    g_filename = filename_param;
    g_k = k_param;
    g_p = p_param;
    if (!g_solver) {
        const int rc = cr_solver_create(0, &g_solver);
        if (rc != CR_OK) {
            report_error("device");
            return true;
        }
    }
    const int rc = cr_load_graph_file(g_solver, filename_param, k_param, p_param);
    if (rc == CR_ERR_EMPTY_GRAPH) {
        std::cerr << "Error: Graph is empty or failed to load.\n";
        return true;
    }
    if (rc != CR_OK) {
        report_error("loadGraphFile");
        return true;
    }
    if (cr_set_variant(g_solver, CR_PLUGIN_VARIANT) != CR_OK) {
        report_error("loadGraphFile");
        return true;
    }
    return false;
}

// buildAuxGraph, k_cops_visibility.cpp:59-83 (v2 :170-198). A readable blob
// under the cache key is loaded (CacheManager::loadAuxGraph, CacheManager.h:
// 77-141) and the build skipped; a missing blob, or one solved under other
// self-edge rules (the reference's stderr line), is a miss.
bool buildAuxGraph() {
    const std::string path = cache_path();
    std::error_code ec;
    if (!path.empty() && std::filesystem::is_regular_file(path, ec)) {
        const int rc = cr_load_aux_graph(g_solver, path.c_str());
        if (rc == CR_OK) {
            std::cout << "Loaded AuxGraph from cache!\n";
            return false;
        }
        if (rc == CR_ERR_CACHE_RULES)
            std::cerr << cr_last_error() << "\n";
        else  // the reference would copy a malformed blob as it is; here it is rebuilt
            std::cerr << "Ignoring cache blob: " << cr_last_error() << "\n";
    }
    if (CR_PLUGIN_VARIANT == CR_VARIANT_MACRO)
        std::cout << "Cache miss. Creating new AuxGraph (single-plane state space)...\n";
    else
        std::cout << "Cache miss. Creating new AuxGraph with " << columns() << " columns...\n";
    const int rc = cr_build_aux_graph(g_solver);
    if (rc != CR_OK) {
        if (rc == CR_ERR_K_RANGE) std::cerr << "FATAL: Number of cops (k) is outside the supported range.\n";
        report_error("buildAuxGraph");
        std::cerr << "Error: Unable to generate aux graph.\n";
        return true;
    }
    cr_info info{};
    cr_get_info(g_solver, &info);
    uint64_t total = 0;
    std::cout << "Building AuxGraph transition table for " << info.config_count << " configurations...\n";
    if (cr_count_transitions(g_solver, &total) != CR_OK) {
        report_error("buildAuxGraph");
        return true;
    }
    std::cout << "Transitions generated. Total edge pointers: " << total << "\n";
    return false;
}

// initializeCaptures, k_cops_visibility.cpp:85-127
bool initializeCaptures() {
    uint64_t base = 0;
    if (cr_initialize_captures(g_solver, &base) != CR_OK) {
        report_error("initializeCaptures");
        return true;
    }
    std::cout << "Initialized " << base << " base captures.\n";
    return false;
}

// mainLoop, k_cops_visibility.cpp:129-299 (v2 :343-565)
bool mainLoop() {
    if (CR_PLUGIN_VARIANT == CR_VARIANT_MACRO)
        std::cout << "Starting Macro-Turn Trajectory Search for " << g_p << "-Visibility...\n";
    else
        std::cout << "Starting Backward Induction Loop for " << g_p << "-Visibility...\n";
    std::vector<uint64_t> passes(1 << 16);
    int32_t n = 0;
    if (cr_main_loop(g_solver, passes.data(), static_cast<int32_t>(passes.size()), &n) != CR_OK) {
        report_error("mainLoop");
        return true;
    }
    for (int32_t i = 0; i < n && i < static_cast<int32_t>(passes.size()); ++i)
        std::cout << "Pass " << (i + 1) << ": Found " << passes[i] << " new winning states.\n";
    return false;
}

// findFinalResult, k_cops_visibility.cpp:301-354 (v2 :567-619)
bool findFinalResult() {
    std::cout << "\n--- FINAL VERDICT ---\n";
    cr_verdict v{};
    if (cr_find_final_result(g_solver, &v) != CR_OK) {
        report_error("findFinalResult");
        return true;
    }
    if (v.win) {
        std::cout << "RESULT: WIN. " << g_k << " Cop(s) CAN win this graph with 1/" << g_p << " visibility.\n";
        std::cout << "Optimal Cop Start Positions: (";
        for (int i = 0; i < g_k; ++i) std::cout << static_cast<int>(v.positions[i]) << (i == g_k - 1 ? "" : ", ");
        std::cout << ")\n";
        if (CR_PLUGIN_VARIANT == CR_VARIANT_MACRO)
            std::cout << "Capture Time: " << v.plys << " plys.\n";
        else
            std::cout << "Capture Time: " << v.plys << " plys (" << (v.plys / columns()) << " full cycles).\n";
    } else {
        std::cout << "RESULT: LOSS. " << g_k << " Cop(s) CANNOT guarantee a win.\n";
        std::cout << "(The Robber has a strategy to survive indefinitely against any start).\n";
    }
    cr_info info{};
    cr_get_info(g_solver, &info);
    print_memory(info.device_bytes);
    return false;
}

// outputData -> CacheManager::saveAuxGraph, k_cops_visibility.cpp:356-373
bool outputData() {
    std::cout << "Saving filled AuxGraph to cache... ";
    std::error_code ec;
    std::filesystem::create_directories("cache", ec);
    const std::string path = cache_path();
    if (path.empty() || cr_save_aux_graph(g_solver, path.c_str()) != CR_OK) {
        std::cout << "Failed!\n";
        return true;
    }
    std::cout << "Success!\n";
    return false;
}

// Releases the device solver (called by the supervisor at exit).
void releaseSolver() {
    cr_solver_destroy(g_solver);
    g_solver = nullptr;
}
