// cli_main.cpp -- supervisor and CLI with the behaviour of app/main.cpp:
// `<exe> <graph_file.txt> <num_cops> [visibility_p]`, the six plug-in steps in
// order under a wall-clock section timer (lib/src/Profiler.cpp report format),
// "[FAILED STEP]" on a failed step, exit status 0 after a run (main.cpp:107-109)
// and 1 on a usage error.
#include <chrono>
#include <iomanip>
#include <iostream>
#include <map>
#include <string>
#include <vector>

bool loadGraphFile(const char* filename, int k, int p);
bool buildAuxGraph();
bool initializeCaptures();
bool mainLoop();
bool findFinalResult();
bool outputData();
void releaseSolver();

namespace {

class SectionTimer {
  public:
    using Clock = std::chrono::steady_clock;
    SectionTimer() : start_(Clock::now()), mark_(start_) {}

    void enter(const std::string& name) {
        const auto now = Clock::now();
        if (!current_.empty()) totals_[current_] += std::chrono::duration<double>(now - mark_).count();
        if (totals_.find(name) == totals_.end()) {
            order_.push_back(name);
            totals_[name] = 0.0;
        }
        current_ = name;
        mark_ = now;
    }

    void print() {
        enter("");
        const double total = std::chrono::duration<double>(Clock::now() - start_).count();
        double tracked = 0;
        for (const auto& name : order_) tracked += totals_[name];
        size_t width = 0;
        for (const auto& name : order_) width = std::max(width, ("  -> " + name + " ").size());
        auto line = [&](const std::string& label, size_t align, double seconds) {
            const size_t dashes = align > label.size() + 2 ? align - label.size() - 2 : 1;
            std::cout << "||" << label << std::string(dashes, '-') << "=> " << std::right << std::fixed
                      << std::setprecision(4) << std::setw(12) << seconds << " s (" << std::setprecision(2)
                      << std::setw(6) << (total > 0 ? 100.0 * seconds / total : 0.0) << "%)\n";
        };
        std::cout << "\n||>>>>>=====-----=====<<<<<     Timing Profiler Report     >>>>>=====-----=====<<<<<\n||\n";
        line("   Total App Uptime ", 37, total);
        line("    -> Tracked Execution ", 32, tracked);
        std::cout << "||\n||\n||  ---===<<<>>>===---   Drill Down   ---===<<<>>>===---\n||\n";
        for (const auto& name : order_)
            if (!name.empty()) line("  -> " + name + " ", width + 3, totals_[name]);
        std::cout << "||\n||>>>>>>>>>>>>>>>>================------------------================<<<<<<<<<<<<<<<<\n\n";
    }

  private:
    Clock::time_point start_, mark_;
    std::string current_;
    std::map<std::string, double> totals_;
    std::vector<std::string> order_;
};

void solveCopsAndRobbers(const char* filename, int k, int p) {
    SectionTimer prof;
    struct Step {
        const char* name;
        bool (*fn)();
    };
    prof.enter("Load Graph File");
    if (loadGraphFile(filename, k, p)) {
        std::cout << "[FAILED STEP]: 'Load Graph File' was unsuccessful!\n";
        return;
    }
    prof.enter("Idle");
    const Step steps[] = {{"Build Aux Graph", buildAuxGraph},
                          {"Initialize Captures", initializeCaptures},
                          {"Main Loop", mainLoop},
                          {"Find Final Result", findFinalResult},
                          {"Output Data", outputData}};
    for (const Step& st : steps) {
        prof.enter(st.name);
        if (st.fn()) {
            std::cout << "[FAILED STEP]: '" << st.name << "' was unsuccessful!\n";
            return;
        }
        prof.enter("Idle");
    }
    prof.print();
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc < 3 || argc > 4) {
        std::cout << "Usage: " << argv[0] << " <graph_file.txt> <num_cops> [visibility_p]\n";
        std::cout << "Example (1/2 Vis): " << argv[0] << " graph3.txt 4 2\n";
        std::cout << "Example (Standard): " << argv[0] << " graph3.txt 4\n";
        return 1;
    }
    int k = 0, p = 1;
    try {
        k = std::stoi(argv[2]);
        p = argc == 4 ? std::stoi(argv[3]) : 1;
    } catch (const std::exception&) {
        std::cout << "Usage: " << argv[0] << " <graph_file.txt> <num_cops> [visibility_p]\n";
        return 1;
    }
    solveCopsAndRobbers(argv[1], k, p);
    releaseSolver();
    return 0;
}
