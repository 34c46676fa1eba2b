// headless_main.cpp — runs the Simulator without a window: the reference's
// UI (src/ui/) starts the worker and toggles continue/pause; here the
// command line does.
//
//   msim_headless --map FILE [--agents N] [--radius R] [--seed S]
//                 [--ticks T] [--csv FILE] [--quiet]
//
// Prints the tick rate and the mean tick, move and collision durations
// when the run ends.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>

#include "simulator.hpp"

namespace {

int usage(const char* argv0) {
    std::fprintf(stderr,
                 "usage: %s --map FILE [--agents N] [--radius R] [--seed S] [--ticks T] [--csv FILE] [--quiet]\n",
                 argv0);
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    msim_host::SimulatorConfig cfg;
    cfg.tick_limit = 100;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        const bool has_value = i + 1 < argc;
        if (a == "--quiet") {
            cfg.log_to_stderr = false;
        } else if (a == "--map" && has_value) {
            cfg.map_path = argv[++i];
        } else if (a == "--agents" && has_value) {
            cfg.agents = static_cast<uint32_t>(std::strtoul(argv[++i], nullptr, 10));
        } else if (a == "--radius" && has_value) {
            cfg.collision_radius = std::strtof(argv[++i], nullptr);
        } else if (a == "--seed" && has_value) {
            cfg.seed = std::strtoull(argv[++i], nullptr, 10);
        } else if (a == "--ticks" && has_value) {
            cfg.tick_limit = std::strtoull(argv[++i], nullptr, 10);
        } else if (a == "--csv" && has_value) {
            cfg.csv_path = argv[++i];
        } else {
            return usage(argv[0]);
        }
    }
    if (cfg.map_path.empty() || cfg.tick_limit == 0) return usage(argv[0]);
    try {
        msim_host::Simulator sim(cfg);
        sim.init();
        sim.start_worker();
        sim.continue_simulation();
        bool timed_out = false;
        const bool ok = sim.wait_ticks(cfg.tick_limit, std::chrono::hours(24), &timed_out);
        sim.pause_simulation();
        sim.stop_worker();
        if (!ok) {
            std::fprintf(stderr, "tick failed: %s\n", sim.last_error().c_str());
            return 1;
        }
        std::printf("ticks %llu  tps %.2f  tick %s  update %s  collision %s\n",
                    static_cast<unsigned long long>(sim.ticks_done()), sim.get_tps().get_ticks(),
                    sim.get_tps_history().get_avg_time_str().c_str(),
                    sim.get_update_tick_history().get_avg_time_str().c_str(),
                    sim.get_collision_detection_tick_history().get_avg_time_str().c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
