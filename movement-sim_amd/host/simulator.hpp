// simulator.hpp — the reference's host driver sim::Simulator
// (src/sim/Simulator.hpp:46-130, Simulator.cpp:44-400) on the HIP pipeline.
//
// Same public surface and behaviour: init() loads the map and adds the
// agents, start_worker()/stop_worker() run one worker thread,
// continue_simulation()/pause_simulation() gate it, each tick is a move
// dispatch and a collision dispatch each timed on its own (the reference's
// evals block), every tick appends "<time>;<tick/2>;<s update>;<s collision>;<s all>"
// to "<agents>.csv" and stderr, get_entities() hands the last read-back
// entity buffer over (the worker reads a fresh one after a hand-over), and
// the tick-rate / tick-duration statistics are kept alike.  Differences:
// the agents are seeded (the reference draws them from random_device), the
// map path, agent count and radius are parameters, the hand-over and the
// run flags are guarded by a mutex, get_grid_cells() replaces
// get_quad_tree_nodes() (per-agent cells of the uniform grid), and an
// optional tick limit pauses the worker after that many ticks.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <fstream>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "map.hpp"
#include "msim/msim.h"
#include "tick_stats.hpp"

namespace msim_host {

enum class SimulatorState { STOPPED, RUNNING, JOINING };

struct SimulatorConfig {
    std::string map_path;
    uint32_t agents = 1000000;      // MAX_ENTITIES
    float collision_radius = 10.0f;  // COLLISION_RADIUS
    uint64_t seed = 1;
    std::string csv_path;            // "" -> "<agents>.csv"
    bool log_to_stderr = true;
    uint64_t tick_limit = 0;         // 0: run until paused or stopped
};

// The reference's add_entities (Simulator.cpp:114-129), seeded: a uniform
// road, pos = its start, target = its end, direction 0, a random colour
// (alpha 1), a random non-zero xorshift state, initialized = 0.
std::vector<msim_entity> make_entities(const Map& m, uint32_t n, uint64_t seed);

// "HH:MM:SS.mmm" of a time of day, as the reference's get_time_stamp()
// prints it (its minutes are not zero-padded: Simulator.cpp's pad loop for
// them tests the milliseconds string).
std::string format_time_of_day(std::chrono::nanoseconds since_midnight);

class Simulator {
 public:
    explicit Simulator(SimulatorConfig cfg);
    ~Simulator();
    Simulator(const Simulator&) = delete;
    Simulator& operator=(const Simulator&) = delete;

    void init();
    bool is_initialized() const { return initialized_; }
    SimulatorState get_state() const { return state_.load(); }
    void start_worker();
    void stop_worker();

    void continue_simulation();
    void pause_simulation();
    bool is_simulating() const { return simulating_.load(); }
    // The statistics are updated by the worker thread; the reference-style
    // getters below are for a stopped or paused simulator, stats() reads
    // them under the worker's lock at any time.
    struct StatsSnapshot {
        double tps;
        std::chrono::nanoseconds tick, update, collision;
    };
    StatsSnapshot stats() const;
    const TickRate& get_tps() const { return tps_; }
    const TickDurationHistory& get_tps_history() const { return tps_history_; }
    const TickDurationHistory& get_update_tick_history() const { return update_history_; }
    const TickDurationHistory& get_collision_detection_tick_history() const { return collision_history_; }
    std::shared_ptr<std::vector<msim_entity>> get_entities();
    std::shared_ptr<std::vector<int32_t>> get_grid_cells();
    std::shared_ptr<const Map> get_map() const { return map_; }
    uint64_t ticks_done() const { return ticks_.load(); }
    // Waits until ticks_done() >= n, the simulation pauses or a tick fails
    // (returns false only on a failure or when the timeout passes first;
    // *timed_out tells the two apart).
    bool wait_ticks(uint64_t n, std::chrono::milliseconds timeout, bool* timed_out);
    bool failed() const { return failed_.load(); }
    std::string last_error() const;
    msim_sim* handle() { return sim_; }

 private:
    void sim_worker();
    bool sim_tick();
    void write_log_csv(uint32_t tick, std::chrono::nanoseconds upd, std::chrono::nanoseconds col,
                       std::chrono::nanoseconds all);
    void set_error(const std::string& e);

    SimulatorConfig cfg_;
    bool initialized_ = false;
    std::ofstream log_;
    std::unique_ptr<std::thread> thread_;
    std::atomic<SimulatorState> state_{SimulatorState::STOPPED};
    std::mutex wait_mutex_;
    std::condition_variable wait_cv_;
    std::atomic<bool> simulating_{false};
    std::atomic<uint64_t> ticks_{0};
    std::atomic<bool> failed_{false};
    std::condition_variable progress_cv_;  // signalled (under wait_mutex_) per tick and on pause

    mutable std::mutex stats_mutex_;  // guards the histories and the tick rate
    TickDurationHistory tps_history_, update_history_, collision_history_;
    TickRate tps_;

    std::shared_ptr<Map> map_;
    msim_sim* sim_ = nullptr;
    msim_push_consts pc_{};

    mutable std::mutex handoff_mutex_;
    std::shared_ptr<std::vector<msim_entity>> entities_;
    std::shared_ptr<std::vector<int32_t>> cells_;
    std::string error_;
};

}  // namespace msim_host
