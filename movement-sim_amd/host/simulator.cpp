// simulator.cpp — see simulator.hpp.
#include "simulator.hpp"

#include <cstdio>
#include <iostream>
#include <random>
#include <sstream>
#include <stdexcept>

namespace msim_host {

namespace {

uint32_t uniform_below(std::mt19937_64& g, uint32_t n) {
    return static_cast<uint32_t>((static_cast<unsigned __int128>(g()) * n) >> 64);
}

float unit_float(std::mt19937_64& g) { return static_cast<float>(g() >> 40) * (1.0f / 16777216.0f); }

// quad-tree constants of the reference's push constants (Simulator.cpp:94-101);
// the grid does not use them, they travel for layout compatibility
constexpr uint32_t kQuadTreeNodes = 21845, kQuadTreeDepth = 8, kQuadTreeCap = 10;

}  // namespace

std::vector<msim_entity> make_entities(const Map& m, uint32_t n, uint64_t seed) {
    if (m.roads.empty() && n) throw std::runtime_error("the map has no roads to place agents on");
    std::mt19937_64 g(seed);
    std::vector<msim_entity> out(n);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t r = uniform_below(g, static_cast<uint32_t>(m.roads.size()));
        msim_entity& e = out[i];
        e.color = msim_rgba{unit_float(g), unit_float(g), unit_float(g), 1.0f};
        do {
            const uint64_t a = g(), b = g();
            e.randomState = msim_vec4u{static_cast<uint32_t>(a), static_cast<uint32_t>(a >> 32),
                                       static_cast<uint32_t>(b), static_cast<uint32_t>(b >> 32)};
        } while ((e.randomState.x | e.randomState.y | e.randomState.z | e.randomState.w) == 0u);
        e.pos = m.roads[r].start.pos;
        e.target = m.roads[r].end.pos;
        e.direction = msim_vec2{0.0f, 0.0f};
        e.roadIndex = r;
        e.initialized = 0u;
    }
    return out;
}

std::string format_time_of_day(std::chrono::nanoseconds t) {
    const long long ms_total = std::chrono::duration_cast<std::chrono::milliseconds>(t).count();
    const long long h = ms_total / 3600000, m = (ms_total / 60000) % 60, s = (ms_total / 1000) % 60,
                    ms = ms_total % 1000;
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%02lld:%lld:%02lld.%03lld", h, m, s, ms);
    return buf;
}

Simulator::Simulator(SimulatorConfig cfg) : cfg_(std::move(cfg)) {
    if (cfg_.csv_path.empty()) cfg_.csv_path = std::to_string(cfg_.agents) + ".csv";
    log_.open(cfg_.csv_path, std::ios::out | std::ios::app);
    if (!log_.is_open()) throw std::runtime_error("cannot open the tick log " + cfg_.csv_path);
}

Simulator::~Simulator() {
    if (state_.load() == SimulatorState::RUNNING) stop_worker();
    if (sim_) msim_destroy(sim_);
}

void Simulator::init() {
    if (initialized_) throw std::logic_error("Simulator::init called twice");
    map_ = load_map(cfg_.map_path);
    if (!map_) throw std::runtime_error("failed to open map " + cfg_.map_path);
    add_guard_entry(*map_);
    const std::vector<msim_entity> ents = make_entities(*map_, cfg_.agents, cfg_.seed);
    pc_.worldSizeX = map_->width;
    pc_.worldSizeY = map_->height;
    pc_.nodeCount = kQuadTreeNodes;
    pc_.maxDepth = kQuadTreeDepth;
    pc_.entityNodeCap = kQuadTreeCap;
    pc_.collisionRadius = cfg_.collision_radius;
    pc_.tick = 1;
    if (msim_create(map_->roads.data(), static_cast<uint32_t>(map_->roads.size()), map_->connections.data(),
                    static_cast<uint32_t>(map_->connections.size()), ents.data(), cfg_.agents, nullptr, &pc_,
                    nullptr, &sim_) != MSIM_OK) {
        throw std::runtime_error(msim_get_last_error());
    }
    entities_ = std::make_shared<std::vector<msim_entity>>(ents);
    cells_ = std::make_shared<std::vector<int32_t>>(2ull * cfg_.agents, 0);
    initialized_ = true;
}

void Simulator::start_worker() {
    if (!initialized_ || state_.load() != SimulatorState::STOPPED || thread_) {
        throw std::logic_error("start_worker: not initialised or already running");
    }
    state_ = SimulatorState::RUNNING;
    thread_ = std::make_unique<std::thread>(&Simulator::sim_worker, this);
}

void Simulator::stop_worker() {
    if (!initialized_ || state_.load() != SimulatorState::RUNNING || !thread_) {
        throw std::logic_error("stop_worker: not running");
    }
    {
        std::lock_guard<std::mutex> lk(wait_mutex_);
        state_ = SimulatorState::JOINING;
    }
    wait_cv_.notify_all();
    progress_cv_.notify_all();
    if (thread_->joinable()) thread_->join();
    thread_ = nullptr;
    state_ = SimulatorState::STOPPED;
}

void Simulator::continue_simulation() {
    {
        std::lock_guard<std::mutex> lk(wait_mutex_);
        if (simulating_ || failed_) return;
        simulating_ = true;
    }
    wait_cv_.notify_all();
}

void Simulator::pause_simulation() {
    {
        std::lock_guard<std::mutex> lk(wait_mutex_);
        if (!simulating_) return;
        simulating_ = false;
    }
    wait_cv_.notify_all();
    progress_cv_.notify_all();
}

bool Simulator::wait_ticks(uint64_t n, std::chrono::milliseconds timeout, bool* timed_out) {
    std::unique_lock<std::mutex> lk(wait_mutex_);
    const bool done = progress_cv_.wait_for(lk, timeout, [&] {
        return ticks_.load() >= n || failed_.load() || !simulating_.load() ||
               state_.load() != SimulatorState::RUNNING;
    });
    if (timed_out) *timed_out = !done;
    return done && !failed_.load();
}

std::shared_ptr<std::vector<msim_entity>> Simulator::get_entities() {
    std::lock_guard<std::mutex> lk(handoff_mutex_);
    return std::move(entities_);
}

std::shared_ptr<std::vector<int32_t>> Simulator::get_grid_cells() {
    std::lock_guard<std::mutex> lk(handoff_mutex_);
    return std::move(cells_);
}

std::string Simulator::last_error() const {
    std::lock_guard<std::mutex> lk(handoff_mutex_);
    return error_;
}

void Simulator::set_error(const std::string& e) {
    std::lock_guard<std::mutex> lk(handoff_mutex_);
    error_ = e;
}

void Simulator::sim_worker() {
    std::unique_lock<std::mutex> lk(wait_mutex_);
    while (state_.load() == SimulatorState::RUNNING) {
        if (!simulating_) {
            wait_cv_.wait(lk, [this] { return simulating_ || state_.load() != SimulatorState::RUNNING; });
            continue;
        }
        lk.unlock();
        const bool ok = sim_tick();
        lk.lock();
        if (!ok) failed_ = true;
        if (!ok || (cfg_.tick_limit && ticks_.load() >= cfg_.tick_limit)) simulating_ = false;
        progress_cv_.notify_all();
    }
    progress_cv_.notify_all();
}

bool Simulator::sim_tick() {
    using clock = std::chrono::high_resolution_clock;
    const auto tick_start = clock::now();
    // move dispatch (even tick), then collision dispatch (odd tick), each
    // waited for like the reference's blocking evals
    pc_.tick++;
    auto t0 = clock::now();
    if (msim_dispatch(sim_, &pc_) != MSIM_OK || msim_sync(sim_) != MSIM_OK) {
        set_error(msim_last_error(sim_));
        return false;
    }
    const std::chrono::nanoseconds upd = clock::now() - t0;
    {
        std::lock_guard<std::mutex> lk(stats_mutex_);
        update_history_.add_time(upd);
    }
    pc_.tick++;
    t0 = clock::now();
    if (msim_dispatch(sim_, &pc_) != MSIM_OK || msim_sync(sim_) != MSIM_OK) {
        set_error(msim_last_error(sim_));
        return false;
    }
    const std::chrono::nanoseconds col = clock::now() - t0;
    {
        std::lock_guard<std::mutex> lk(stats_mutex_);
        collision_history_.add_time(col);
    }
    write_log_csv(pc_.tick, upd, col, upd + col);

    bool want_entities, want_cells;
    {
        std::lock_guard<std::mutex> lk(handoff_mutex_);
        want_entities = !entities_;
        want_cells = !cells_;
    }
    if (want_entities) {
        auto fresh = std::make_shared<std::vector<msim_entity>>(cfg_.agents);
        if (msim_read_entities(sim_, fresh->data(), 0, cfg_.agents) != MSIM_OK) {
            set_error(msim_last_error(sim_));
            return false;
        }
        std::lock_guard<std::mutex> lk(handoff_mutex_);
        entities_ = std::move(fresh);
    }
    if (want_cells) {
        auto fresh = std::make_shared<std::vector<int32_t>>(2ull * cfg_.agents);
        if (msim_read_cells(sim_, fresh->data()) != MSIM_OK) {
            set_error(msim_last_error(sim_));
            return false;
        }
        std::lock_guard<std::mutex> lk(handoff_mutex_);
        cells_ = std::move(fresh);
    }
    uint32_t debug[10];
    if (msim_read_debug(sim_, debug) != MSIM_OK) {
        set_error(msim_last_error(sim_));
        return false;
    }
    {
        std::lock_guard<std::mutex> lk(stats_mutex_);
        tps_history_.add_time(clock::now() - tick_start);
        tps_.tick();
    }
    ticks_++;
    return true;
}

Simulator::StatsSnapshot Simulator::stats() const {
    std::lock_guard<std::mutex> lk(stats_mutex_);
    return StatsSnapshot{tps_.get_ticks(), tps_history_.get_avg_time(), update_history_.get_avg_time(),
                         collision_history_.get_avg_time()};
}

void Simulator::write_log_csv(uint32_t tick, std::chrono::nanoseconds upd, std::chrono::nanoseconds col,
                              std::chrono::nanoseconds all) {
    using std::chrono::duration_cast;
    using std::chrono::milliseconds;
    const auto now = std::chrono::system_clock::now().time_since_epoch();
    const auto day = std::chrono::hours(24);
    const std::chrono::nanoseconds tod = std::chrono::duration_cast<std::chrono::nanoseconds>(now % day);
    std::ostringstream line;
    line << format_time_of_day(tod) << ";" << (tick / 2) << ";"
         << static_cast<double>(duration_cast<milliseconds>(upd).count()) / 1000 << ";"
         << static_cast<double>(duration_cast<milliseconds>(col).count()) / 1000 << ";"
         << static_cast<double>(duration_cast<milliseconds>(all).count()) / 1000 << "\n";
    log_ << line.str();
    log_.flush();
    if (cfg_.log_to_stderr) std::cerr << line.str();
}

}  // namespace msim_host
