// host_capi.cpp — extern "C" entry points of include/msim/msim_host.h over
// the C++ host driver (map.hpp, tick_stats.hpp, simulator.hpp).  Exceptions
// never cross the ABI: each one becomes a status code and a per-thread
// message.
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "map.hpp"
#include "msim/msim_host.h"
#include "simulator.hpp"
#include "tick_stats.hpp"

struct msim_host_map {
    std::shared_ptr<msim_host::Map> map;
};
struct msim_tick_rate {
    msim_host::TickRate rate{msim_host::TickRate::clock::time_point{}};
};
struct msim_tick_history {
    msim_host::TickDurationHistory hist;
};
struct msim_host_sim {
    std::unique_ptr<msim_host::Simulator> sim;
};

namespace {

thread_local std::string g_host_error;

int fail(int code, const std::string& msg) {
    g_host_error = msg;
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(MSIM_ERR_OOM, "host allocation failed");
    } catch (const std::logic_error& e) {
        return fail(MSIM_ERR_STATE, e.what());
    } catch (const std::exception& e) {
        return fail(MSIM_ERR_INVALID, e.what());
    }
}

int copy_str(const std::string& s, char* buf, size_t len) {
    if (!buf || len <= s.size()) return fail(MSIM_ERR_INVALID, "buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return MSIM_OK;
}

}  // namespace

extern "C" {

int msim_host_map_load(const char* path, msim_host_map** out) {
    if (!path || !out) return fail(MSIM_ERR_INVALID, "null argument");
    *out = nullptr;
    return guarded([&] {
        auto m = msim_host::load_map(path);
        if (!m) return fail(MSIM_ERR_INVALID, std::string("cannot open map file ") + path);
        *out = new msim_host_map{std::move(m)};
        return static_cast<int>(MSIM_OK);
    });
}

int msim_host_map_info(const msim_host_map* m, float* width, float* height, uint32_t* n_roads,
                       uint32_t* n_connections, uint32_t* skipped) {
    if (!m) return fail(MSIM_ERR_INVALID, "null map");
    if (width) *width = m->map->width;
    if (height) *height = m->map->height;
    if (n_roads) *n_roads = static_cast<uint32_t>(m->map->roads.size());
    if (n_connections) *n_connections = static_cast<uint32_t>(m->map->connections.size());
    if (skipped) *skipped = m->map->skipped;
    return MSIM_OK;
}

int msim_host_map_read(const msim_host_map* m, msim_road* roads, uint32_t* connections) {
    if (!m) return fail(MSIM_ERR_INVALID, "null map");
    if (roads && !m->map->roads.empty())
        std::memcpy(roads, m->map->roads.data(), m->map->roads.size() * sizeof(msim_road));
    if (connections && !m->map->connections.empty())
        std::memcpy(connections, m->map->connections.data(), m->map->connections.size() * sizeof(uint32_t));
    return MSIM_OK;
}

int msim_host_map_add_guard(msim_host_map* m) {
    if (!m) return fail(MSIM_ERR_INVALID, "null map");
    msim_host::add_guard_entry(*m->map);
    return MSIM_OK;
}

void msim_host_map_free(msim_host_map* m) { delete m; }

int msim_host_make_entities(const msim_host_map* m, uint32_t n, uint64_t seed, msim_entity* out) {
    if (!m || (n && !out)) return fail(MSIM_ERR_INVALID, "null argument");
    return guarded([&] {
        const auto v = msim_host::make_entities(*m->map, n, seed);
        if (n) std::memcpy(out, v.data(), v.size() * sizeof(msim_entity));
        return static_cast<int>(MSIM_OK);
    });
}

msim_tick_rate* msim_tick_rate_new(void) { return new (std::nothrow) msim_tick_rate; }

void msim_tick_rate_tick_at(msim_tick_rate* r, int64_t now_ns) {
    if (r) r->rate.tick_at(msim_host::TickRate::clock::time_point{std::chrono::nanoseconds(now_ns)});
}

double msim_tick_rate_get(const msim_tick_rate* r) { return r ? r->rate.get_ticks() : 0.0; }

void msim_tick_rate_free(msim_tick_rate* r) { delete r; }

msim_tick_history* msim_tick_history_new(void) { return new (std::nothrow) msim_tick_history; }

void msim_tick_history_add(msim_tick_history* h, int64_t ns) {
    if (h) h->hist.add_time(std::chrono::nanoseconds(ns));
}

int64_t msim_tick_history_avg_ns(const msim_tick_history* h) {
    return h ? static_cast<int64_t>(h->hist.get_avg_time().count()) : 0;
}

int msim_tick_history_avg_str(const msim_tick_history* h, char* buf, size_t len) {
    if (!h) return fail(MSIM_ERR_INVALID, "null history");
    return copy_str(h->hist.get_avg_time_str(), buf, len);
}

void msim_tick_history_free(msim_tick_history* h) { delete h; }

int msim_host_format_time_of_day(uint64_t ms_since_midnight, char* buf, size_t len) {
    return copy_str(msim_host::format_time_of_day(std::chrono::milliseconds(ms_since_midnight)), buf, len);
}

int msim_host_sim_create(const char* map_path, uint32_t agents, float collision_radius, uint64_t seed,
                         const char* csv_path, int log_stderr, uint64_t tick_limit, msim_host_sim** out) {
    if (!map_path || !out) return fail(MSIM_ERR_INVALID, "null argument");
    *out = nullptr;
    return guarded([&] {
        msim_host::SimulatorConfig cfg;
        cfg.map_path = map_path;
        cfg.agents = agents;
        cfg.collision_radius = collision_radius;
        cfg.seed = seed;
        cfg.csv_path = csv_path ? csv_path : "";
        cfg.log_to_stderr = log_stderr != 0;
        cfg.tick_limit = tick_limit;
        auto s = std::make_unique<msim_host_sim>();
        s->sim = std::make_unique<msim_host::Simulator>(cfg);
        try {
            s->sim->init();
        } catch (const std::logic_error& e) {
            return fail(MSIM_ERR_INVALID, e.what());
        }
        *out = s.release();
        return static_cast<int>(MSIM_OK);
    });
}

int msim_host_sim_start_worker(msim_host_sim* s) {
    if (!s) return fail(MSIM_ERR_INVALID, "null simulator");
    return guarded([&] {
        s->sim->start_worker();
        return static_cast<int>(MSIM_OK);
    });
}

int msim_host_sim_stop_worker(msim_host_sim* s) {
    if (!s) return fail(MSIM_ERR_INVALID, "null simulator");
    return guarded([&] {
        s->sim->stop_worker();
        return static_cast<int>(MSIM_OK);
    });
}

int msim_host_sim_continue(msim_host_sim* s) {
    if (!s) return fail(MSIM_ERR_INVALID, "null simulator");
    if (s->sim->failed()) return fail(MSIM_ERR_STATE, "a tick failed: " + s->sim->last_error());
    s->sim->continue_simulation();
    return MSIM_OK;
}

int msim_host_sim_pause(msim_host_sim* s) {
    if (!s) return fail(MSIM_ERR_INVALID, "null simulator");
    s->sim->pause_simulation();
    return MSIM_OK;
}

int msim_host_sim_is_simulating(const msim_host_sim* s) { return s && s->sim->is_simulating() ? 1 : 0; }

int msim_host_sim_wait_ticks(msim_host_sim* s, uint64_t n, uint32_t timeout_ms) {
    if (!s) return fail(MSIM_ERR_INVALID, "null simulator");
    bool timed_out = false;
    if (s->sim->wait_ticks(n, std::chrono::milliseconds(timeout_ms), &timed_out)) return MSIM_OK;
    if (timed_out) return fail(MSIM_ERR_TIMEOUT, "timed out waiting for ticks");
    return fail(MSIM_ERR_STATE, "a tick failed: " + s->sim->last_error());
}

int msim_host_sim_stats(const msim_host_sim* s, uint64_t* ticks, double* tps, int64_t* tick_ns,
                        int64_t* update_ns, int64_t* collision_ns) {
    if (!s) return fail(MSIM_ERR_INVALID, "null simulator");
    const auto st = s->sim->stats();  // taken under the worker's lock
    if (ticks) *ticks = s->sim->ticks_done();
    if (tps) *tps = st.tps;
    if (tick_ns) *tick_ns = st.tick.count();
    if (update_ns) *update_ns = st.update.count();
    if (collision_ns) *collision_ns = st.collision.count();
    return MSIM_OK;
}

int msim_host_sim_take_entities(msim_host_sim* s, msim_entity* out, uint32_t n, int* taken) {
    if (!s || !taken || (n && !out)) return fail(MSIM_ERR_INVALID, "null argument");
    auto v = s->sim->get_entities();
    *taken = 0;
    if (!v) return MSIM_OK;
    if (v->size() != n) return fail(MSIM_ERR_INVALID, "agent count mismatch");
    if (n) std::memcpy(out, v->data(), v->size() * sizeof(msim_entity));
    *taken = 1;
    return MSIM_OK;
}

msim_sim* msim_host_sim_handle(msim_host_sim* s) { return s ? s->sim->handle() : nullptr; }

int msim_host_sim_destroy(msim_host_sim* s) {
    delete s;
    return MSIM_OK;
}

const char* msim_host_last_error(void) { return g_host_error.c_str(); }

}  // extern "C"
