/*
 * msim/msim_host.h — the host driver around the agent pipeline, C-ABI.
 *
 * The reference's host side (sim::Map, sim::Simulator, utils::TickRate,
 * utils::TickDurationHistory) is C++ called in-process by its UI; these
 * entry points expose the C++ mirror in movement-sim_amd/host/ to C, Python
 * (ctypes) and any other FFI.  Return codes are msim.h's MSIM_* values;
 * msim_host_last_error() describes the last failure of the calling thread.
 */
#ifndef MSIM_MSIM_HOST_H
#define MSIM_MSIM_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "msim/msim.h"
#include "msim/types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- maps (sim::Map, src/sim/Map.cpp:28-150) ----------------------------- */
typedef struct msim_host_map msim_host_map;

/* Map::load_from_file: MSIM_ERR_INVALID with "cannot open" when the file
 * cannot be read (the reference returns nullptr), MSIM_ERR_INVALID with
 * "Failed to parse map. '<field>' field missing." when a field is absent.
 * Zero-length roads are skipped; the connection list is kept as read. */
int msim_host_map_load(const char* path, msim_host_map** out);

/* Sizes of a loaded map; any pointer may be NULL. */
int msim_host_map_info(const msim_host_map* m, float* width, float* height, uint32_t* n_roads,
                       uint32_t* n_connections, uint32_t* skipped);

/* Copies the road and connection arrays (msim_host_map_info's sizes). */
int msim_host_map_read(const msim_host_map* m, msim_road* roads, uint32_t* connections);

/* Appends one guard entry (road 0) when a group's top draw would read past
 * the list (random_move.comp:808); a no-op otherwise. */
int msim_host_map_add_guard(msim_host_map* m);

void msim_host_map_free(msim_host_map* m);

/* ---- agents (Simulator::add_entities, src/sim/Simulator.cpp:114-129) ----
 * n agents on uniformly drawn roads: pos = road start, target = road end,
 * direction 0, random colour with alpha 1, random non-zero xorshift state,
 * initialized = 0.  Deterministic in seed (the reference uses
 * std::random_device). */
int msim_host_make_entities(const msim_host_map* m, uint32_t n, uint64_t seed, msim_entity* out);

/* ---- statistics (src/utils/TickRate.cpp, src/utils/TickDurationHistory.cpp) */
typedef struct msim_tick_rate msim_tick_rate;
typedef struct msim_tick_history msim_tick_history;

/* A tick-rate counter whose clock starts at time 0 ns; msim_tick_rate_tick_at
 * feeds it ticks at explicit times (ns, non-decreasing). */
msim_tick_rate* msim_tick_rate_new(void);
void msim_tick_rate_tick_at(msim_tick_rate* r, int64_t now_ns);
double msim_tick_rate_get(const msim_tick_rate* r);
void msim_tick_rate_free(msim_tick_rate* r);

/* Mean of the last 100 durations; the string is "%.2f" plus ns/us/ms/s. */
msim_tick_history* msim_tick_history_new(void);
void msim_tick_history_add(msim_tick_history* h, int64_t ns);
int64_t msim_tick_history_avg_ns(const msim_tick_history* h);
int msim_tick_history_avg_str(const msim_tick_history* h, char* buf, size_t len);
void msim_tick_history_free(msim_tick_history* h);

/* Simulator::get_time_stamp for a time of day in ms: "HH:M:SS.mmm" (the
 * reference pads hours, seconds and milliseconds, not minutes). */
int msim_host_format_time_of_day(uint64_t ms_since_midnight, char* buf, size_t len);

/* ---- the simulator (sim::Simulator, src/sim/Simulator.hpp:46-130) --------
 * init() on create: loads the map, adds the guard entry when needed, adds
 * the agents and creates the msim_sim handle.  The worker thread runs
 * sim_tick (a move dispatch and a collision dispatch, each waited for and
 * timed) while the simulation is continued, appends
 * "<time>;<tick/2>;<s update>;<s collision>;<s all>" per tick to csv_path
 * ("<agents>.csv" when NULL or empty) and, if log_stderr, to stderr, and
 * refreshes the entity and grid-cell copies after each hand-over.
 * tick_limit > 0 pauses the simulation after that many ticks. */
typedef struct msim_host_sim msim_host_sim;

int msim_host_sim_create(const char* map_path, uint32_t agents, float collision_radius, uint64_t seed,
                         const char* csv_path, int log_stderr, uint64_t tick_limit, msim_host_sim** out);
int msim_host_sim_start_worker(msim_host_sim* s);
int msim_host_sim_stop_worker(msim_host_sim* s);
int msim_host_sim_continue(msim_host_sim* s);
int msim_host_sim_pause(msim_host_sim* s);
int msim_host_sim_is_simulating(const msim_host_sim* s);

/* Blocks until the worker has completed at least n ticks in total, the
 * simulation is paused, or timeout_ms passes (MSIM_ERR_TIMEOUT).  Reports a
 * failed tick as MSIM_ERR_STATE. */
int msim_host_sim_wait_ticks(msim_host_sim* s, uint64_t n, uint32_t timeout_ms);

/* Ticks done, ticks per second, and the mean whole-tick, move-dispatch and
 * collision-dispatch durations (ns) of the last 100 ticks; any may be NULL. */
int msim_host_sim_stats(const msim_host_sim* s, uint64_t* ticks, double* tps, int64_t* tick_ns,
                        int64_t* update_ns, int64_t* collision_ns);

/* Simulator::get_entities: takes the worker's last entity copy (n agents,
 * 64-byte records) into out.  *taken is 1 when a copy was available (the
 * worker then reads a fresh one after its next tick) and 0 otherwise. */
int msim_host_sim_take_entities(msim_host_sim* s, msim_entity* out, uint32_t n, int* taken);

/* The pipeline handle (for msim_read_* calls while the worker is paused). */
msim_sim* msim_host_sim_handle(msim_host_sim* s);

int msim_host_sim_destroy(msim_host_sim* s);

const char* msim_host_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MSIM_MSIM_HOST_H */
