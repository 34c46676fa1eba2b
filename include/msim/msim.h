/*
 * msim/msim.h — C-ABI of the MI355X (gfx950) per-tick agent pipeline.
 *
 * This is the drop-in boundary for movement-sim's compute-shader tick.  The
 * reference builds a Kompute pipeline from random_move.comp, binds seven
 * storage buffers and calls it twice per tick; every entry point below names
 * the reference call it replaces.  All functions are plain C: POD pointers and
 * sizes in, an int status out (MSIM_OK == 0), no exceptions across the ABI.
 * On failure msim_last_error(handle) (or msim_get_last_error() when no handle
 * exists yet) returns a message.
 *
 * Ownership: the caller owns every host array it passes; msim_create copies
 * them into device memory owned by the handle.  Readbacks write into caller
 * buffers.  A handle is bound to one device and one HIP stream and is not
 * thread-safe (the reference drives its pipeline from one worker thread,
 * src/sim/Simulator.cpp:159-211); distinct handles may be used from distinct
 * threads.
 */
#ifndef MSIM_MSIM_H
#define MSIM_MSIM_H

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MSIM_OK = 0,
    MSIM_ERR_INVALID = 1,  /* bad argument or malformed map / agent data */
    MSIM_ERR_HIP = 2,      /* a HIP runtime call failed */
    MSIM_ERR_OOM = 3,      /* device allocation failed */
    MSIM_ERR_STATE = 4,    /* call not valid in the handle's current state */
    MSIM_ERR_TIMEOUT = 5   /* host watchdog expired waiting for the stream */
};

typedef struct msim_sim msim_sim;

/* Tuning knobs; zero-initialise and set struct_size = sizeof(msim_options).
 * Every zero field selects the default noted beside it. */
typedef struct msim_options {
    uint32_t struct_size;
    float cell_size;           /* 0: smallest power of two >= collisionRadius  */
    uint32_t table_log2;       /* 0: auto (about one bucket per agent)         */
    uint32_t block_size;       /* 0: 256 threads (4 wave64s)                   */
    uint32_t reorder_interval; /* 0: never; k: spatially renumber every k ticks*/
    uint32_t graph_ticks;      /* ticks per captured hipGraph segment, rounded
                                  to whole reorder periods (0: 32; batched: 100)*/
    uint32_t history_capacity; /* 0: 4096 entries of per-tick pair counts      */
    int32_t device;            /* -1 (or 0 with struct_size 0): current device */
    uint32_t watchdog_ms;      /* 0: 600000 ms host-side wait limit            */
    uint32_t table_width_log2; /* 0: auto; buckets per table row = 2^this      */
    uint32_t segment_log2;     /* 0: auto; buckets per rebinning workgroup
                                  (batched: for every scenario, 9..its table) */
    uint32_t binning;          /* 0: segment-owned rebinning (LDS histograms,
                                  per-segment offsets, count and place passes);
                                  1: global-atomic counting; 2: segment-owned
                                  in one pass (count, look-back over the
                                  earlier segments' totals, place); 3: sorted
                                  state: agent state kept in bucket order and
                                  advanced and rebinned by one launch per move
                                  dispatch (reorder_interval is ignored);
                                  batched: 0, 2 or 3                            */
    uint32_t route_queue;      /* 0: arrivals routed inside the advance kernel;
                                  1: queued and routed by a separate pass       */
    uint32_t collide_mode;     /* 0: each wave stages its candidates in LDS with
                                  16-byte loads, 2 slots per lane; 1: every
                                  lane loads its own; 2: staged, 1 slot per
                                  lane                                          */
    uint32_t bin_unroll;       /* window slots each rebinning thread loads per
                                  round: 4 or 8 (0: 8 for segments of up to
                                  2^13 buckets, else 4); binning 3: 2, 4 or 8
                                  (0: 4)                                        */
} msim_options;

typedef struct msim_info {
    uint32_t n_agents;
    uint32_t n_roads;
    uint32_t n_connections;
    uint32_t tick;             /* PushConsts.tick after the last dispatch     */
    uint64_t ticks_simulated;  /* completed move dispatches                   */
    float cell_size;
    float collision_radius;
    uint32_t table_width;      /* toroidal bucket table, buckets per row      */
    uint32_t table_height;
    uint32_t uniform_speed;    /* 1 when no per-agent speed array was given   */
    uint32_t n_uninitialized;  /* agents whose Entity.initialized is still 0  */
    uint64_t device_bytes;     /* device memory held by the handle            */
    uint32_t kernels_per_tick; /* launches of one fused tick (msim_time_ticks) */
    uint32_t binning;          /* msim_options.binning in effect              */
} msim_info;

/* Replaces Simulator::init's buffer/pipeline creation (src/sim/Simulator.cpp:
 * 44-106: mgr->tensor x7, mgr->algorithm) and the one-time upload
 * OpTensorSyncDevice (:189-193).
 *   roads/n_roads, connections/n_connections: the Map buffers bound at
 *     random_move.comp:41-42.  Every endpoint group must satisfy
 *     connectedIndex + connectedCount < n_connections (one entry after the
 *     group is readable, as the shader reads connections[idx + count] when the
 *     draw returns the top offset), and every entry must be < n_roads.
 *   entities/n_entities: the Entity buffer (binding 0).
 *   speed: per-agent speed in m/tick, or NULL for the shader's constant 1.4f
 *     (random_move.comp:750).
 *   pc: PushConsts (PushConsts.hpp:9-20); tick is the value before the first
 *     dispatch (the reference starts at 1, Simulator.cpp:101).
 *   opt: NULL for defaults. */
int msim_create(const msim_road* roads, uint32_t n_roads,
                const uint32_t* connections, uint32_t n_connections,
                const msim_entity* entities, uint32_t n_entities,
                const float* speed, const msim_push_consts* pc,
                const msim_options* opt, msim_sim** out);

/* Replaces one calcSeq->eval<kp::OpAlgoDispatch>(algo, pushConsts)
 * (Simulator.cpp:224 and :235).  pc->tick selects the phase as the shader's
 * main() does (random_move.comp:860-879): even = move (advance, route choice,
 * spatial-index rebuild), odd = collision (neighbour pass).  Agents whose
 * Entity.initialized is 0 only enter the spatial index and become initialised
 * (random_move.comp:863-867).  Asynchronous on the handle's stream. */
int msim_dispatch(msim_sim* sim, const msim_push_consts* pc);

/* n_ticks iterations of Simulator::sim_tick (Simulator.cpp:213-279) without
 * the per-tick readbacks: each is tick+=1 -> dispatch, tick+=1 -> dispatch.
 * Runs as replays of a captured hipGraph of fused ticks.  Asynchronous. */
int msim_run_ticks(msim_sim* sim, uint32_t n_ticks);

/* Blocks until the handle's stream drains (bounded by the watchdog), then
 * reports MSIM_ERR_STATE if a device-side invariant check failed (every agent
 * binned exactly once; bounded look-back waits). */
int msim_sync(msim_sim* sim);

/* Replaces OpTensorSyncLocal({tensorEntities}) + tensor->vector<Entity>()
 * (Simulator.cpp:197,248-263): writes entities [first, first+count) in the
 * 64-byte layout.  color is (0,0,1,1) for an agent that was in at least one
 * pair in the last collision dispatch and (0,1,0,1) otherwise
 * (random_move.comp:545-548,876), or the uploaded colour before any
 * collision dispatch.  Synchronous. */
int msim_read_entities(msim_sim* sim, msim_entity* out, uint32_t first, uint32_t count);

/* Replaces the debugData readback (Simulator.cpp:199,273).  out[0] counts
 * spatial-index insertions (initial inserts plus cell changes; the
 * reference's quadtree count at random_move.comp:311 is implementation
 * specific), out[1] the cumulative neighbour-pair count modulo 2^32
 * (random_move.comp:548), out[2..9] are 0.  Synchronous. */
int msim_read_debug(msim_sim* sim, uint32_t out[10]);

/* Structure-of-arrays readback in agent order; any pointer may be NULL.
 * pos/target/direction: 2 floats per agent; rand_state: 4 words per agent;
 * neighbour_count: neighbours (both ends of every pair) found for each agent
 * in the last collision dispatch.  Synchronous. */
int msim_read_state(msim_sim* sim, float* pos, float* target, float* direction,
                    uint32_t* road_index, uint32_t* rand_state, uint32_t* neighbour_count);

/* Cumulative pair total (64-bit) and per-tick pair counts of the most recent
 * min(n, history) collision dispatches, oldest first; *n_out receives the
 * number written.  Synchronous. */
int msim_read_pairs(msim_sim* sim, uint64_t* total, uint32_t* per_tick, uint32_t n, uint32_t* n_out);

/* Spatial-index export, the uniform-grid counterpart of get_quad_tree_nodes
 * (Simulator.cpp:149-153; node layout GpuQuadTree.hpp:14-41).  cell_xy gets
 * the integer cell (floor(x/cell), floor(y/cell)) of every agent as binned by
 * the last move dispatch (2 ints per agent).  Synchronous. */
int msim_read_cells(msim_sim* sim, int32_t* cell_xy);

/* Grid occupancy for the UI overlay, the counterpart of the quadtree nodes
 * get_quad_tree_nodes hands the UI (Simulator.cpp:149-153, Node layout
 * GpuQuadTree.hpp:14-41, drawn as rectangles by QuadTreeGridGlObject.cpp:
 * 30-50): every occupied cell of the last binning, cell_xy = (cx, cy) (the
 * world rectangle [cx, cx + 1) x [cy, cy + 1) times msim_info.cell_size) and
 * counts = its agents, computed on the device.  *n_cells receives the number
 * of occupied cells; min(*n_cells, capacity) entries are written, in no
 * particular order.  Synchronous. */
int msim_read_grid_occupancy(msim_sim* sim, int32_t* cell_xy, uint32_t* counts, uint32_t capacity,
                             uint32_t* n_cells);

/* msim_read_cells without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The cells are staged
 * in a static device array of 2^21 agents (16 MiB, held by each GPU the
 * library runs on until the process ends); handles with more agents are
 * rejected with MSIM_ERR_INVALID.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (from different threads) race.
 * Synchronous. */
int msim_read_cells_noalloc(msim_sim* sim, int32_t* cell_xy);

int msim_get_info(msim_sim* sim, msim_info* out);

/* Time n_ticks fused ticks on the handle's stream with HIP events.
 * ms_total receives the elapsed time; kernel_ms (may be NULL) receives the
 * summed duration of each kernel of the tick, measured by bracketing every
 * launch with events in a second, separate pass of n_ticks ticks, in launch
 * order (see msim_kernel_name; n_kernel_ms entries at most).  Both passes
 * advance the simulation. */
int msim_time_ticks(msim_sim* sim, uint32_t n_ticks, float* ms_total,
                    float* kernel_ms, uint32_t n_kernel_ms);

/* Name of the k-th kernel of a fused tick in msim_time_ticks' order, or NULL
 * for k >= msim_info.kernels_per_tick. */
const char* msim_kernel_name(const msim_sim* sim, uint32_t k);

/* Stream the handle enqueues on (a hipStream_t), for callers that order
 * their own work after it. */
int msim_get_stream(msim_sim* sim, void** stream);

int msim_destroy(msim_sim* sim);

/* ---- batched form ---------------------------------------------------------
 * Many independent scenarios (own map, own agents, same PushConsts) advanced
 * together: every kernel launch of a tick covers all scenarios, and ticks run
 * as captured hipGraph segments.  No reference counterpart (the reference runs
 * one map); each scenario's state is bit-identical to a plain msim_sim run of
 * it.  Arrays are packed scenario after scenario; road_off / conn_off /
 * agent_off (n_scen + 1 entries each) delimit them.  Road indices inside
 * roads, connections and entities are local to their scenario, and each
 * scenario's connection block must satisfy msim_create's rules on its own
 * (including the guard entry).  All agents must share one Entity.initialized
 * value; pc->tick must be odd (the next dispatch is a move), and only
 * pc->collisionRadius and pc->tick are used.  speed may be NULL (1.4 m/tick for
 * all).  Options: cell_size, reorder_interval (default 20), graph_ticks
 * (default 100), segment_log2, binning (0, 2 or 3), bin_unroll, collide_mode,
 * device and watchdog_ms apply; each scenario's table size is chosen from
 * its agent count, and its row width from its extent unless
 * table_width_log2 is set (then clamped per scenario to [2, table_log2 - 2]).
 * msim_batch_last_error(NULL) reports a failed create. */
typedef struct msim_batch msim_batch;

int msim_batch_create(uint32_t n_scen, const uint32_t* road_off, const msim_road* roads,
                      const uint32_t* conn_off, const uint32_t* connections,
                      const uint32_t* agent_off, const msim_entity* entities, const float* speed,
                      const msim_push_consts* pc, const msim_options* opt, msim_batch** out);

/* n_ticks sim_ticks of every scenario (graph segments of graph_ticks).  With
 * every agent uninitialised, the first tick is the reference's insert-only
 * move dispatch followed by a collision dispatch. */
int msim_batch_run_ticks(msim_batch* b, uint32_t n_ticks);

int msim_batch_sync(msim_batch* b);

/* One scenario's state in its local agent order (any pointer may be NULL);
 * road indices are local to the scenario. */
int msim_batch_read_state(msim_batch* b, uint32_t scen, float* pos, float* target, float* direction,
                          uint32_t* road_index, uint32_t* rand_state, uint32_t* neighbour_count);

/* Cumulative pair totals (n_scen entries) and the pairs of each scenario's
 * last collision dispatch (n_scen entries); either may be NULL. */
int msim_batch_read_pairs(msim_batch* b, uint64_t* totals, uint32_t* last);

/* Elapsed milliseconds of n_ticks batched ticks (HIP events).  kernel_ms (may
 * be NULL) receives each kernel's summed duration over a second, separate
 * pass of n_ticks ticks in which every launch is bracketed by events, in
 * msim_batch_kernel_name order.  Both passes advance the simulation. */
int msim_batch_time_ticks(msim_batch* b, uint32_t n_ticks, float* ms_total, float* kernel_ms,
                          uint32_t n_kernel_ms);

/* Name of the k-th kernel of the handle's batched tick, or NULL past the
 * last one. */
const char* msim_batch_kernel_name(const msim_batch* b, uint32_t k);

int msim_batch_destroy(msim_batch* b);
const char* msim_batch_last_error(const msim_batch* b);

/* ---- differentiable tick on a batched handle ------------------------------
 * No reference counterpart.  Per tick: speed v_i * theta[road_i], the plain
 * move (arrivals snap and take the plain route choice), then a soft-avoidance
 * push p += alpha * sum_j (1 - |x_ij|^2/R^2)^3 x_ij over the neighbours of the
 * collision dispatch (x_ij = p_i - p_j).  theta == 1, alpha == 0 is the plain
 * tick.  Every run starts from the handle's state at setup; after setup the
 * handle runs differentiable ticks only (msim_batch_run_ticks refuses), and
 * msim_batch_read_state shows the state after the last simulated tick.
 * Pointers marked (device) are device memory on the handle's device; calls
 * block until their work is done. */

/* Allocates checkpoints (every checkpoint_every ticks) and a segment tape for
 * runs of n_ticks ticks; keep_schedule != 0 also keeps every tick's decisions
 * (targets, roads, arrival flags) for replays and readback. */
int msim_batch_diff_setup(msim_batch* b, uint32_t n_ticks, uint32_t checkpoint_every, int keep_schedule);

/* n_ticks differentiable ticks with theta (device, one multiplier per packed
 * road) and alpha.  replay != 0 reuses the decisions recorded by the last
 * forward run instead of making them (frozen route choices; finite
 * differences of the adjoint's function).  final_pos (device, nullable)
 * receives the end positions, 2 floats per agent in packed caller order. */
int msim_batch_diff_forward(msim_batch* b, const float* theta, float alpha, int replay, float* final_pos);

/* Adjoint of the last (non-replay) forward run: grad_pos (device, 2 floats
 * per agent, caller order) is dL/d(final positions); grad_theta (device, one
 * double per packed road) and grad_alpha (device, 1 double) receive dL/dtheta
 * and dL/dalpha with the run's decisions and neighbour cells frozen. */
int msim_batch_diff_backward(msim_batch* b, const float* grad_pos, double* grad_theta, double* grad_alpha);

/* Number of recorded (non-replay) forward runs so far.  msim_batch_diff_backward
 * differentiates the most recent one; a caller holding the result of an
 * earlier run compares this number to tell.  No reference counterpart. */
int msim_batch_diff_generation(msim_batch* b, uint64_t* out);

/* One tick's recorded decisions (host, caller order, packed road indices):
 * target (2 floats per agent), road, flag (bit 0 arrived, bit 1 no move).
 * tick == n_ticks gives the decisions in force after the run (flag 0). */
int msim_batch_diff_read_schedule(msim_batch* b, uint32_t tick, float* target, uint32_t* road, uint8_t* flag);

/* How msim_batch_diff_backward treats arrivals: 0 (default) differentiates
 * the frozen-decision run, in which an arrival's snap to its target cuts the
 * gradient; 1 differentiates each arrival as if it carried its leftover time
 * onto the next road, p' = T + (1 - L/s) s2 u2 (L: distance to the target T
 * before the move, s and s2: the two roads' speeds, u2: the next road's
 * direction), so earlier roads' multipliers see how they shift later
 * positions.  Mode 1 is a surrogate: the run itself still snaps. */
int msim_batch_diff_set_arrival_gradient(msim_batch* b, int mode);

const char* msim_last_error(const msim_sim* sim);
const char* msim_get_last_error(void);

/* Library build identification: "msim <version> gfx950 fp-contract=off". */
const char* msim_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MSIM_MSIM_H */
