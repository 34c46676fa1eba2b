/*
 * msim_oracle.h — CPU restatement of movement-sim's per-tick agent pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against: only tests, the smoke check and the benchmark's CPU-baseline leg
 * may load it.  It is written independently of the HIP sources (nothing is
 * shared with the kernels) from the reference shader
 * src/sim/shader/random_move.comp, in plain C with IEEE binary32 arithmetic,
 * no contraction (-ffp-contract=off), correctly rounded sqrtf and division.
 *
 * Layouts: orc_road is the reference's Road (src/sim/Map.hpp:12-25, 32 B).
 * Agent state is structure-of-arrays: pos/tgt/dir are 2 floats per agent,
 * rng is 4 words per agent in the shader's uvec4 element order (x, y, z, w).
 */
#ifndef MSIM_ORACLE_H
#define MSIM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_road {
    float sx, sy;
    uint32_t sidx, scnt;
    float ex, ey;
    uint32_t eidx, ecnt;
} orc_road;

/* Error codes returned by the checked entry points. */
#define ORC_OK 0
#define ORC_ERR_ROAD 1   /* roadIndex out of range                     */
#define ORC_ERR_CONN 2   /* connection-list read out of range          */

uint32_t orc_rng_next(uint32_t s[4]);
float orc_rng_float(uint32_t s[4]);
uint32_t orc_rng_range(uint32_t s[4], uint32_t lo, uint32_t hi);

float orc_length(float x, float y);
void orc_update_direction(const float tgt[2], const float pos[2], float speed, float dir[2]);

/* Route choice for one agent (random_move.comp:778-828). */
int orc_new_target(float tgt[2], uint32_t* road, uint32_t rng[4],
                   const orc_road* roads, uint32_t n_roads,
                   const uint32_t* conns, uint32_t n_conns);

/* One agent's move-phase update (random_move.comp:869-874 with 830-852).
 * Returns ORC_OK or an error; *arrived is set to 1 on arrival. */
int orc_move_agent(float pos[2], float tgt[2], float dir[2], uint32_t* road,
                   uint32_t rng[4], float speed,
                   const orc_road* roads, uint32_t n_roads,
                   const uint32_t* conns, uint32_t n_conns, int* arrived);

/* Move dispatch over all agents.  speed may be NULL (constant 1.4f).
 * initialized (may be NULL = all initialised): agents with 0 do not move and
 * are set to 1.  arrivals (may be NULL) receives the number of arrivals. */
int orc_move_phase(uint32_t n, float* pos, float* tgt, float* dir, uint32_t* road,
                   uint32_t* rng, const float* speed, uint32_t* initialized,
                   const orc_road* roads, uint32_t n_roads,
                   const uint32_t* conns, uint32_t n_conns, uint64_t* arrivals);

/* Pair predicate quad_tree_in_range(v1 = pos_j, v2 = pos_i, R)
 * (random_move.comp:551-562). */
int orc_in_range(const float pj[2], const float pi[2], float radius);

/* Cell size used by the grid restatement: the smallest power of two >= R. */
float orc_cell_size(float radius);
void orc_cells(uint32_t n, const float* pos, float cell, int32_t* cell_xy);

/* Collision dispatch.  Pair (i, j), i > j, is counted when the predicate
 * holds and agent i searches this dispatch (searching = initialized flag at
 * dispatch start; NULL = everyone).  ncount[k] receives the number of counted
 * pairs agent k is in.  Returns the number of counted pairs.
 * _grid bins agents into cells of orc_cell_size(R) and scans 3x3 cells;
 * _brute tests every pair.  Both mark every agent initialised. */
int64_t orc_collide_grid(uint32_t n, const float* pos, float radius,
                         uint32_t* initialized, uint32_t* ncount);
int64_t orc_collide_brute(uint32_t n, const float* pos, float radius,
                          uint32_t* initialized, uint32_t* ncount);

/* The same collision pass split over n_threads POSIX threads (cell binning is
 * done once, the per-agent search is split); used as the timed CPU baseline. */
int64_t orc_collide_grid_mt(uint32_t n, const float* pos, float radius,
                            uint32_t* ncount, int n_threads);
int orc_move_phase_mt(uint32_t n, float* pos, float* tgt, float* dir, uint32_t* road,
                      uint32_t* rng, const float* speed,
                      const orc_road* roads, uint32_t n_roads,
                      const uint32_t* conns, uint32_t n_conns, int n_threads);

/* Binary64 replay of the advance from the same binary32 inputs: the exact
 * step the binary32 arithmetic approximates.  pos_out gets 2 doubles per
 * agent, arrive_out 1 where the binary64 step snaps to the target. */
void orc_advance_f64(uint32_t n, const float* pos, const float* tgt,
                     const float* speed, double* pos_out, int32_t* arrive_out);

/* ---- differentiable tick (msim_diff_oracle.c) ----------------------------
 * Soft-avoidance weight w(q) = (1 - q/R^2)^3 over in-range neighbours.
 * Agents of scenario s are [agent_off[s], agent_off[s+1]); neighbours never
 * cross scenarios.  theta has one multiplier per road.  The schedule arrays
 * are [n_ticks][n] (targets: 2 floats per agent): the target and road each
 * move used and a flag (bit 0 arrived, bit 1 no move; the first tick does not
 * move when no_move_first). */
/* Threads for orc_diff_record, orc_diff_replay64* and orc_diff_adjoint64*
 * (default 1). */
void orc_diff_set_threads(int n);

/* binary32 run recording its decisions; the push terms are summed in fixed
 * point (orc_diff_fix_scale), the same order-independent sum as the device's.
 * traj (nullable) receives each tick's pre-move and post-move (pre-push)
 * positions, [n_ticks][2][n][2]. */
int orc_diff_record(uint32_t n, uint32_t n_scen, const uint32_t* agent_off, float* pos, float* tgt,
                    uint32_t* road, uint32_t* rng, const float* speed, const orc_road* roads, uint32_t n_roads,
                    const uint32_t* conns, uint32_t n_conns, const float* theta, float alpha, float radius,
                    uint32_t n_ticks, int no_move_first, float* sched_tgt, uint32_t* sched_road,
                    uint8_t* sched_flag, float* traj);
/* The record's fixed-point unit for the push terms: 2^(40 - max(0, ceil(log2 R))). */
double orc_diff_fix_scale(float radius);
/* binary64 replay under a schedule: final positions (2 doubles per agent). */
int orc_diff_replay64(uint32_t n, uint32_t n_scen, const uint32_t* agent_off, const float* pos0,
                      const float* speed, const double* theta, double alpha, double radius, uint32_t n_ticks,
                      const float* sched_tgt, const uint32_t* sched_road, const uint8_t* sched_flag,
                      double* pos_out);
/* Reverse mode of orc_diff_replay64: dL/dtheta (n_roads) and dL/dalpha from
 * dL/d(final positions). */
int orc_diff_adjoint64(uint32_t n, uint32_t n_scen, const uint32_t* agent_off, const float* pos0,
                       const float* speed, const double* theta, uint32_t n_roads, double alpha, double radius,
                       uint32_t n_ticks, const float* sched_tgt, const uint32_t* sched_road,
                       const uint8_t* sched_flag, const double* grad_pos, double* grad_theta, double* grad_alpha);
/* Arrival carry-over (the surrogate the adjoint can use through arrivals):
 * orc_diff_replay64_carry replays with every arrival carrying its leftover
 * time onto the next road, p' = T + (1 - L/s) s2 u2; final_tgt/final_road are
 * the decisions after the last tick.  orc_diff_adjoint64_ex runs the reverse
 * pass on the snapped (traj_carry = 0) or carried (1) trajectory, with
 * arrivals cutting the gradient (adj_carry = 0) or differentiated as the
 * carry-over (1).  traj32 (nullable): the binary32 trajectory to linearise
 * around instead of the binary64 replay, as orc_diff_record's traj gives it
 * ([n_ticks][2][n][2]: each tick's pre-move and post-move positions), with
 * the binary32 neighbour predicate. */
int orc_diff_replay64_carry(uint32_t n, uint32_t n_scen, const uint32_t* agent_off, const float* pos0,
                            const float* speed, const double* theta, double alpha, double radius, uint32_t n_ticks,
                            const float* sched_tgt, const uint32_t* sched_road, const uint8_t* sched_flag,
                            const float* final_tgt, const uint32_t* final_road, double* pos_out);
int orc_diff_adjoint64_ex(uint32_t n, uint32_t n_scen, const uint32_t* agent_off, const float* pos0,
                          const float* speed, const double* theta, uint32_t n_roads, double alpha, double radius,
                          uint32_t n_ticks, const float* sched_tgt, const uint32_t* sched_road,
                          const uint8_t* sched_flag, const float* final_tgt, const uint32_t* final_road,
                          int traj_carry, int adj_carry, const float* traj32, const double* grad_pos,
                          double* grad_theta, double* grad_alpha);

#ifdef __cplusplus
}
#endif

#endif
