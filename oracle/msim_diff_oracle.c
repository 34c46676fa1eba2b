/*
 * msim_diff_oracle.c — CPU restatement of the differentiable tick and its
 * adjoint.  TEST INFRASTRUCTURE ONLY (see msim_oracle.h).
 *
 * The differentiable tick extends the reference tick (random_move.comp:841-852
 * move, :778-828 route choice, :551-562 neighbour predicate) with a per-road
 * speed multiplier theta and a soft-avoidance push:
 *   s_i = v_i * theta[road_i];  arrived (|T - p| <= s): p~ = T, route choice;
 *   else p~ = p + ((T - p) / |T - p|) * s;
 *   p' = p~ + alpha * sum_{j in scenario, j != i, in range} w(|x|^2) x,
 *   x = p~_i - p~_j, w(q) = (1 - q / R^2)^3 (C^2 across the range boundary).
 * orc_diff_record runs it in binary32 (the move exactly as orc_move_agent;
 * the push summed in fixed point, independent of the neighbours' order) and
 * records the decisions.  orc_diff_replay64 re-runs it in binary64 under recorded
 * decisions; orc_diff_adjoint64 is the reverse mode of that replay.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "msim_oracle.h"

#define ORC_SPEED 1.4f

/* The neighbour predicate: binary32 (orc_in_range) or binary64. */
typedef int (*pred_fn)(const void* pts, uint32_t i, uint32_t j, double r);

static int pred_f32(const void* pts, uint32_t i, uint32_t j, double r) {
    const float* p = (const float*)pts;
    return orc_in_range(&p[2 * j], &p[2 * i], (float)r);
}

static int pred_f64(const void* pts, uint32_t i, uint32_t j, double r) {
    const double* p = (const double*)pts;
    double dx = p[2 * i] - p[2 * j], dy = p[2 * i + 1] - p[2 * j + 1];
    return dx * dx + dy * dy < r * r;
}

/* ---- threads ------------------------------------------------------------
 * The binary64 replay and adjoint split their per-agent loops over
 * g_threads POSIX threads (orc_diff_set_threads).  Every agent's value is
 * computed by the same operations in the same order whatever the split; the
 * gradient sums are kept per thread and added in thread order at the end,
 * which only reassociates binary64 sums. */
static int g_threads = 1;

void orc_diff_set_threads(int n) { g_threads = n < 1 ? 1 : (n > 256 ? 256 : n); }

typedef void (*range_fn)(void* ctx, uint32_t lo, uint32_t hi, int worker);

typedef struct {
    range_fn f;
    void* ctx;
    uint32_t lo, hi;
    int worker;
} par_job;

static void* par_run(void* arg) {
    par_job* j = (par_job*)arg;
    j->f(j->ctx, j->lo, j->hi, j->worker);
    return NULL;
}

/* f(ctx, lo, hi, worker) over [0, n) in g_threads contiguous slices. */
static void par_for(uint32_t n, range_fn f, void* ctx) {
    int t = g_threads;
    if (t <= 1 || n < 2048) {
        f(ctx, 0, n, 0);
        return;
    }
    pthread_t th[256];
    par_job jobs[256];
    for (int k = 0; k < t; k++) {
        jobs[k] = (par_job){f, ctx, (uint32_t)(((uint64_t)n * k) / t), (uint32_t)(((uint64_t)n * (k + 1)) / t), k};
        pthread_create(&th[k], NULL, par_run, &jobs[k]);
    }
    for (int k = 0; k < t; k++) pthread_join(th[k], NULL);
}

/* ---- neighbour lists -------------------------------------------------- */

/* The points binned into a toroidal table of 2^lw x 2^lh buckets (at least
 * 4 x 4, about one bucket per point): bucket of cell (cx, cy) =
 * (cy mod 2^lh) * 2^lw + (cx mod 2^lw); the 3 x 3 cells around any cell fall
 * in 9 distinct buckets, so every neighbour is found once (the exact
 * predicate removes the other cells folded onto the same buckets).  A
 * counting sort: ids ascending within a bucket. */
typedef struct {
    uint32_t n;
    const void* pts;
    int f64;
    double r, cell;
    uint32_t lw, lh;
    uint32_t* start;  /* [m + 1] */
    uint32_t* ids;    /* [n] */
    uint32_t* off;    /* pass 1: counts at off[i + 1]; pass 2: the CSR offsets */
    uint32_t* idx;    /* pass 2 */
    int fill;
} list_ctx;

static void point_cell(const list_ctx* c, uint32_t i, int64_t* cx, int64_t* cy) {
    const float* pf = (const float*)c->pts;
    const double* pd = (const double*)c->pts;
    double x = c->f64 ? pd[2 * i] : (double)pf[2 * i];
    double y = c->f64 ? pd[2 * i + 1] : (double)pf[2 * i + 1];
    *cx = (int64_t)floor(x / c->cell);
    *cy = (int64_t)floor(y / c->cell);
}

static uint32_t bucket_at(const list_ctx* c, int64_t cx, int64_t cy) {
    return ((uint32_t)((uint64_t)cy & ((1u << c->lh) - 1u)) << c->lw) | (uint32_t)((uint64_t)cx & ((1u << c->lw) - 1u));
}

/* The j != i within range of point i, over the 3x3 cells, in ascending j
 * (into out when given); returns their number. */
static uint32_t neighbours_of(const list_ctx* c, uint32_t i, uint32_t* out) {
    pred_fn pred = c->f64 ? pred_f64 : pred_f32;
    int64_t cx, cy;
    point_cell(c, i, &cx, &cy);
    uint32_t m = 0;
    for (int64_t dy = -1; dy <= 1; dy++) {
        for (int64_t dx = -1; dx <= 1; dx++) {
            uint32_t b = bucket_at(c, cx + dx, cy + dy);
            for (uint32_t k = c->start[b]; k < c->start[b + 1]; k++) {
                uint32_t j = c->ids[k];
                if (j != i && pred(c->pts, i, j, c->r)) {
                    if (out) out[m] = j;
                    m++;
                }
            }
        }
    }
    if (out) { /* ascending j */
        for (uint32_t a = 1; a < m; a++) {
            uint32_t v = out[a], q = a;
            while (q > 0 && out[q - 1] > v) {
                out[q] = out[q - 1];
                q--;
            }
            out[q] = v;
        }
    }
    return m;
}

static void list_range(void* c_, uint32_t lo, uint32_t hi, int w) {
    (void)w;
    list_ctx* c = (list_ctx*)c_;
    for (uint32_t i = lo; i < hi; i++) {
        if (c->fill) {
            (void)neighbours_of(c, i, c->idx + c->off[i]);
        } else {
            c->off[i + 1] = neighbours_of(c, i, NULL);
        }
    }
}

/* Neighbour lists of n points (cell >= R): for each i, the j != i with
 * pred(i, j), in ascending j.  Returns CSR (off[n+1], idx) via malloc.
 * (Counted, then filled, each pass split over the threads.) */
static int neighbour_lists_mt(uint32_t n, const void* pts, int f64, double r, double cell, uint32_t** off_out,
                              uint32_t** idx_out) {
    uint32_t lg = 4;
    while (lg < 26 && (1u << lg) < n) lg++;
    list_ctx c = {n, pts, f64, r, cell, lg / 2, lg - lg / 2, NULL, NULL, NULL, NULL, 0};
    const uint32_t m = 1u << lg;
    c.start = (uint32_t*)calloc((size_t)m + 1, sizeof(uint32_t));
    c.ids = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* bk = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* off = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
    if (!c.start || !c.ids || !bk || !off) {
        free(c.start);
        free(c.ids);
        free(bk);
        free(off);
        return -1;
    }
    for (uint32_t i = 0; i < n; i++) {
        int64_t cx, cy;
        point_cell(&c, i, &cx, &cy);
        bk[i] = bucket_at(&c, cx, cy);
        c.start[bk[i] + 1]++;
    }
    for (uint32_t q = 0; q < m; q++) c.start[q + 1] += c.start[q];
    uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * m);
    if (!cur) {
        free(c.start);
        free(c.ids);
        free(bk);
        free(off);
        return -1;
    }
    memcpy(cur, c.start, sizeof(uint32_t) * m);
    for (uint32_t i = 0; i < n; i++) c.ids[cur[bk[i]]++] = i;  /* ascending ids within each bucket */
    free(cur);
    free(bk);
    c.off = off;
    par_for(n, list_range, &c);
    for (uint32_t i = 0; i < n; i++) off[i + 1] += off[i];
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * (off[n] ? off[n] : 1));
    if (!idx) {
        free(c.start);
        free(c.ids);
        free(off);
        return -1;
    }
    c.idx = idx;
    c.fill = 1;
    par_for(n, list_range, &c);
    free(c.start);
    free(c.ids);
    *off_out = off;
    *idx_out = idx;
    return 0;
}

/* The binary32 record, split over the threads per agent (the move, and the
 * push of each scenario); every agent's result is independent of the split. */
typedef struct {
    uint32_t n, t;
    int no_move_first;
    float *pos, *tgt, *ptil;
    uint32_t *road, *rng;
    const float* speed;
    const orc_road* roads;
    uint32_t n_roads;
    const uint32_t* conns;
    uint32_t n_conns;
    const float* theta;
    float* sched_tgt;
    uint32_t* sched_road;
    uint8_t* sched_flag;
    int err[256];
} rec_ctx;

static void rec_move_range(void* c_, uint32_t lo, uint32_t hi, int w) {
    rec_ctx* c = (rec_ctx*)c_;
    const uint32_t n = c->n, t = c->t;
    float dir[2];
    for (uint32_t i = lo; i < hi; i++) {
        uint32_t r = c->road[i];
        if (r >= c->n_roads) {
            c->err[w] = ORC_ERR_ROAD;
            return;
        }
        uint8_t flag = 0;
        c->sched_tgt[((size_t)t * n + i) * 2] = c->tgt[2 * i];
        c->sched_tgt[((size_t)t * n + i) * 2 + 1] = c->tgt[2 * i + 1];
        c->sched_road[(size_t)t * n + i] = r;
        if (t == 0 && c->no_move_first) {
            flag = 2;
        } else {
            float s = (c->speed ? c->speed[i] : ORC_SPEED) * c->theta[r];
            int arrived = 0;
            int rc = orc_move_agent(&c->pos[2 * i], &c->tgt[2 * i], dir, &c->road[i], &c->rng[4 * i], s, c->roads,
                                    c->n_roads, c->conns, c->n_conns, &arrived);
            if (rc != ORC_OK) {
                c->err[w] = rc;
                return;
            }
            flag = arrived ? 1 : 0;
        }
        c->sched_flag[(size_t)t * n + i] = flag;
        c->ptil[2 * i] = c->pos[2 * i];
        c->ptil[2 * i + 1] = c->pos[2 * i + 1];
    }
}

typedef struct {
    const float* ptil;
    float* pos;
    uint32_t o;
    const uint32_t *off, *idx;
    float alpha, inv_r2;
    double fix_scale;
} rec_push_ctx;

/* p' = p~ + alpha A, A = sum_j w x_ij: each binary32 term w x rounded to a
 * multiple of 1/fix_scale (ties to even) and the terms summed as 64-bit
 * integers, so that A does not depend on the order of the neighbours. */
static void rec_push_range(void* c_, uint32_t lo, uint32_t hi, int w) {
    (void)w;
    const rec_push_ctx* c = (const rec_push_ctx*)c_;
    for (uint32_t i = lo; i < hi; i++) {
        const float* me = &c->ptil[2 * (size_t)(c->o + i)];
        long long ix = 0, iy = 0;
        for (uint32_t k = c->off[i]; k < c->off[i + 1]; k++) {
            const float* q = &c->ptil[2 * (size_t)(c->o + c->idx[k])];
            float x = me[0] - q[0], y = me[1] - q[1];
            float h = 1.0f - (x * x + y * y) * c->inv_r2;
            float wv = h * h * h;
            ix += llrint((double)(wv * x) * c->fix_scale);
            iy += llrint((double)(wv * y) * c->fix_scale);
        }
        float ax = (float)((double)ix / c->fix_scale), ay = (float)((double)iy / c->fix_scale);
        float px = c->alpha * ax, py = c->alpha * ay;
        c->pos[2 * (size_t)(c->o + i)] = me[0] + px;
        c->pos[2 * (size_t)(c->o + i) + 1] = me[1] + py;
    }
}

double orc_diff_fix_scale(float radius) {
    int e = (int)ceilf(log2f(radius));
    return ldexp(1.0, 40 - (e > 0 ? e : 0));
}

int orc_diff_record(uint32_t n, uint32_t n_scen, const uint32_t* agent_off, float* pos, float* tgt,
                    uint32_t* road, uint32_t* rng, const float* speed, const orc_road* roads, uint32_t n_roads,
                    const uint32_t* conns, uint32_t n_conns, const float* theta, float alpha, float radius,
                    uint32_t n_ticks, int no_move_first, float* sched_tgt, uint32_t* sched_road,
                    uint8_t* sched_flag, float* traj) {
    const double cell = (double)orc_cell_size(radius);
    float* ptil = (float*)malloc(sizeof(float) * 2 * (n ? n : 1));
    if (!ptil) return -1;
    rec_ctx c;
    memset(&c, 0, sizeof(c));
    c.n = n;
    c.no_move_first = no_move_first;
    c.pos = pos;
    c.tgt = tgt;
    c.ptil = ptil;
    c.road = road;
    c.rng = rng;
    c.speed = speed;
    c.roads = roads;
    c.n_roads = n_roads;
    c.conns = conns;
    c.n_conns = n_conns;
    c.theta = theta;
    c.sched_tgt = sched_tgt;
    c.sched_road = sched_road;
    c.sched_flag = sched_flag;
    for (uint32_t t = 0; t < n_ticks; t++) {
        if (traj) memcpy(traj + (size_t)t * 4 * n, pos, sizeof(float) * 2 * n);
        c.t = t;
        par_for(n, rec_move_range, &c);
        for (int w = 0; w < 256; w++) {
            if (c.err[w] != 0) {
                free(ptil);
                return c.err[w];
            }
        }
        if (traj) memcpy(traj + (size_t)t * 4 * n + 2 * (size_t)n, ptil, sizeof(float) * 2 * n);
        if (alpha != 0.0f) {
            for (uint32_t sc = 0; sc < n_scen; sc++) {
                uint32_t o = agent_off[sc], m = agent_off[sc + 1] - o;
                uint32_t *off = NULL, *idx = NULL;
                if (neighbour_lists_mt(m, ptil + 2 * (size_t)o, 0, radius, cell, &off, &idx) != 0) {
                    free(ptil);
                    return -1;
                }
                rec_push_ctx pc = {ptil, pos, o, off, idx, alpha, 1.0f / (radius * radius),
                                   orc_diff_fix_scale(radius)};
                par_for(m, rec_push_range, &pc);
                free(off);
                free(idx);
            }
        }
    }
    free(ptil);
    return ORC_OK;
}

/* binary64 forward under a recorded schedule; traj (nullable) receives the
 * pre-move and post-move positions of every tick: [T][2][n][2]. */
/* The decision in force after tick t for agent i: the next tick's target and
 * road, or the run's final ones after its last tick. */
static void next_decision(uint32_t n, uint32_t n_ticks, uint32_t t, uint32_t i, const float* sched_tgt,
                          const uint32_t* sched_road, const float* final_tgt, const uint32_t* final_road,
                          double* tx, double* ty, uint32_t* r) {
    if (t + 1 < n_ticks) {
        *tx = sched_tgt[((size_t)(t + 1) * n + i) * 2];
        *ty = sched_tgt[((size_t)(t + 1) * n + i) * 2 + 1];
        *r = sched_road[(size_t)(t + 1) * n + i];
    } else {
        *tx = final_tgt[2 * (size_t)i];
        *ty = final_tgt[2 * (size_t)i + 1];
        *r = final_road[i];
    }
}

/* Carry-over geometry of an arrival from p at target T, the next target T2:
 * the distance L to T, the unit direction u towards T (0 when L = 0) and u2
 * along the next road.  Returns 0 when the next road has no direction
 * (T2 == T), where nothing is carried. */
static int carry_geom(double px, double py, double Tx, double Ty, double T2x, double T2y, double* L, double* ux,
                      double* uy, double* u2x, double* u2y) {
    double dx = Tx - px, dy = Ty - py;
    double ex = T2x - Tx, ey = T2y - Ty;
    double E = sqrt(ex * ex + ey * ey);
    *L = sqrt(dx * dx + dy * dy);
    if (!(E > 0.0)) return 0;
    *ux = *L > 0.0 ? dx / *L : 0.0;
    *uy = *L > 0.0 ? dy / *L : 0.0;
    *u2x = ex / E;
    *u2y = ey / E;
    return 1;
}

/* carry != 0: an arrival carries its leftover time onto the next road,
 * p' = T + (1 - L/s) s2 u2 (s2: the next road's speed), instead of stopping
 * at T; final_tgt/final_road give the decisions after the last tick. */
/* Per-agent step of the binary64 replay, tick c->t: p -> pt. */
typedef struct {
    uint32_t n, n_ticks, t;
    const float* speed;
    const double* theta;
    const float *sched_tgt;
    const uint32_t* sched_road;
    const uint8_t* sched_flag;
    const float* final_tgt;
    const uint32_t* final_road;
    int carry;
    double *p, *pt;
    uint32_t pad;
    double alpha, radius;
} replay_ctx;

static void replay_move_range(void* c_, uint32_t lo, uint32_t hi, int w) {
    (void)w;
    const replay_ctx* c = (const replay_ctx*)c_;
    const uint32_t n = c->n, t = c->t;
    const double* theta = c->theta;
    double* p = c->p;
    double* pt = c->pt;
    for (uint32_t i = lo; i < hi; i++) {
        uint8_t f = c->sched_flag[(size_t)t * n + i];
        double Tx = c->sched_tgt[((size_t)t * n + i) * 2], Ty = c->sched_tgt[((size_t)t * n + i) * 2 + 1];
        if (f & 2) {
            pt[2 * i] = p[2 * i];
            pt[2 * i + 1] = p[2 * i + 1];
        } else if (f & 1) {
            pt[2 * i] = Tx;
            pt[2 * i + 1] = Ty;
            double T2x, T2y, L, ux, uy, u2x, u2y;
            uint32_t r2;
            if (c->carry) {
                next_decision(n, c->n_ticks, t, i, c->sched_tgt, c->sched_road, c->final_tgt, c->final_road, &T2x,
                              &T2y, &r2);
                if (carry_geom(p[2 * i], p[2 * i + 1], Tx, Ty, T2x, T2y, &L, &ux, &uy, &u2x, &u2y)) {
                    double v = (double)(c->speed ? c->speed[i] : ORC_SPEED);
                    double sp = v * theta[c->sched_road[(size_t)t * n + i]], s2 = v * theta[r2];
                    pt[2 * i] = Tx + (1.0 - L / sp) * s2 * u2x;
                    pt[2 * i + 1] = Ty + (1.0 - L / sp) * s2 * u2y;
                }
            }
        } else {
            double s = (double)(c->speed ? c->speed[i] : ORC_SPEED) * theta[c->sched_road[(size_t)t * n + i]];
            double dx = Tx - p[2 * i], dy = Ty - p[2 * i + 1];
            double len = sqrt(dx * dx + dy * dy);
            pt[2 * i] = p[2 * i] + dx / len * s;
            pt[2 * i + 1] = p[2 * i + 1] + dy / len * s;
        }
    }
}

/* Avoidance push of one scenario's agents (offset o): p = pt + alpha A. */
typedef struct {
    const replay_ctx* c;
    uint32_t o;
    const uint32_t *off, *idx;
} avoid_ctx;

static void replay_avoid_range(void* a_, uint32_t lo, uint32_t hi, int w) {
    (void)w;
    const avoid_ctx* a = (const avoid_ctx*)a_;
    const double* pt = a->c->pt;
    double* p = a->c->p;
    const double radius = a->c->radius, alpha = a->c->alpha;
    for (uint32_t i = lo; i < hi; i++) {
        const double* me = &pt[2 * (size_t)(a->o + i)];
        double ax = 0.0, ay = 0.0;
        for (uint32_t k = a->off[i]; k < a->off[i + 1]; k++) {
            const double* q = &pt[2 * (size_t)(a->o + a->idx[k])];
            double x = me[0] - q[0], y = me[1] - q[1];
            double h = 1.0 - (x * x + y * y) / (radius * radius);
            ax += h * h * h * x;
            ay += h * h * h * y;
        }
        p[2 * (size_t)(a->o + i)] = me[0] + alpha * ax;
        p[2 * (size_t)(a->o + i) + 1] = me[1] + alpha * ay;
    }
}

/* The binary64 replay under recorded decisions; traj (nullable) receives
 * every tick's pre-move and post-move (pre-push) positions. */
static int replay64(uint32_t n, uint32_t n_scen, const uint32_t* agent_off, const float* pos0, const float* speed,
                    const double* theta, double alpha, double radius, uint32_t n_ticks, const float* sched_tgt,
                    const uint32_t* sched_road, const uint8_t* sched_flag, const float* final_tgt,
                    const uint32_t* final_road, int carry, double* pos_out, double* traj) {
    const double cell = (double)orc_cell_size((float)radius);
    double* p = (double*)malloc(sizeof(double) * 2 * (n ? n : 1));
    double* pt = (double*)malloc(sizeof(double) * 2 * (n ? n : 1));
    if (!p || !pt) {
        free(p);
        free(pt);
        return -1;
    }
    for (uint32_t i = 0; i < 2 * n; i++) p[i] = (double)pos0[i];
    replay_ctx c = {n, n_ticks, 0, speed, theta, sched_tgt, sched_road, sched_flag, final_tgt, final_road, carry,
                    p, pt, 0, alpha, radius};
    for (uint32_t t = 0; t < n_ticks; t++) {
        if (traj) memcpy(traj + (size_t)t * 4 * n, p, sizeof(double) * 2 * n);
        c.t = t;
        par_for(n, replay_move_range, &c);
        if (traj) memcpy(traj + (size_t)t * 4 * n + 2 * (size_t)n, pt, sizeof(double) * 2 * n);
        memcpy(p, pt, sizeof(double) * 2 * n);
        if (alpha != 0.0) {
            for (uint32_t sc = 0; sc < n_scen; sc++) {
                uint32_t o = agent_off[sc], m = agent_off[sc + 1] - o;
                uint32_t *off = NULL, *idx = NULL;
                if (neighbour_lists_mt(m, pt + 2 * (size_t)o, 1, radius, cell, &off, &idx) != 0) {
                    free(p);
                    free(pt);
                    return -1;
                }
                avoid_ctx a = {&c, o, off, idx};
                par_for(m, replay_avoid_range, &a);
                free(off);
                free(idx);
            }
        }
    }
    memcpy(pos_out, p, sizeof(double) * 2 * n);
    free(p);
    free(pt);
    return 0;
}

int orc_diff_replay64(uint32_t n, uint32_t n_scen, const uint32_t* agent_off, const float* pos0,
                      const float* speed, const double* theta, double alpha, double radius, uint32_t n_ticks,
                      const float* sched_tgt, const uint32_t* sched_road, const uint8_t* sched_flag,
                      double* pos_out) {
    return replay64(n, n_scen, agent_off, pos0, speed, theta, alpha, radius, n_ticks, sched_tgt, sched_road,
                    sched_flag, NULL, NULL, 0, pos_out, NULL);
}

int orc_diff_replay64_carry(uint32_t n, uint32_t n_scen, const uint32_t* agent_off, const float* pos0,
                            const float* speed, const double* theta, double alpha, double radius, uint32_t n_ticks,
                            const float* sched_tgt, const uint32_t* sched_road, const uint8_t* sched_flag,
                            const float* final_tgt, const uint32_t* final_road, double* pos_out) {
    return replay64(n, n_scen, agent_off, pos0, speed, theta, alpha, radius, n_ticks, sched_tgt, sched_road,
                    sched_flag, final_tgt, final_road, 1, pos_out, NULL);
}

int orc_diff_adjoint64(uint32_t n, uint32_t n_scen, const uint32_t* agent_off, const float* pos0,
                       const float* speed, const double* theta, uint32_t n_roads, double alpha, double radius,
                       uint32_t n_ticks, const float* sched_tgt, const uint32_t* sched_road,
                       const uint8_t* sched_flag, const double* grad_pos, double* grad_theta, double* grad_alpha) {
    return orc_diff_adjoint64_ex(n, n_scen, agent_off, pos0, speed, theta, n_roads, alpha, radius, n_ticks, sched_tgt,
                                 sched_road, sched_flag, NULL, NULL, 0, 0, NULL, grad_pos, grad_theta, grad_alpha);
}

/* Reverse sweep of one tick: the avoidance part (per scenario) and the move
 * part, each split over the threads; per-worker sums ga[w], gt + w * n_roads. */
typedef struct {
    uint32_t n, n_ticks, t, o, n_roads;
    const double *pre, *pt, *lam;
    double* lt;
    double* lam_out;
    const uint32_t *off, *idx;
    double alpha, inv_r2;
    double* ga;  /* [threads] */
    double* gt;  /* [threads][n_roads] */
    const float* speed;
    const double* theta;
    const float* sched_tgt;
    const uint32_t* sched_road;
    const uint8_t* sched_flag;
    const float* final_tgt;
    const uint32_t* final_road;
    int adj_carry;
} adj_ctx;

static void adj_avoid_range(void* c_, uint32_t lo, uint32_t hi, int w) {
    adj_ctx* c = (adj_ctx*)c_;
    const double *pt = c->pt, *lam = c->lam;
    double ga = 0.0;
    for (uint32_t i = lo; i < hi; i++) {
        size_t gi = c->o + i;
        double gx = 0.0, gy = 0.0, ax = 0.0, ay = 0.0;
        for (uint32_t k = c->off[i]; k < c->off[i + 1]; k++) {
            size_t gj = c->o + c->idx[k];
            double x = pt[2 * gi] - pt[2 * gj], y = pt[2 * gi + 1] - pt[2 * gj + 1];
            double h = 1.0 - (x * x + y * y) * c->inv_r2;
            double wv = h * h * h, wp = -3.0 * h * h * c->inv_r2;
            double dlx = lam[2 * gi] - lam[2 * gj], dly = lam[2 * gi + 1] - lam[2 * gj + 1];
            double xd = 2.0 * wp * (x * dlx + y * dly);
            gx += wv * dlx + xd * x;
            gy += wv * dly + xd * y;
            ax += wv * x;
            ay += wv * y;
        }
        c->lt[2 * gi] = lam[2 * gi] + c->alpha * gx;
        c->lt[2 * gi + 1] = lam[2 * gi + 1] + c->alpha * gy;
        ga += lam[2 * gi] * ax + lam[2 * gi + 1] * ay;
    }
    c->ga[w] += ga;
}

static void adj_move_range(void* c_, uint32_t lo, uint32_t hi, int w) {
    adj_ctx* c = (adj_ctx*)c_;
    const uint32_t n = c->n, t = c->t;
    const double *pre = c->pre, *lt = c->lt, *theta = c->theta;
    double* lam = c->lam_out;
    double* gt = c->gt + (size_t)w * c->n_roads;
    for (uint32_t i = lo; i < hi; i++) {
        uint8_t f = c->sched_flag[(size_t)t * n + i];
        if (f & 1) {
            /* frozen: the snap is constant; carry: the derivatives of
             * p' = T + (1 - L/s) s2 u2 (L = |T - p|, s, s2 the two roads' speeds) */
            double T2x, T2y, L, ux, uy, u2x, u2y;
            uint32_t r2;
            double lx = lt[2 * i], ly = lt[2 * i + 1];
            lam[2 * i] = 0.0;
            lam[2 * i + 1] = 0.0;
            if (c->adj_carry) {
                next_decision(n, c->n_ticks, t, i, c->sched_tgt, c->sched_road, c->final_tgt, c->final_road, &T2x,
                              &T2y, &r2);
                double Tx = c->sched_tgt[((size_t)t * n + i) * 2], Ty = c->sched_tgt[((size_t)t * n + i) * 2 + 1];
                if (carry_geom(pre[2 * i], pre[2 * i + 1], Tx, Ty, T2x, T2y, &L, &ux, &uy, &u2x, &u2y)) {
                    uint32_t r = c->sched_road[(size_t)t * n + i];
                    double v = (double)(c->speed ? c->speed[i] : ORC_SPEED);
                    double sp = v * theta[r], s2 = v * theta[r2];
                    if (sp > 0.0) {  /* a zero speed carries nothing: the frozen snap */
                        double g2 = u2x * lx + u2y * ly;
                        lam[2 * i] = ux * (s2 / sp) * g2;
                        lam[2 * i + 1] = uy * (s2 / sp) * g2;
                        gt[r] += v * (L * s2 / (sp * sp)) * g2;
                        gt[r2] += v * (1.0 - L / sp) * g2;
                    }
                }
            }
            continue;
        }
        if (f & 2) {
            lam[2 * i] = lt[2 * i];
            lam[2 * i + 1] = lt[2 * i + 1];
            continue;
        }
        uint32_t r = c->sched_road[(size_t)t * n + i];
        double v = (double)(c->speed ? c->speed[i] : ORC_SPEED);
        double s = v * theta[r];
        double dx = (double)c->sched_tgt[((size_t)t * n + i) * 2] - pre[2 * i];
        double dy = (double)c->sched_tgt[((size_t)t * n + i) * 2 + 1] - pre[2 * i + 1];
        double len = sqrt(dx * dx + dy * dy);
        double ux = dx / len, uy = dy / len;
        double ul = ux * lt[2 * i] + uy * lt[2 * i + 1];
        double fct = s / len;
        lam[2 * i] = lt[2 * i] - fct * (lt[2 * i] - ux * ul);
        lam[2 * i + 1] = lt[2 * i + 1] - fct * (lt[2 * i + 1] - uy * ul);
        gt[r] += v * ul;
    }
}

int orc_diff_adjoint64_ex(uint32_t n, uint32_t n_scen, const uint32_t* agent_off, const float* pos0,
                          const float* speed, const double* theta, uint32_t n_roads, double alpha, double radius,
                          uint32_t n_ticks, const float* sched_tgt, const uint32_t* sched_road,
                          const uint8_t* sched_flag, const float* final_tgt, const uint32_t* final_road,
                          int traj_carry, int adj_carry, const float* traj32, const double* grad_pos,
                          double* grad_theta, double* grad_alpha) {
    const double cell = (double)orc_cell_size((float)radius);
    const int nw = g_threads;
    const size_t n2 = 2 * (size_t)(n ? n : 1);
    double* traj = (double*)malloc(sizeof(double) * (traj32 ? 2 * n2 : 2 * n2 * (n_ticks ? n_ticks : 1)));
    double* fin = (double*)malloc(sizeof(double) * n2);
    double* lam = (double*)malloc(sizeof(double) * n2);
    double* lt = (double*)malloc(sizeof(double) * n2);
    double* gt = (double*)calloc((size_t)nw * (n_roads ? n_roads : 1), sizeof(double));
    double* ga = (double*)calloc((size_t)nw, sizeof(double));
    int rc = -1;
    if (!traj || !fin || !lam || !lt || !gt || !ga) goto done;
    if (!traj32 && replay64(n, n_scen, agent_off, pos0, speed, theta, alpha, radius, n_ticks, sched_tgt, sched_road,
                            sched_flag, final_tgt, final_road, traj_carry, fin, traj) != 0) {
        goto done;
    }
    memcpy(lam, grad_pos, sizeof(double) * 2 * n);
    for (uint32_t t = n_ticks; t-- > 0;) {
        const double* pre = traj32 ? traj : traj + (size_t)t * 4 * n;
        const double* pt = pre + 2 * (size_t)n;
        const float* pt32 = NULL;
        if (traj32) {  /* the given binary32 trajectory, widened */
            const float* src = traj32 + (size_t)t * 4 * n;
            for (size_t k = 0; k < 4 * (size_t)n; k++) traj[k] = (double)src[k];
            pt32 = src + 2 * (size_t)n;
        }
        adj_ctx c = {n, n_ticks, t, 0, n_roads, pre, pt, lam, lt, lam, NULL, NULL, alpha, 1.0 / (radius * radius),
                     ga, gt, speed, theta, sched_tgt, sched_road, sched_flag, final_tgt, final_road, adj_carry};
        /* avoidance: lt = lam + alpha sum J (lam_i - lam_j); dL/dalpha += lam . A */
        memcpy(lt, lam, sizeof(double) * 2 * n);
        for (uint32_t sc = 0; sc < n_scen; sc++) {
            uint32_t o = agent_off[sc], m = agent_off[sc + 1] - o;
            uint32_t *off = NULL, *idx = NULL;
            /* neighbours by the binary32 predicate on a binary32 trajectory */
            int lists = pt32 ? neighbour_lists_mt(m, pt32 + 2 * (size_t)o, 0, radius, cell, &off, &idx)
                             : neighbour_lists_mt(m, pt + 2 * (size_t)o, 1, radius, cell, &off, &idx);
            if (lists != 0) goto done;
            c.o = o;
            c.off = off;
            c.idx = idx;
            par_for(m, adj_avoid_range, &c);
            free(off);
            free(idx);
        }
        /* move (writes lam from lt) */
        par_for(n, adj_move_range, &c);
    }
    memset(grad_theta, 0, sizeof(double) * n_roads);
    *grad_alpha = 0.0;
    for (int w = 0; w < nw; w++) {
        for (uint32_t r = 0; r < n_roads; r++) grad_theta[r] += gt[(size_t)w * n_roads + r];
        *grad_alpha += ga[w];
    }
    rc = 0;
done:
    free(traj);
    free(fin);
    free(lam);
    free(lt);
    free(gt);
    free(ga);
    return rc;
}
