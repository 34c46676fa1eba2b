/*
 * msim_oracle.c — CPU restatement of movement-sim's per-tick agent pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (see msim_oracle.h).  Build with
 * -ffp-contract=off -fno-fast-math so every binary32 operation rounds once,
 * in the order the shader writes it.
 */
#include "msim_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* random_move.comp:750  float SPEED = 1.4; */
#define ORC_SPEED 1.4f

/* random_move.comp:725-736  uint next(inout uvec4 state) */
uint32_t orc_rng_next(uint32_t s[4]) {
    uint32_t t = s[3];
    uint32_t x = s[0];
    s[3] = s[2];
    s[2] = s[1];
    s[1] = x;
    t ^= t << 11;
    t ^= t >> 8;
    s[0] = t ^ x ^ (x >> 19);
    return s[0];
}

/* random_move.comp:738-741  float(next(state)) * (1.0 / 4294967296.0) */
float orc_rng_float(uint32_t s[4]) {
    float u = (float)orc_rng_next(s);
    return u * 2.3283064365386963e-10f;
}

/* random_move.comp:743-746
 *   uint(ceil(float(min) + (next_float(state) * float(max - min + 1)))) - 1 */
uint32_t orc_rng_range(uint32_t s[4], uint32_t lo, uint32_t hi) {
    float f = orc_rng_float(s);
    float span = (float)(hi - lo + 1u);
    float prod = f * span;
    float sum = (float)lo + prod;
    return (uint32_t)ceilf(sum) - 1u;
}

/* GLSL length(v) = sqrt(dot(v, v)), dot = x*x + y*y. */
float orc_length(float x, float y) {
    float xx = x * x;
    float yy = y * y;
    return sqrtf(xx + yy);
}

/* random_move.comp:830-839  update_direction(index, pos) */
void orc_update_direction(const float tgt[2], const float pos[2], float speed, float dir[2]) {
    float dx = tgt[0] - pos[0];
    float dy = tgt[1] - pos[1];
    float len = orc_length(dx, dy);
    if (len == 0.0f) {
        dir[0] = 0.0f;
        dir[1] = 0.0f;
        return;
    }
    float nx = dx / len;
    float ny = dy / len;
    dir[0] = nx * speed;
    dir[1] = ny * speed;
}

/* random_move.comp:778-828  new_target(index) */
int orc_new_target(float tgt[2], uint32_t* road, uint32_t rng[4],
                   const orc_road* roads, uint32_t n_roads,
                   const uint32_t* conns, uint32_t n_conns) {
    if (*road >= n_roads) {
        return ORC_ERR_ROAD;
    }
    const orc_road* cur = &roads[*road];
    uint32_t idx;
    uint32_t cnt;
    if (tgt[0] == cur->sx && tgt[1] == cur->sy) { /* :783 target == start.pos */
        idx = cur->sidx;
        cnt = cur->scnt;
        if (cnt <= 1) { /* :786-789 dead end: turn around */
            tgt[0] = cur->ex;
            tgt[1] = cur->ey;
            return ORC_OK;
        }
    } else {
        idx = cur->eidx;
        cnt = cur->ecnt;
        if (cnt <= 1) { /* :794-797 */
            tgt[0] = cur->sx;
            tgt[1] = cur->sy;
            return ORC_OK;
        }
    }
    uint32_t off;
    if (cnt == 2) { /* :802-804 the single alternative, no draw */
        off = 1;
    } else {        /* :808 offset in {1..cnt} */
        off = orc_rng_range(rng, 1u, cnt);
    }
    uint64_t at = (uint64_t)idx + off;
    if (at >= n_conns) {
        return ORC_ERR_CONN;
    }
    uint32_t nr = conns[at];
    if (nr >= n_roads) {
        return ORC_ERR_ROAD;
    }
    const orc_road* nw = &roads[nr];
    if (nw->sx == tgt[0] && nw->sy == tgt[1]) { /* :814-819 far endpoint */
        tgt[0] = nw->ex;
        tgt[1] = nw->ey;
    } else {
        tgt[0] = nw->sx;
        tgt[1] = nw->sy;
    }
    *road = nr; /* :820 */
    return ORC_OK;
}

/* random_move.comp:869-874 (update_direction, move) and 841-852 (move). */
int orc_move_agent(float pos[2], float tgt[2], float dir[2], uint32_t* road,
                   uint32_t rng[4], float speed,
                   const orc_road* roads, uint32_t n_roads,
                   const uint32_t* conns, uint32_t n_conns, int* arrived) {
    orc_update_direction(tgt, pos, speed, dir);                /* :870 */
    float dist = orc_length(pos[0] - tgt[0], pos[1] - tgt[1]); /* :842 distance(pos, target) */
    if (dist > speed) {                                        /* :844-846 */
        pos[0] = pos[0] + dir[0];
        pos[1] = pos[1] + dir[1];
        *arrived = 0;
        return ORC_OK;
    }
    float np[2] = {tgt[0], tgt[1]}; /* :848 snap; leftover distance dropped */
    int rc = orc_new_target(tgt, road, rng, roads, n_roads, conns, n_conns);
    if (rc != ORC_OK) {
        return rc;
    }
    orc_update_direction(tgt, np, speed, dir); /* :850 */
    pos[0] = np[0];
    pos[1] = np[1];
    *arrived = 1;
    return ORC_OK;
}

int orc_move_phase(uint32_t n, float* pos, float* tgt, float* dir, uint32_t* road,
                   uint32_t* rng, const float* speed, uint32_t* initialized,
                   const orc_road* roads, uint32_t n_roads,
                   const uint32_t* conns, uint32_t n_conns, uint64_t* arrivals) {
    uint64_t arr = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (initialized && initialized[i] == 0) { /* :863-867 insert only */
            initialized[i] = 1;
            continue;
        }
        int a = 0;
        float s = speed ? speed[i] : ORC_SPEED;
        int rc = orc_move_agent(&pos[2 * i], &tgt[2 * i], &dir[2 * i], &road[i], &rng[4 * i], s,
                                roads, n_roads, conns, n_conns, &a);
        if (rc != ORC_OK) {
            return rc;
        }
        arr += (uint64_t)a;
    }
    if (arrivals) {
        *arrivals = arr;
    }
    return ORC_OK;
}

/* random_move.comp:551-562  quad_tree_in_range(v1, v2, maxDistance),
 * called as in_range(entities[j].pos, ePos_i, R) at :574. */
int orc_in_range(const float pj[2], const float pi[2], float radius) {
    float dx = fabsf(pi[0] - pj[0]);
    if (dx > radius) {
        return 0;
    }
    float dy = fabsf(pi[1] - pj[1]);
    if (dy > radius) {
        return 0;
    }
    return orc_length(pj[0] - pi[0], pj[1] - pi[1]) < radius; /* distance(v1, v2) */
}

float orc_cell_size(float radius) {
    int e = 0;
    float m = frexpf(radius, &e); /* radius = m * 2^e, m in [0.5, 1) */
    if (m == 0.5f) {
        return radius;
    }
    return ldexpf(1.0f, e);
}

static int32_t orc_cell_coord(float v, float cell) {
    float c = floorf(v / cell);
    if (!(c >= -1073741824.0f)) {
        c = -1073741824.0f;
    }
    if (c > 1073741823.0f) {
        c = 1073741823.0f;
    }
    return (int32_t)c;
}

void orc_cells(uint32_t n, const float* pos, float cell, int32_t* cell_xy) {
    for (uint32_t i = 0; i < n; i++) {
        cell_xy[2 * i] = orc_cell_coord(pos[2 * i], cell);
        cell_xy[2 * i + 1] = orc_cell_coord(pos[2 * i + 1], cell);
    }
}

typedef struct {
    uint64_t key;
    uint32_t id;
} orc_binned;

static uint64_t orc_key(int32_t cx, int32_t cy) {
    return ((uint64_t)(uint32_t)(cy + 0x40000000) << 32) | (uint64_t)(uint32_t)(cx + 0x40000000);
}

static int orc_cmp_binned(const void* a, const void* b) {
    const orc_binned* x = (const orc_binned*)a;
    const orc_binned* y = (const orc_binned*)b;
    if (x->key != y->key) {
        return x->key < y->key ? -1 : 1;
    }
    return x->id < y->id ? -1 : (x->id > y->id ? 1 : 0);
}

/* Agents sorted by (cy, cx, id): the restatement's cell lists. */
static orc_binned* orc_bin(uint32_t n, const float* pos, float cell) {
    orc_binned* b = (orc_binned*)malloc(sizeof(orc_binned) * (n ? n : 1));
    if (!b) {
        return NULL;
    }
    for (uint32_t i = 0; i < n; i++) {
        b[i].key = orc_key(orc_cell_coord(pos[2 * i], cell), orc_cell_coord(pos[2 * i + 1], cell));
        b[i].id = i;
    }
    qsort(b, n, sizeof(orc_binned), orc_cmp_binned);
    return b;
}

static uint32_t orc_lower_bound(const orc_binned* b, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (b[mid].key < key) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    return lo;
}

/* One agent's search over the 3x3 cells around it (the quadtree walk at
 * random_move.comp:686-719 restated on a uniform grid whose cell edge is at
 * least R, so every partner lies in the 3x3 block).  Returns the number of
 * counted pairs for which i is the larger index; *mine gets agent i's count. */
static uint64_t orc_search(uint32_t i, uint32_t n, const float* pos, float radius, float cell,
                           const orc_binned* b, const uint32_t* searching, uint32_t* mine) {
    int32_t cx = orc_cell_coord(pos[2 * i], cell);
    int32_t cy = orc_cell_coord(pos[2 * i + 1], cell);
    uint64_t pairs = 0;
    uint32_t cnt = 0;
    for (int32_t dy = -1; dy <= 1; dy++) {
        for (int32_t dx = -1; dx <= 1; dx++) {
            uint64_t key = orc_key(cx + dx, cy + dy);
            for (uint32_t k = orc_lower_bound(b, n, key); k < n && b[k].key == key; k++) {
                uint32_t j = b[k].id;
                if (j == i || !orc_in_range(&pos[2 * j], &pos[2 * i], radius)) {
                    continue;
                }
                if (j < i) {
                    if (!searching || searching[i]) {
                        cnt++;
                        pairs++;
                    }
                } else if (!searching || searching[j]) {
                    cnt++;
                }
            }
        }
    }
    *mine = cnt;
    return pairs;
}

int64_t orc_collide_grid(uint32_t n, const float* pos, float radius,
                         uint32_t* initialized, uint32_t* ncount) {
    float cell = orc_cell_size(radius);
    orc_binned* b = orc_bin(n, pos, cell);
    if (!b) {
        return -1;
    }
    uint64_t pairs = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t c = 0;
        pairs += orc_search(i, n, pos, radius, cell, b, initialized, &c);
        if (ncount) {
            ncount[i] = c;
        }
    }
    free(b);
    if (initialized) {
        for (uint32_t i = 0; i < n; i++) {
            initialized[i] = 1;
        }
    }
    return (int64_t)pairs;
}

int64_t orc_collide_brute(uint32_t n, const float* pos, float radius,
                          uint32_t* initialized, uint32_t* ncount) {
    uint64_t pairs = 0;
    if (ncount) {
        memset(ncount, 0, sizeof(uint32_t) * n);
    }
    for (uint32_t i = 0; i < n; i++) {
        for (uint32_t j = 0; j < i; j++) {
            if (!orc_in_range(&pos[2 * j], &pos[2 * i], radius)) {
                continue;
            }
            if (initialized && !initialized[i]) {
                continue;
            }
            pairs++;
            if (ncount) {
                ncount[i]++;
                ncount[j]++;
            }
        }
    }
    if (initialized) {
        for (uint32_t i = 0; i < n; i++) {
            initialized[i] = 1;
        }
    }
    return (int64_t)pairs;
}

/* ---- threaded variants (CPU baseline) ---------------------------------- */

/* orc_bin on n_threads threads: each thread keys and sorts one slice, then
 * sorted runs are merged pairwise, the merges of a round in parallel.  The
 * order (cy, cx, id) is total, so the result equals orc_bin's. */
typedef struct {
    orc_binned* b;
    uint32_t lo, hi;
    const float* pos;
    float cell;
} orc_sort_job;

static void* orc_sort_worker(void* arg) {
    orc_sort_job* j = (orc_sort_job*)arg;
    for (uint32_t i = j->lo; i < j->hi; i++) {
        j->b[i].key = orc_key(orc_cell_coord(j->pos[2 * i], j->cell), orc_cell_coord(j->pos[2 * i + 1], j->cell));
        j->b[i].id = i;
    }
    qsort(j->b + j->lo, j->hi - j->lo, sizeof(orc_binned), orc_cmp_binned);
    return NULL;
}

typedef struct {
    const orc_binned* src;
    orc_binned* dst;
    uint32_t lo, mid, hi;
} orc_merge_job;

static void* orc_merge_worker(void* arg) {
    orc_merge_job* j = (orc_merge_job*)arg;
    uint32_t a = j->lo, b = j->mid, o = j->lo;
    while (a < j->mid && b < j->hi) {
        j->dst[o++] = (orc_cmp_binned(&j->src[b], &j->src[a]) < 0) ? j->src[b++] : j->src[a++];
    }
    while (a < j->mid) j->dst[o++] = j->src[a++];
    while (b < j->hi) j->dst[o++] = j->src[b++];
    return NULL;
}

static orc_binned* orc_bin_mt(uint32_t n, const float* pos, float cell, int n_threads) {
    if (n_threads <= 1 || n < 4096) {
        return orc_bin(n, pos, cell);
    }
    orc_binned* b = (orc_binned*)malloc(sizeof(orc_binned) * n);
    orc_binned* t = (orc_binned*)malloc(sizeof(orc_binned) * n);
    uint32_t* cut = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n_threads + 1));
    pthread_t* th = (pthread_t*)calloc((size_t)n_threads, sizeof(pthread_t));
    orc_sort_job* sj = (orc_sort_job*)calloc((size_t)n_threads, sizeof(orc_sort_job));
    orc_merge_job* mj = (orc_merge_job*)calloc((size_t)n_threads, sizeof(orc_merge_job));
    if (!b || !t || !cut || !th || !sj || !mj) {
        free(b);
        free(t);
        free(cut);
        free(th);
        free(sj);
        free(mj);
        return NULL;
    }
    for (int k = 0; k <= n_threads; k++) cut[k] = (uint32_t)(((uint64_t)n * k) / n_threads);
    for (int k = 0; k < n_threads; k++) {
        sj[k] = (orc_sort_job){b, cut[k], cut[k + 1], pos, cell};
        pthread_create(&th[k], NULL, orc_sort_worker, &sj[k]);
    }
    for (int k = 0; k < n_threads; k++) pthread_join(th[k], NULL);
    int runs = n_threads;
    while (runs > 1) {  /* merge runs 2i and 2i+1 (cut[] holds the run bounds) */
        int m = 0;
        for (int k = 0; k + 1 < runs; k += 2) {
            mj[m] = (orc_merge_job){b, t, cut[k], cut[k + 1], cut[k + 2]};
            pthread_create(&th[m], NULL, orc_merge_worker, &mj[m]);
            m++;
        }
        if (runs & 1) { /* the odd run out is copied through */
            memcpy(t + cut[runs - 1], b + cut[runs - 1], sizeof(orc_binned) * (cut[runs] - cut[runs - 1]));
        }
        for (int k = 0; k < m; k++) pthread_join(th[k], NULL);
        int nr = 0;
        for (int k = 0; k < runs; k += 2) cut[nr++] = cut[k];
        cut[nr] = n;
        runs = nr;
        orc_binned* x = b;
        b = t;
        t = x;
    }
    free(t);
    free(cut);
    free(th);
    free(sj);
    free(mj);
    return b;
}

typedef struct {
    uint32_t lo, hi, n;
    const float* pos;
    float radius, cell;
    const orc_binned* b;
    uint32_t* ncount;
    uint64_t pairs;
} orc_collide_job;

static void* orc_collide_worker(void* arg) {
    orc_collide_job* j = (orc_collide_job*)arg;
    uint64_t pairs = 0;
    for (uint32_t i = j->lo; i < j->hi; i++) {
        uint32_t c = 0;
        pairs += orc_search(i, j->n, j->pos, j->radius, j->cell, j->b, NULL, &c);
        j->ncount[i] = c;
    }
    j->pairs = pairs;
    return NULL;
}

int64_t orc_collide_grid_mt(uint32_t n, const float* pos, float radius,
                            uint32_t* ncount, int n_threads) {
    if (n_threads < 1) {
        n_threads = 1;
    }
    float cell = orc_cell_size(radius);
    orc_binned* b = orc_bin_mt(n, pos, cell, n_threads);
    if (!b) {
        return -1;
    }
    orc_collide_job* jobs = (orc_collide_job*)calloc((size_t)n_threads, sizeof(orc_collide_job));
    pthread_t* th = (pthread_t*)calloc((size_t)n_threads, sizeof(pthread_t));
    for (int t = 0; t < n_threads; t++) {
        jobs[t].lo = (uint32_t)(((uint64_t)n * t) / n_threads);
        jobs[t].hi = (uint32_t)(((uint64_t)n * (t + 1)) / n_threads);
        jobs[t].n = n;
        jobs[t].pos = pos;
        jobs[t].radius = radius;
        jobs[t].cell = cell;
        jobs[t].b = b;
        jobs[t].ncount = ncount;
        pthread_create(&th[t], NULL, orc_collide_worker, &jobs[t]);
    }
    uint64_t pairs = 0;
    for (int t = 0; t < n_threads; t++) {
        pthread_join(th[t], NULL);
        pairs += jobs[t].pairs;
    }
    free(th);
    free(jobs);
    free(b);
    return (int64_t)pairs;
}

typedef struct {
    uint32_t lo, hi;
    float *pos, *tgt, *dir;
    uint32_t *road, *rng;
    const float* speed;
    const orc_road* roads;
    uint32_t n_roads;
    const uint32_t* conns;
    uint32_t n_conns;
    int rc;
} orc_move_job;

static void* orc_move_worker(void* arg) {
    orc_move_job* j = (orc_move_job*)arg;
    uint32_t n = j->hi - j->lo;
    j->rc = orc_move_phase(n, j->pos + 2 * (size_t)j->lo, j->tgt + 2 * (size_t)j->lo,
                           j->dir + 2 * (size_t)j->lo, j->road + j->lo, j->rng + 4 * (size_t)j->lo,
                           j->speed ? j->speed + j->lo : NULL, NULL, j->roads, j->n_roads,
                           j->conns, j->n_conns, NULL);
    return NULL;
}

int orc_move_phase_mt(uint32_t n, float* pos, float* tgt, float* dir, uint32_t* road,
                      uint32_t* rng, const float* speed,
                      const orc_road* roads, uint32_t n_roads,
                      const uint32_t* conns, uint32_t n_conns, int n_threads) {
    if (n_threads < 1) {
        n_threads = 1;
    }
    orc_move_job* jobs = (orc_move_job*)calloc((size_t)n_threads, sizeof(orc_move_job));
    pthread_t* th = (pthread_t*)calloc((size_t)n_threads, sizeof(pthread_t));
    for (int t = 0; t < n_threads; t++) {
        orc_move_job* j = &jobs[t];
        j->lo = (uint32_t)(((uint64_t)n * t) / n_threads);
        j->hi = (uint32_t)(((uint64_t)n * (t + 1)) / n_threads);
        j->pos = pos;
        j->tgt = tgt;
        j->dir = dir;
        j->road = road;
        j->rng = rng;
        j->speed = speed;
        j->roads = roads;
        j->n_roads = n_roads;
        j->conns = conns;
        j->n_conns = n_conns;
        pthread_create(&th[t], NULL, orc_move_worker, j);
    }
    int rc = ORC_OK;
    for (int t = 0; t < n_threads; t++) {
        pthread_join(th[t], NULL);
        if (jobs[t].rc != ORC_OK && rc == ORC_OK) {
            rc = jobs[t].rc;
        }
    }
    free(th);
    free(jobs);
    return rc;
}

/* ---- binary64 replay ---------------------------------------------------- */

void orc_advance_f64(uint32_t n, const float* pos, const float* tgt,
                     const float* speed, double* pos_out, int32_t* arrive_out) {
    for (uint32_t i = 0; i < n; i++) {
        double px = pos[2 * i], py = pos[2 * i + 1];
        double tx = tgt[2 * i], ty = tgt[2 * i + 1];
        double s = speed ? (double)speed[i] : (double)ORC_SPEED;
        double dx = tx - px, dy = ty - py;
        double len = sqrt(dx * dx + dy * dy);
        if (len > s) {
            pos_out[2 * i] = px + dx / len * s;
            pos_out[2 * i + 1] = py + dy / len * s;
            arrive_out[i] = 0;
        } else {
            pos_out[2 * i] = tx;
            pos_out[2 * i + 1] = ty;
            arrive_out[i] = 1;
        }
    }
}
