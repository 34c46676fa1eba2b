// msim_diff.hip — the differentiable tick on a batched handle and its
// hand-written adjoint.
//
// Differentiable tick (binary32, the move arithmetic of k_advance):
//   speed      s_i = v_i * theta[road_i]          (theta: per-road multiplier)
//   move       arrived: p~_i = T_i (snap; route choice as in the plain tick)
//              else:    p~_i = p_i + ((T_i - p_i) / |T_i - p_i|) * s_i
//   avoidance  p'_i = p~_i + alpha * sum_{j != i, |x_ij| < R} w(|x_ij|^2) x_ij,
//              x_ij = p~_i - p~_j, w(q) = (1 - q / R^2)^3
// over the neighbour relation of the plain collision dispatch (same exact
// predicate, same scenario tables).  theta = 1, alpha = 0 is the reference
// tick.  w, w' and w'' vanish at q = R^2, so the tick is C^2 in the positions.
//
// The adjoint freezes the discrete decisions of the forward run (arrival
// flags, targets, roads, neighbour cells): a forward pass keeps checkpoints
// (agent state and binning) every C ticks; the backward pass replays each
// segment from its checkpoint, recording per tick the pre-move positions,
// targets, roads, flags and the binning, then walks the segment backwards:
//   avoidance  lam~_k = lam_k + alpha * sum_j J(x_kj) (lam_k - lam_j),
//              J(x) = w(q) I + 2 w'(q) x x^T,   dL/dalpha += lam_k . A_k
//   move       lam_k = lam~_k - (s/L) (lam~_k - u (u . lam~_k)), u = (T-p)/L
//              dL/dtheta[road_k] += v_k (u . lam~_k);  arrived: lam_k = 0
// The reverse pass runs in binary64 on the run's binary32 values.
#include "msim_batch.hpp"

#include <cmath>
#include <cstring>

using namespace msim;

namespace msim {

enum DiffFlag : uint8_t { kDiffArrived = 1, kDiffNoMove = 2 };

struct DiffTapeDev {
    float2* pos;     // [n] position before the move
    float2* tgt;     // [n] target of the move
    uint32_t* road;  // [n] road of the move (theta index)
    uint8_t* flag;   // [n] DiffFlag bits
};

// Forward move of the differentiable tick; REPLAY takes the decisions from
// `in` instead of making them (no route choice, agent state untouched but
// for pos).  `rec` (nullable) receives the tick's decisions.
template <bool REPLAY, bool UNIFORM>
static __global__ __launch_bounds__(kBlock) void k_dadvance(AgentsDev a, const float* __restrict__ theta,
                                                     uint32_t no_move, DiffTapeDev in, DiffTapeDev rec,
                                                     const msim_road* __restrict__ roads,
                                                     const uint32_t* __restrict__ conns) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= a.n) {
        return;
    }
    const float2 p = a.pos[k];
    const float2 T = REPLAY ? in.tgt[k] : a.tgt[k];
    const uint32_t r = REPLAY ? in.road[k] : a.road[k];
    const float s = (UNIFORM ? a.spd_const : a.spd[k]) * theta[r];
    uint8_t flag = 0;
    float2 np = p;
    if (no_move) {
        flag = kDiffNoMove;
    } else {
        const float dx = T.x - p.x;
        const float dy = T.y - p.y;
        const float len = vlen(dx, dy);
        const bool arrived = REPLAY ? (in.flag[k] & kDiffArrived) != 0 : !(len > s);
        if (!arrived) {
            np = make_float2(p.x + (dx / len) * s, p.y + (dy / len) * s);
        } else {
            np = T;
            flag = kDiffArrived;
            if (!REPLAY) {
                float2 nt = T;
                uint32_t nr = r;
                uint4 st = a.rng[k];
                choose_route(nt, nr, st, roads, conns);
                a.tgt[k] = nt;
                a.road[k] = nr;
                a.rng[k] = st;
            }
        }
    }
    a.pos[k] = np;
    if (rec.pos) {
        rec.pos[k] = p;
        rec.tgt[k] = T;
        rec.road[k] = r;
        rec.flag[k] = flag;
    }
}

__device__ __forceinline__ float avoid_w(float q, float inv_r2) {
    const float h = 1.0f - q * inv_r2;
    return h * h * h;
}

struct DiffCells {
    const BlockScen* blk;
    const ScenDev* scen;
    const uint32_t* agent_off;
    uint32_t n;
    const uint32_t* start;
    const float2* bpos;
    const uint32_t* bid;
    float inv_cell;
    float d2_limit;
    float inv_r2;
    double inv_r2_d;
    float fix_scale;   // forward avoidance: terms summed as integers in units of 1/fix_scale
    float push_allow;  // largest push per axis the rebinning windows were sized for
};

// Calls f(slot) for every other slot within the d2 limit of slot s, over the
// 3x3 cells of s's scenario table.  The order of the slots inside a bucket
// follows the rebinning's LDS atomics and is not reproducible: callers sum in
// an order-independent way.
template <class F>
__device__ __forceinline__ void for_neighbours(const DiffCells& c, uint32_t s, float2 me, F&& f) {
    const BlockScen bs = c.blk[s / kBlock];
    uint32_t sid = bs.sid;
    uint32_t base = bs.start_base, lw = bs.lw, wmask = bs.wmask, hmask = bs.hmask;
    if (s >= bs.end) {
        sid++;
        while (s >= c.agent_off[sid + 1]) sid++;
        const ScenDev& sc = c.scen[sid];
        base = sc.start_base;
        lw = sc.lw;
        wmask = sc.wmask;
        hmask = sc.hmask;
    }
    const uint32_t* start = c.start + base;
    const int32_t cx = cell_coord(me.x, c.inv_cell);
    const int32_t cy = cell_coord(me.y, c.inv_cell);
    const uint32_t c0 = static_cast<uint32_t>(cx - 1) & wmask;
    const uint32_t c2 = static_cast<uint32_t>(cx + 1) & wmask;
    const bool wrap = c0 > c2;
#pragma unroll 1
    for (int r = -1; r <= 1; r++) {
        const uint32_t rowb = (static_cast<uint32_t>(cy + r) & hmask) << lw;
        const uint32_t lo = start[rowb + c0];
        const uint32_t hi = start[rowb + (wrap ? wmask + 1u : c2 + 1u)];
        for (uint32_t k = lo; k < hi; k++) {
            if (k != s && in_range(c.bpos[k], me, c.d2_limit)) f(k);
        }
        if (wrap) {
            const uint32_t hi2 = start[rowb + c2 + 1u];
            for (uint32_t k = start[rowb]; k < hi2; k++) {
                if (k != s && in_range(c.bpos[k], me, c.d2_limit)) f(k);
            }
        }
    }
}

// Forward avoidance: pos[agent] = p~ + alpha * A over the tick's binning.
// Each binary32 term w x is rounded to a multiple of 1/fix_scale (a power of
// two with |w x| * fix_scale < 2^40, at most 2^-40 R) and the terms are
// summed as 64-bit integers, so A does not depend on the order in which the
// neighbours are visited and two runs with the same parameters are
// bit-identical.  A push beyond the windows' allowance raises kFaultBinCount
// (the run then fails instead of rebinning an agent outside its window).
static __global__ __launch_bounds__(kBlock) void k_davoid(DiffCells c, float alpha, float2* __restrict__ pos,
                                                   unsigned int* __restrict__ faults) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= c.n) {
        return;
    }
    const float2 me = c.bpos[s];
    long long ix = 0, iy = 0;
    const double sc = c.fix_scale;
    for_neighbours(c, s, me, [&](uint32_t k) {
        const float2 q = c.bpos[k];
        const float x = me.x - q.x;
        const float y = me.y - q.y;
        const float w = avoid_w(x * x + y * y, c.inv_r2);
        ix += __double2ll_rn(static_cast<double>(w * x) * sc);
        iy += __double2ll_rn(static_cast<double>(w * y) * sc);
    });
    const float ax = static_cast<float>(static_cast<double>(ix) / sc);
    const float ay = static_cast<float>(static_cast<double>(iy) / sc);
    const float px = alpha * ax, py = alpha * ay;
    if (!(fabsf(px) <= c.push_allow) || !(fabsf(py) <= c.push_allow)) {
        atomicOr(faults, kFaultBinCount);
    }
    pos[c.bid[s]] = make_float2(me.x + px, me.y + py);
}

// Avoidance adjoint over the tick's binning, in binary64 from the run's
// binary32 positions: lam_out = lam + alpha * sum J dl, and the alpha
// derivative sum lam . A (one atomic per block).
static __global__ __launch_bounds__(kBlock) void k_davoid_adj(DiffCells c, double alpha,
                                                       const double2* __restrict__ lam,
                                                       double2* __restrict__ lam_out,
                                                       double* __restrict__ galpha) {
    __shared__ double red[kBlock / 64];
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    double ga = 0.0;
    if (s < c.n) {
        const float2 me = c.bpos[s];
        const uint32_t i = c.bid[s];
        const double2 li = lam[i];
        const double inv_r2 = c.inv_r2_d;
        double gx = 0.0, gy = 0.0, ax = 0.0, ay = 0.0;
        for_neighbours(c, s, me, [&](uint32_t k) {
            const float2 q = c.bpos[k];
            const double x = static_cast<double>(me.x) - q.x;
            const double y = static_cast<double>(me.y) - q.y;
            const double h = 1.0 - (x * x + y * y) * inv_r2;
            const double w = h * h * h;
            const double wp = -3.0 * inv_r2 * h * h;  // dw/dq
            const double2 lj = lam[c.bid[k]];
            const double dlx = li.x - lj.x;
            const double dly = li.y - lj.y;
            const double xd = 2.0 * wp * (x * dlx + y * dly);
            gx += w * dlx + xd * x;
            gy += w * dly + xd * y;
            ax += w * x;
            ay += w * y;
        });
        lam_out[i] = make_double2(li.x + alpha * gx, li.y + alpha * gy);
        ga = li.x * ax + li.y * ay;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ga += __shfl_xor(ga, o, 64);
    }
    if (__lane_id() == 0) red[threadIdx.x >> 6] = ga;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (uint32_t k = 0; k < kBlock / 64; k++) t += red[k];
        if (t != 0.0) atomicAdd(galpha, t);
    }
}

// Move adjoint in agent order, binary64 from the run's binary32 values.
// An arrival's snap is constant under frozen decisions (λ = 0).  With carry,
// it is differentiated as if the arrival carried its leftover time onto the
// next road (nxt_tgt, nxt_road: the decisions after this tick),
// p' = T + (1 - L/s) s2 u2 with L = |T - p|, u2 the next road's direction and
// s, s2 the two roads' speeds: λ = (s2/s) u (u2·λ'), ∂/∂θ[road] += v L s2/s²
// (u2·λ'), ∂/∂θ[next road] += v (1 - L/s)(u2·λ').
template <bool UNIFORM>
static __global__ __launch_bounds__(kBlock) void k_dmove_adj(uint32_t n, const float* __restrict__ spd,
                                                      float spd_const, const float* __restrict__ theta,
                                                      DiffTapeDev tp, const float2* __restrict__ nxt_tgt,
                                                      const uint32_t* __restrict__ nxt_road, uint32_t carry,
                                                      const double2* __restrict__ lam_in,
                                                      double2* __restrict__ lam_out,
                                                      double* __restrict__ gtheta) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) {
        return;
    }
    const double2 l = lam_in[k];
    const uint8_t flag = tp.flag[k];
    if (flag & kDiffArrived) {
        double2 out = make_double2(0.0, 0.0);
        if (carry) {
            const float2 p = tp.pos[k];
            const float2 T = tp.tgt[k];
            const float2 T2 = nxt_tgt[k];
            const double ex = static_cast<double>(T2.x) - T.x;
            const double ey = static_cast<double>(T2.y) - T.y;
            const double E = sqrt(ex * ex + ey * ey);
            const uint32_t r = tp.road[k];
            const uint32_t r2 = nxt_road[k];
            const float v = UNIFORM ? spd_const : spd[k];
            const double s = static_cast<double>(v * theta[r]);
            const double s2 = static_cast<double>(v * theta[r2]);
            // a zero speed (theta[r] == 0) carries nothing: the frozen snap
            if (E > 0.0 && s > 0.0) {
                const double dx = static_cast<double>(T.x) - p.x;
                const double dy = static_cast<double>(T.y) - p.y;
                const double L = sqrt(dx * dx + dy * dy);
                const double ux = L > 0.0 ? dx / L : 0.0;
                const double uy = L > 0.0 ? dy / L : 0.0;
                const double g2 = (ex / E) * l.x + (ey / E) * l.y;
                out = make_double2(ux * (s2 / s) * g2, uy * (s2 / s) * g2);
                atomicAdd(&gtheta[r], static_cast<double>(v) * (L * s2 / (s * s)) * g2);
                atomicAdd(&gtheta[r2], static_cast<double>(v) * (1.0 - L / s) * g2);
            }
        }
        lam_out[k] = out;
        return;
    }
    if (flag & kDiffNoMove) {
        lam_out[k] = l;
        return;
    }
    const float2 p = tp.pos[k];
    const float2 T = tp.tgt[k];
    const uint32_t r = tp.road[k];
    const float v = UNIFORM ? spd_const : spd[k];
    const double s = static_cast<double>(v * theta[r]);  // the run's binary32 speed
    const double dx = static_cast<double>(T.x) - p.x;
    const double dy = static_cast<double>(T.y) - p.y;
    const double len = sqrt(dx * dx + dy * dy);
    const double ux = dx / len;
    const double uy = dy / len;
    const double ul = ux * l.x + uy * l.y;
    const double f = s / len;
    lam_out[k] = make_double2(l.x - f * (l.x - ux * ul), l.y - f * (l.y - uy * ul));
    atomicAdd(&gtheta[r], static_cast<double>(v) * ul);
}

// Caller order <-> internal order (perm: internal slot -> caller index).
static __global__ __launch_bounds__(kBlock) void k_dscatter(uint32_t n, const uint32_t* __restrict__ perm,
                                                     const float2* __restrict__ in, float2* __restrict__ out) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k < n) out[perm[k]] = in[k];
}

static __global__ __launch_bounds__(kBlock) void k_dgather(uint32_t n, const uint32_t* __restrict__ perm,
                                                    const float2* __restrict__ in, double2* __restrict__ out) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k < n) {
        const float2 v = in[perm[k]];
        out[k] = make_double2(v.x, v.y);
    }
}

}  // namespace msim

// ---------------------------------------------------------------------------

struct DiffSnapshot {
    uint32_t* perm = nullptr;  // initial snapshot only: agent order and speeds
    float* spd = nullptr;
    float2* pos = nullptr;
    float2* tgt = nullptr;
    uint32_t* road = nullptr;
    uint4* rng = nullptr;
    uint32_t* start = nullptr;
    uint32_t* bid = nullptr;
    float2* bpos = nullptr;
};

struct DiffState {
    uint32_t T = 0, C = 50, n_ck = 0;
    bool keep_schedule = false;
    bool no_move_first = false;
    bool carry = false;  // arrivals differentiated as carrying their leftover time
    bool have_forward = false;
    float alpha = 0.0f;      // parameters of the last recorded forward run
    float theta_max = 1.0f;
    float* theta = nullptr;  // [n_roads]
    float alpha_rep = 0.0f;  // parameters of the last replay
    float* theta_rep = nullptr;
    DiffSnapshot init;
    std::vector<DiffSnapshot> ck;
    // per-tick tape of one segment (or of the whole run with keep_schedule)
    DiffTapeDev tape{};      // [C][n] each
    uint32_t* tp_start = nullptr;  // [C][start_len]
    uint32_t* tp_bid = nullptr;    // [C][n]
    float2* tp_bpos = nullptr;     // [C][n + 2]
    DiffTapeDev sched{};     // [T + 1][n] decisions of the whole run and those after it (keep_schedule)
    double2* lam = nullptr;  // adjoint state, agent order
    double2* lam2 = nullptr;
    double* gtheta = nullptr;
    double* galpha = nullptr;
    ScenDev* scen = nullptr;  // scenario tables with the differentiable tick's margins
    std::vector<ScenDev> scen_h;
    float smax = 0.0f;  // largest base speed
    uint32_t max_block = 1;  // most agents in one 3x3 cell block of the initial binning
    float push_allow = 0.0f;  // avoidance push per axis the window margins allow for
    uint64_t generation = 0;  // recorded (non-replay) forward runs so far
};

namespace {

DiffTapeDev tape_at(const DiffTapeDev& t, size_t n, uint32_t k) {
    return DiffTapeDev{t.pos + n * k, t.tgt + n * k, t.road + n * k, t.flag + n * k};
}

int alloc_snapshot(msim_batch* b, DiffSnapshot& s) {
    int rc = balloc_n(b, &s.pos, b->n);
    if (rc == MSIM_OK) rc = balloc_n(b, &s.tgt, b->n);
    if (rc == MSIM_OK) rc = balloc_n(b, &s.road, b->n);
    if (rc == MSIM_OK) rc = balloc_n(b, &s.rng, b->n);
    if (rc == MSIM_OK) rc = balloc_n(b, &s.start, b->start_len);
    if (rc == MSIM_OK) rc = balloc_n(b, &s.bid, b->n);
    if (rc == MSIM_OK) rc = balloc_n(b, &s.bpos, static_cast<size_t>(b->n) + 2);
    return rc;
}

template <class T>
void dcopy(msim_batch* b, T* dst, const T* src, size_t count) {
    if (count) (void)hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToDevice, b->stream);
}

// Working agent state (buffer cur) and the current binning (bcur) <-> snapshot.
void save_snapshot(msim_batch* b, DiffSnapshot& s) {
    const BBuffers& w = b->buf[b->cur];
    dcopy(b, s.pos, w.pos, b->n);
    dcopy(b, s.tgt, w.tgt, b->n);
    dcopy(b, s.road, w.road, b->n);
    dcopy(b, s.rng, w.rng, b->n);
    dcopy(b, s.start, b->bstart[b->bcur], b->start_len);
    dcopy(b, s.bid, b->bbid[b->bcur], b->n);
    dcopy(b, s.bpos, b->bbpos[b->bcur], b->n);
}

void load_agents(msim_batch* b, const DiffSnapshot& s) {
    const BBuffers& w = b->buf[b->cur];
    dcopy(b, w.pos, s.pos, b->n);
    dcopy(b, w.tgt, s.tgt, b->n);
    dcopy(b, w.road, s.road, b->n);
    dcopy(b, w.rng, s.rng, b->n);
}

// The initial state: agent order, speeds, agent state and binning.
void load_initial(msim_batch* b, const DiffSnapshot& s) {
    const BBuffers& w = b->buf[b->cur];
    dcopy(b, w.perm, s.perm, b->n);
    if (!b->uniform) dcopy(b, w.spd, s.spd, b->n);
    load_agents(b, s);
    dcopy(b, b->bstart[b->bcur], s.start, b->start_len);
    dcopy(b, b->bbid[b->bcur], s.bid, b->n);
    dcopy(b, b->bbpos[b->bcur], s.bpos, b->n);
}

// Margins of the rebinning windows for the largest displacement of a
// differentiable tick: speed * theta_max plus the avoidance push.  Per axis
// the push is |alpha| sum_j w x_j <= |alpha| k max_u (1 - u^2)^3 u R
// (u = |x| / R; the maximum, at u^2 = 1/7, is below 0.24) for k neighbours;
// the margin allows for k = the most agents found in one 3x3 block of cells
// at set-up, and at least 2R.  A larger push (agents bunched tighter later)
// is caught on the device (kFaultBinCount, the run fails).
int diff_margins(msim_batch* b, DiffState& d, float alpha) {
    const float cell = 1.0f / b->inv_cell;
    const double push = std::max(2.0 * b->radius, 0.24 * std::fabs(static_cast<double>(alpha)) * d.max_block *
                                                      static_cast<double>(b->radius));
    d.push_allow = static_cast<float>(push);
    const double disp = static_cast<double>(d.smax) * d.theta_max + push;
    const uint32_t D = static_cast<uint32_t>(std::floor(disp * (1.0 + 1.0 / 1048576.0) / cell)) + 1u;
    for (ScenDev& s : d.scen_h) s.margin = (D + 1u) * (s.wmask + 1u) + D;
    hipError_t e = hipMemcpy(d.scen, d.scen_h.data(), sizeof(ScenDev) * d.scen_h.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bfail(b, MSIM_ERR_HIP, "margin upload failed: %s", hipGetErrorString(e));
    return MSIM_OK;
}

// Rebinning of the working positions: old binning -> new binning.
void dbins(msim_batch* b, const DiffState& d, const uint32_t* old_start, const uint32_t* old_bid,
           uint32_t* new_start, float2* new_bpos, uint32_t* new_bid) {
    BatchBin bb{};
    bb.scen = d.scen;
    bb.seg_scen = b->seg_scen;
    bb.old_start = old_start;
    bb.old_bid = old_bid;
    bb.new_start = new_start;
    bb.new_bpos = new_bpos;
    bb.new_bid = new_bid;
    bb.seg_total = b->seg_total;
    bb.faults = b->faults;
    bb.inv_cell = b->inv_cell;
    const size_t lds = sizeof(uint32_t) << b->max_seg_log;
    const float2* pos = b->buf[b->cur].pos;
    hipLaunchKernelGGL((k_bbin_count<4, kBinThreads>), dim3(b->n_seg), dim3(kBinThreads), lds, b->stream, pos, bb);
    hipLaunchKernelGGL((k_bbin_place<4, kBinThreads>), dim3(b->n_seg), dim3(kBinThreads), lds, b->stream, pos, bb);
}

DiffCells dcells(msim_batch* b, const uint32_t* start, const float2* bpos, const uint32_t* bid) {
    DiffCells c{};
    c.blk = b->blk_scen;
    c.scen = b->scen;
    c.agent_off = b->agent_off;
    c.n = b->n;
    c.start = start;
    c.bpos = bpos;
    c.bid = bid;
    c.inv_cell = b->inv_cell;
    c.d2_limit = b->d2_limit;
    c.inv_r2 = 1.0f / (b->radius * b->radius);
    c.inv_r2_d = 1.0 / (static_cast<double>(b->radius) * b->radius);
    // |w x| < R: terms below 2^40 units, sums of up to 2^22 terms in range
    c.fix_scale = std::ldexp(1.0f, 40 - std::max(0, static_cast<int>(std::ceil(std::log2(b->radius)))));
    c.push_allow = b->diff ? b->diff->push_allow : 0.0f;
    return c;
}

// One differentiable tick t on the working state: move (decide or replay),
// rebin into (start, bpos, bid), avoidance.
void dtick(msim_batch* b, DiffState& d, const float* theta, float alpha, uint32_t t, bool replay, DiffTapeDev rec,
           const uint32_t* old_start, const uint32_t* old_bid, uint32_t* new_start, float2* new_bpos,
           uint32_t* new_bid) {
    const AgentsDev a = b->agents(b->cur);
    const uint32_t no_move = (t == 0 && d.no_move_first) ? 1u : 0u;
    const DiffTapeDev in = replay ? tape_at(d.sched, b->n, t) : DiffTapeDev{};
    const dim3 g(bblocks(b->n));
    if (replay) {
        if (b->uniform) {
            hipLaunchKernelGGL((k_dadvance<true, true>), g, dim3(kBlock), 0, b->stream, a, theta, no_move, in, rec,
                               b->roads, b->conns);
        } else {
            hipLaunchKernelGGL((k_dadvance<true, false>), g, dim3(kBlock), 0, b->stream, a, theta, no_move, in,
                               rec, b->roads, b->conns);
        }
    } else {
        if (b->uniform) {
            hipLaunchKernelGGL((k_dadvance<false, true>), g, dim3(kBlock), 0, b->stream, a, theta, no_move, in,
                               rec, b->roads, b->conns);
        } else {
            hipLaunchKernelGGL((k_dadvance<false, false>), g, dim3(kBlock), 0, b->stream, a, theta, no_move, in,
                               rec, b->roads, b->conns);
        }
    }
    dbins(b, d, old_start, old_bid, new_start, new_bpos, new_bid);
    if (alpha != 0.0f) {
        hipLaunchKernelGGL(k_davoid, g, dim3(kBlock), 0, b->stream, dcells(b, new_start, new_bpos, new_bid), alpha,
                           a.pos, b->faults);
    }
}

int check_diff_faults(msim_batch* b) {
    unsigned int f = 0;
    hipError_t e = hipMemcpyAsync(&f, b->faults, sizeof(f), hipMemcpyDeviceToHost, b->stream);
    if (e != hipSuccess) return bfail(b, MSIM_ERR_HIP, "fault readback: %s", hipGetErrorString(e));
    int rc = bwait(b);
    if (rc != MSIM_OK) return rc;
    if (f & kFaultBinCount) {
        return bfail(b, MSIM_ERR_STATE,
                     "an agent moved beyond its rebinning window in a differentiable tick (theta or alpha too "
                     "large for the window margins)");
    }
    if (f) return bfail(b, MSIM_ERR_STATE, "device-side check failed (fault bits 0x%x)", f);
    return MSIM_OK;
}

}  // namespace

void diff_release(msim_batch* b) {
    delete b->diff;  // device memory is in b->allocs
    b->diff = nullptr;
}

extern "C" int msim_batch_diff_setup(msim_batch* b, uint32_t n_ticks, uint32_t checkpoint_every,
                                     int keep_schedule) {
    if (!b) return bfail(b, MSIM_ERR_INVALID, "NULL handle");
    if (b->diff) return bfail(b, MSIM_ERR_STATE, "the differentiable tick is already set up on this handle");
    if (n_ticks == 0 || checkpoint_every == 0) return bfail(b, MSIM_ERR_INVALID, "n_ticks and checkpoint_every must be > 0");
    int rc = bleave_sorted(b);  // the differentiable tick runs on the classic layout
    if (rc == MSIM_OK) rc = msim_batch_sync(b);
    if (rc != MSIM_OK) return rc;
    DiffState* d = new DiffState();
    b->diff = d;
    d->T = n_ticks;
    d->C = std::min(checkpoint_every, n_ticks);
    d->n_ck = (n_ticks + d->C - 1) / d->C;
    d->keep_schedule = keep_schedule != 0;
    d->no_move_first = b->first_tick_insert_only;
    const size_t n = b->n;
    rc = balloc_n(b, &d->theta, b->n_roads);
    if (rc == MSIM_OK) rc = balloc_n(b, &d->theta_rep, b->n_roads);
    if (rc == MSIM_OK) rc = alloc_snapshot(b, d->init);
    if (rc == MSIM_OK) rc = balloc_n(b, &d->init.perm, n);
    if (rc == MSIM_OK && !b->uniform) rc = balloc_n(b, &d->init.spd, n);
    d->ck.resize(d->n_ck);
    for (uint32_t k = 0; k < d->n_ck && rc == MSIM_OK; k++) rc = alloc_snapshot(b, d->ck[k]);
    if (rc == MSIM_OK) rc = balloc_n(b, &d->tape.pos, n * d->C);
    if (rc == MSIM_OK) rc = balloc_n(b, &d->tape.tgt, n * d->C);
    if (rc == MSIM_OK) rc = balloc_n(b, &d->tape.road, n * d->C);
    if (rc == MSIM_OK) rc = balloc_n(b, &d->tape.flag, n * d->C);
    if (rc == MSIM_OK) rc = balloc_n(b, &d->tp_start, static_cast<size_t>(b->start_len) * d->C);
    if (rc == MSIM_OK) rc = balloc_n(b, &d->tp_bid, n * d->C);
    if (rc == MSIM_OK) rc = balloc_n(b, &d->tp_bpos, (n + 2) * d->C);
    if (rc == MSIM_OK && d->keep_schedule) {
        rc = balloc_n(b, &d->sched.pos, n * (n_ticks + 1u));
        if (rc == MSIM_OK) rc = balloc_n(b, &d->sched.tgt, n * (n_ticks + 1u));
        if (rc == MSIM_OK) rc = balloc_n(b, &d->sched.road, n * (n_ticks + 1u));
        if (rc == MSIM_OK) rc = balloc_n(b, &d->sched.flag, n * (n_ticks + 1u));
    }
    if (rc == MSIM_OK) rc = balloc_n(b, &d->lam, n);
    if (rc == MSIM_OK) rc = balloc_n(b, &d->lam2, n);
    if (rc == MSIM_OK) rc = balloc_n(b, &d->gtheta, b->n_roads);
    if (rc == MSIM_OK) rc = balloc_n(b, &d->galpha, 1);
    if (rc == MSIM_OK) rc = balloc_n(b, &d->scen, b->n_scen);
    if (rc == MSIM_OK) {
        d->scen_h.resize(b->n_scen);
        hipError_t e = hipMemcpy(d->scen_h.data(), b->scen, sizeof(ScenDev) * b->n_scen, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = bfail(b, MSIM_ERR_HIP, "scenario table readback: %s", hipGetErrorString(e));
    }
    if (rc == MSIM_OK && b->start_len) {
        // the push allowance: at most 9 x the fullest bucket in a 3x3 block
        std::vector<uint32_t> st(b->start_len);
        hipError_t e = hipMemcpy(st.data(), b->bstart[b->bcur], sizeof(uint32_t) * b->start_len,
                                 hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = bfail(b, MSIM_ERR_HIP, "binning readback: %s", hipGetErrorString(e));
        uint32_t most = 1;
        for (const ScenDev& sc : d->scen_h) {
            for (uint32_t j = 0; j < sc.m && rc == MSIM_OK; j++) {
                most = std::max(most, st[sc.start_base + j + 1] - st[sc.start_base + j]);
            }
        }
        d->max_block = 9u * most;
    }
    if (rc == MSIM_OK) {
        d->smax = kShaderSpeed;
        if (!b->uniform) {
            std::vector<float> v(n);
            hipError_t e = n ? hipMemcpy(v.data(), b->buf[b->cur].spd, sizeof(float) * n, hipMemcpyDeviceToHost)
                             : hipSuccess;
            if (e != hipSuccess) rc = bfail(b, MSIM_ERR_HIP, "speed readback: %s", hipGetErrorString(e));
            d->smax = 0.0f;
            for (float s : v) d->smax = std::max(d->smax, s);
        }
    }
    if (rc == MSIM_OK) {
        save_snapshot(b, d->init);
        dcopy(b, d->init.perm, b->buf[b->cur].perm, n);
        if (!b->uniform) dcopy(b, d->init.spd, b->buf[b->cur].spd, n);
        rc = bwait(b);
    }
    if (rc != MSIM_OK) diff_release(b);
    return rc;
}

namespace {

// Parameters: theta (device, n_roads packed, >= 0 and finite) into dst,
// alpha; the window margins follow the largest multiplier.
int set_params(msim_batch* b, DiffState& d, const float* theta, float alpha, float* dst) {
    if (!theta) return bfail(b, MSIM_ERR_INVALID, "NULL theta");
    if (!std::isfinite(alpha)) return bfail(b, MSIM_ERR_INVALID, "alpha must be finite");
    std::vector<float> th(b->n_roads);
    if (b->n_roads) {
        hipError_t e = hipMemcpy(th.data(), theta, sizeof(float) * b->n_roads, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return bfail(b, MSIM_ERR_HIP, "theta readback: %s", hipGetErrorString(e));
    }
    float mx = 0.0f;
    for (uint32_t r = 0; r < b->n_roads; r++) {
        if (!(th[r] >= 0.0f) || !std::isfinite(th[r])) {
            return bfail(b, MSIM_ERR_INVALID, "theta[%u] = %g is not a finite non-negative multiplier", r,
                         static_cast<double>(th[r]));
        }
        mx = std::max(mx, th[r]);
    }
    hipError_t e = b->n_roads ? hipMemcpy(dst, th.data(), sizeof(float) * b->n_roads, hipMemcpyHostToDevice)
                              : hipSuccess;
    if (e != hipSuccess) return bfail(b, MSIM_ERR_HIP, "theta upload: %s", hipGetErrorString(e));
    d.theta_max = mx;
    int rc = diff_margins(b, d, alpha);
    // the uploads ran on the null stream, which the handle's stream does not wait for
    if (rc == MSIM_OK && hipStreamSynchronize(nullptr) != hipSuccess) {
        rc = bfail(b, MSIM_ERR_HIP, "parameter upload failed");
    }
    return rc;
}

}  // namespace

extern "C" int msim_batch_diff_forward(msim_batch* b, const float* theta, float alpha, int replay,
                                       float* final_pos) {
    if (!b || !b->diff) return bfail(b, MSIM_ERR_STATE, "msim_batch_diff_setup first");
    DiffState& d = *b->diff;
    if (replay && !(d.keep_schedule && d.have_forward)) {
        return bfail(b, MSIM_ERR_STATE, "replay needs keep_schedule and a recorded forward run");
    }
    int rc = msim_batch_sync(b);
    if (rc != MSIM_OK) return rc;
    float* th = replay ? d.theta_rep : d.theta;
    rc = set_params(b, d, theta, alpha, th);
    if (rc != MSIM_OK) return rc;
    if (replay) {
        d.alpha_rep = alpha;
    } else {
        d.alpha = alpha;
        d.have_forward = false;
    }
    load_initial(b, d.init);
    for (uint32_t t = 0; t < d.T && rc == MSIM_OK; t++) {
        if (!replay && t % d.C == 0) save_snapshot(b, d.ck[t / d.C]);
        const DiffTapeDev rec = (!replay && d.keep_schedule) ? tape_at(d.sched, b->n, t) : DiffTapeDev{};
        const int nb = 1 - b->bcur;
        dtick(b, d, th, alpha, t, replay != 0, rec, b->bstart[b->bcur], b->bbid[b->bcur], b->bstart[nb],
              b->bbpos[nb], b->bbid[nb]);
        b->bcur = nb;
        rc = bcheck(b, "differentiable tick");
    }
    if (rc == MSIM_OK && !replay && d.keep_schedule && b->n) {
        // the decisions in force after the last tick (schedule entry T)
        const DiffTapeDev fin = tape_at(d.sched, b->n, d.T);
        const AgentsDev a = b->agents(b->cur);
        hipError_t e = hipMemcpyAsync(fin.pos, a.pos, sizeof(float2) * b->n, hipMemcpyDeviceToDevice, b->stream);
        if (e == hipSuccess) {
            e = hipMemcpyAsync(fin.tgt, a.tgt, sizeof(float2) * b->n, hipMemcpyDeviceToDevice, b->stream);
        }
        if (e == hipSuccess) {
            e = hipMemcpyAsync(fin.road, a.road, sizeof(uint32_t) * b->n, hipMemcpyDeviceToDevice, b->stream);
        }
        if (e == hipSuccess) e = hipMemsetAsync(fin.flag, 0, b->n, b->stream);
        if (e != hipSuccess) rc = bfail(b, MSIM_ERR_HIP, "schedule copy: %s", hipGetErrorString(e));
    }
    if (rc == MSIM_OK && final_pos) {
        hipLaunchKernelGGL(k_dscatter, dim3(bblocks(b->n)), dim3(kBlock), 0, b->stream, b->n, b->buf[b->cur].perm,
                           b->buf[b->cur].pos, reinterpret_cast<float2*>(final_pos));
        rc = bcheck(b, "k_dscatter");
    }
    if (rc == MSIM_OK) rc = check_diff_faults(b);
    if (rc == MSIM_OK && !replay) {
        d.have_forward = true;
        d.generation++;
    }
    return rc;
}

extern "C" int msim_batch_diff_generation(msim_batch* b, uint64_t* out) {
    if (!b || !b->diff || !out) return bfail(b, MSIM_ERR_STATE, "msim_batch_diff_setup first (or NULL argument)");
    *out = b->diff->generation;
    return MSIM_OK;
}

extern "C" int msim_batch_diff_backward(msim_batch* b, const float* grad_pos, double* grad_theta,
                                        double* grad_alpha) {
    if (!b || !b->diff) return bfail(b, MSIM_ERR_STATE, "msim_batch_diff_setup first");
    DiffState& d = *b->diff;
    if (!d.have_forward) return bfail(b, MSIM_ERR_STATE, "run msim_batch_diff_forward (not a replay) first");
    if (!grad_pos || !grad_theta || !grad_alpha) return bfail(b, MSIM_ERR_INVALID, "NULL argument");
    int rc = msim_batch_sync(b);
    if (rc != MSIM_OK) return rc;
    rc = set_params(b, d, d.theta, d.alpha, d.theta);  // margins of the recorded run
    if (rc != MSIM_OK) return rc;
    const size_t n = b->n;
    const dim3 g(bblocks(b->n));
    const uint32_t* perm = b->buf[b->cur].perm;
    hipLaunchKernelGGL(k_dgather, g, dim3(kBlock), 0, b->stream, b->n, perm,
                       reinterpret_cast<const float2*>(grad_pos), d.lam);
    (void)hipMemsetAsync(d.gtheta, 0, sizeof(double) * std::max(1u, b->n_roads), b->stream);
    (void)hipMemsetAsync(d.galpha, 0, sizeof(double), b->stream);
    for (int seg = static_cast<int>(d.n_ck) - 1; seg >= 0 && rc == MSIM_OK; seg--) {
        const uint32_t t0 = static_cast<uint32_t>(seg) * d.C;
        const uint32_t t1 = std::min(d.T, t0 + d.C);
        // replay the segment from its checkpoint, recording the tape
        load_agents(b, d.ck[seg]);
        for (uint32_t t = t0; t < t1; t++) {
            const uint32_t j = t - t0;
            const uint32_t* os = j ? d.tp_start + static_cast<size_t>(b->start_len) * (j - 1) : d.ck[seg].start;
            const uint32_t* ob = j ? d.tp_bid + n * (j - 1) : d.ck[seg].bid;
            dtick(b, d, d.theta, d.alpha, t, false, tape_at(d.tape, n, j), os, ob,
                  d.tp_start + static_cast<size_t>(b->start_len) * j, d.tp_bpos + (n + 2) * j, d.tp_bid + n * j);
        }
        rc = bcheck(b, "differentiable tick (segment replay)");
        // reverse through the segment
        for (uint32_t t = t1; t-- > t0 && rc == MSIM_OK;) {
            const uint32_t j = t - t0;
            const DiffCells c = dcells(b, d.tp_start + static_cast<size_t>(b->start_len) * j, d.tp_bpos + (n + 2) * j,
                                       d.tp_bid + n * j);
            hipLaunchKernelGGL(k_davoid_adj, g, dim3(kBlock), 0, b->stream, c, static_cast<double>(d.alpha), d.lam,
                               d.lam2, d.galpha);
            const AgentsDev a = b->agents(b->cur);
            // the decisions after tick t: the tape's next entry, or after the
            // segment's last tick the agent state its replay ended in
            const float2* nt = t + 1 < t1 ? tape_at(d.tape, n, j + 1).tgt : a.tgt;
            const uint32_t* nr = t + 1 < t1 ? tape_at(d.tape, n, j + 1).road : a.road;
            const uint32_t carry = d.carry ? 1u : 0u;
            if (b->uniform) {
                hipLaunchKernelGGL(k_dmove_adj<true>, g, dim3(kBlock), 0, b->stream, b->n, a.spd, a.spd_const,
                                   d.theta, tape_at(d.tape, n, j), nt, nr, carry, d.lam2, d.lam, d.gtheta);
            } else {
                hipLaunchKernelGGL(k_dmove_adj<false>, g, dim3(kBlock), 0, b->stream, b->n, a.spd, a.spd_const,
                                   d.theta, tape_at(d.tape, n, j), nt, nr, carry, d.lam2, d.lam, d.gtheta);
            }
            rc = bcheck(b, "adjoint tick");
        }
    }
    if (rc == MSIM_OK) {
        dcopy(b, grad_theta, d.gtheta, b->n_roads);
        dcopy(b, grad_alpha, d.galpha, 1);
        rc = check_diff_faults(b);
    }
    return rc;
}

extern "C" int msim_batch_diff_set_arrival_gradient(msim_batch* b, int mode) {
    if (!b || !b->diff) return bfail(b, MSIM_ERR_STATE, "msim_batch_diff_setup first");
    if (mode != 0 && mode != 1) return bfail(b, MSIM_ERR_INVALID, "arrival gradient mode must be 0 or 1");
    b->diff->carry = mode == 1;
    return MSIM_OK;
}

extern "C" int msim_batch_diff_read_schedule(msim_batch* b, uint32_t tick, float* target, uint32_t* road,
                                             uint8_t* flag) {
    if (!b || !b->diff) return bfail(b, MSIM_ERR_STATE, "msim_batch_diff_setup first");
    DiffState& d = *b->diff;
    if (!(d.keep_schedule && d.have_forward)) return bfail(b, MSIM_ERR_STATE, "no recorded schedule");
    if (tick > d.T) return bfail(b, MSIM_ERR_INVALID, "tick %u beyond the run's %u ticks", tick, d.T);
    int rc = msim_batch_sync(b);
    if (rc != MSIM_OK) return rc;
    const size_t n = b->n;
    std::vector<float2> t(n);
    std::vector<uint32_t> r(n), perm(n);
    std::vector<uint8_t> f(n);
    const DiffTapeDev s = tape_at(d.sched, n, tick);
    hipError_t e = hipSuccess;
    if (n) {
        e = hipMemcpy(t.data(), s.tgt, sizeof(float2) * n, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(r.data(), s.road, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(f.data(), s.flag, n, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(perm.data(), b->buf[b->cur].perm, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) return bfail(b, MSIM_ERR_HIP, "schedule readback: %s", hipGetErrorString(e));
    for (size_t k = 0; k < n; k++) {
        const uint32_t i = perm[k];
        if (target) {
            target[2 * i] = t[k].x;
            target[2 * i + 1] = t[k].y;
        }
        if (road) road[i] = r[k];
        if (flag) flag[i] = f[k];
    }
    return MSIM_OK;
}
