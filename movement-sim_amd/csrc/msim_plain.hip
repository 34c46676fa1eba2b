// msim_plain.hip — C-ABI of the plain (one scenario per handle) pipeline.
//
// Replaces the Kompute pipeline that Simulator drives
// (src/sim/Simulator.cpp:44-106 creation, :189-199 upload/sequences,
// :213-279 the two dispatches per tick and the readbacks).  See
// include/msim/msim.h for the contract of each entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "msim/msim.h"
#include "msim_bin.hpp"
#include "msim_host.hpp"
#include "msim_kernels.hpp"
#include "msim_sorted.hpp"

using namespace msim;

namespace {

thread_local std::string g_last_error;

struct Buffers {
    float2* pos = nullptr;
    float2* tgt = nullptr;
    float2* dir = nullptr;
    float* spd = nullptr;
    uint32_t* road = nullptr;
    uint4* rng = nullptr;
    uint32_t* init = nullptr;
    uint32_t* ncount = nullptr;
    uint32_t* perm = nullptr;
};

}  // namespace

struct msim_sim {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t n = 0;
    uint32_t n_roads = 0;
    uint32_t n_conns = 0;
    msim_road* roads = nullptr;
    uint32_t* conns = nullptr;
    Buffers buf[2];
    int cur = 0;
    bool uniform = true;
    float spd_const = kShaderSpeed;
    GridDev grid{};  // table parameters and the atomic path's scratch arrays
    // Double-buffered binning: [bcur] is the current one (read by the
    // collision dispatch), the fast move dispatch writes [1 - bcur].
    uint32_t* bstart[2] = {nullptr, nullptr};
    float2* bbpos[2] = {nullptr, nullptr};
    uint32_t* bbid[2] = {nullptr, nullptr};
    int bcur = 0;
    bool stats_pending = false;
    int binning = 0;  // 0: segment-owned (k_bin_count + k_bin_place), 1: atomic counting,
                      // 2: segment-owned in one pass (k_bin_fused, look-back), 3: sorted state
                      // (k_move_sorted: advance and rebinning in one pass, no slot map)
    // Sorted state (binning 3), entered at the first move dispatch of
    // initialised agents: pos, tgt, spd, perm (and dir, ncount) of buf[cur]
    // follow the slot order of binning bcur, which flips with cur on every
    // move; road and PRNG state stay in caller order in cold (cold_word, rng).
    bool sorted = false;
    uint32_t* nxt[2] = {nullptr, nullptr};  // per buffer: bucket of each slot's next position
    uint4* rec[2] = {nullptr, nullptr};     // per buffer: each slot's target, speed and caller index
    uint4* cold = nullptr;                  // [2n] caller order
    uint32_t* ncount_caller = nullptr;      // [n] neighbour counts kept in caller order
    uint4* arr_rec = nullptr;               // [n] arrival records, listed per segment
    float* arr_spd = nullptr;               // [n] their speeds (per-agent speeds only)
    uint32_t* seg_start = nullptr;          // [n_seg + 1] first slot of each segment (k_seg_scan)
    uint32_t* seg_cnt[2] = {nullptr, nullptr};  // per buffer: agents per segment for the move leaving it
    uint32_t* seg_arr = nullptr;            // [n_seg] arrivals of each segment
    int ncount_at = 0;  // buffer whose slot order ncount follows (-1: ncount_caller)
    bool route_inline = false;  // arrivals routed inside k_advance instead of k_route
    bool capturing = false;  // inside build_graph: renumberings skip the output-only arrays
    int collide_mode = 0;  // 0: staged, 2 slots per lane; 1: per lane; 2: staged, 1 slot per lane
    uint32_t bin_unroll = 4;  // window slots per rebinning thread and load round
    // msim_time_ticks probe: launch number probe_kernel of the tick is
    // bracketed by probe_ev[0] / probe_ev[1]; -1 = no probe.
    int probe_kernel = -1;
    int launch_no = 0;
    hipEvent_t probe_ev[2] = {nullptr, nullptr};
    uint32_t seg_log = 0, n_seg = 0, margin = 0;
    uint32_t* seg_total = nullptr;  // [n_seg] per-segment agent totals
    unsigned long long* bin_ticket = nullptr;  // [1] k_bin_fused segment tickets (never reset)
    unsigned long long* seg_status = nullptr;  // [n_seg] k_bin_fused published words
    // Arrival queues of the fast move dispatch; counters are double-buffered by
    // tick parity (k_advance zeroes next tick's set).
    uint32_t* q_items = nullptr;
    uint32_t* q_count = nullptr;  // [2][kQueues * kQueueStride]
    uint32_t q_cap = 0;
    uint32_t q_blocks = 1;  // k_route workgroups per queue
    int qpar = 0;
    Counters* ctr = nullptr;
    unsigned long long* part_stats = nullptr;  // [2][kSlots * kSlotStride] advance statistics
    int stats_par = 0;  // half of part_stats the last move dispatch added to (sorted state: alternates)
    unsigned long long* part_pairs = nullptr;  // [kSlots * kSlotStride] neighbour sums
    uint32_t* hist = nullptr;
    uint32_t hist_cap = 0;
    msim_push_consts pc{};
    uint32_t n_uninit = 0;
    bool any_collide = false;
    float4* color0 = nullptr;       // [n] uploaded colours, caller order
    msim_entity* ent_stage = nullptr;  // [n] Entity records (upload at creation, readback staging)
    uint64_t ticks_simulated = 0;
    uint32_t reorder_interval = 0;
    uint32_t since_reorder = 0;
    uint32_t graph_ticks = 32;
    // captured segments keyed by the start buffers [cur][bcur][qpar]; each
    // holds whole reorder periods and records the buffers it ends on
    struct Segment {
        hipGraphExec_t exec = nullptr;
        int cur = 0, bcur = 0, qpar = 0;
    };
    Segment gexec[2][2][2] = {};
    uint32_t seg_ticks = 32;
    uint32_t watchdog_ms = 600000;
    uint64_t bytes = 0;
    std::vector<void*> allocs;
    std::string err;

    GridDev grid_cur() const {
        GridDev g = grid;
        g.start = bstart[bcur];
        g.bpos = bbpos[bcur];
        g.bid = bbid[bcur];
        return g;
    }

    AgentsDev agents(int which) const {
        const Buffers& b = buf[which];
        return AgentsDev{n, b.pos, b.tgt, b.dir, uniform ? nullptr : b.spd, spd_const,
                         b.road, b.rng, b.init, b.ncount, b.perm};
    }
};

namespace {
unsigned long long* stats_buf(const msim_sim* s) { return s->part_stats + s->stats_par * (kSlots * kSlotStride); }
const char* const kSegKernels[] = {"k_bin_count", "k_bin_place"};
const char* const kAtomicKernels[] = {"k_scan_reduce", "k_scan_partials", "k_scan_down", "k_scatter"};
const char* const kFusedKernels[] = {"k_bin_fused"};
uint32_t bin_kernels(const msim_sim* s) { return s->binning == 1 ? 4u : (s->binning == 2 ? 1u : 2u); }
uint32_t kernels_per_tick(const msim_sim* s) {
    return s->binning == 3 ? 5u : (s->route_inline ? 1u : 2u) + bin_kernels(s) + 2u;
}
// Name of launch k of a fused tick (the kernel's name, template arguments
// dropped).
const char* tick_kernel_name(const msim_sim* s, uint32_t k) {
    if (s->binning == 3) {
        if (k == 0) return "k_seg_scan";
        if (k == 1) return "k_move_sorted";
        if (k == 2) return "k_route_sorted";
        return k == 3 ? (s->collide_mode == 1 ? "k_collide" : "k_collide_staged") : "k_collide_finish";
    }
    if (k == 0) return (s->route_inline && s->binning != 1) ? "k_advance4" : "k_advance";
    if (!s->route_inline) {
        if (k == 1) return "k_route";
        k--;
    }
    k--;
    const uint32_t nbin = bin_kernels(s);
    if (k < nbin) return s->binning == 1 ? kAtomicKernels[k] : (s->binning == 2 ? kFusedKernels[k] : kSegKernels[k]);
    k -= nbin;
    if (k == 0) return s->collide_mode == 1 ? "k_collide" : "k_collide_staged";
    return "k_collide_finish";
}
}  // namespace

namespace {

int fail(msim_sim* s, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (s) {
        s->err = buf;
    }
    g_last_error = buf;
    return code;
}

#define MSIM_TRY_HIP(s, call)                                                             \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            return fail((s), MSIM_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
        }                                                                                 \
    } while (0)

int dalloc(msim_sim* s, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) {
        bytes = 16;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(s, MSIM_ERR_OOM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    s->allocs.push_back(*p);
    s->bytes += bytes;
    return MSIM_OK;
}

template <class T>
int dalloc_n(msim_sim* s, T** p, size_t count) {
    return dalloc(s, reinterpret_cast<void**>(p), count * sizeof(T));
}

uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

uint32_t agent_blocks(uint32_t n) { return std::max(1u, ceil_div(n, kBlock)); }

int check_launch(msim_sim* s, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        return fail(s, MSIM_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    }
    return MSIM_OK;
}

template <int MODE>
void launch_advance_mode(msim_sim* s, bool write_dir) {
    const AgentsDev a = s->agents(s->cur);
    const dim3 g(agent_blocks(s->n));
    if (s->uniform) {
        if (write_dir) {
            hipLaunchKernelGGL((k_advance_bin<MODE, true, true>), g, dim3(kBlock), 0, s->stream, a,
                               s->grid_cur(), s->roads, s->conns, s->part_stats);
        } else {
            hipLaunchKernelGGL((k_advance_bin<MODE, true, false>), g, dim3(kBlock), 0, s->stream, a,
                               s->grid_cur(), s->roads, s->conns, s->part_stats);
        }
    } else {
        if (write_dir) {
            hipLaunchKernelGGL((k_advance_bin<MODE, false, true>), g, dim3(kBlock), 0, s->stream, a,
                               s->grid_cur(), s->roads, s->conns, s->part_stats);
        } else {
            hipLaunchKernelGGL((k_advance_bin<MODE, false, false>), g, dim3(kBlock), 0, s->stream, a,
                               s->grid_cur(), s->roads, s->conns, s->part_stats);
        }
    }
}

// msim_time_ticks brackets one launch of a tick with events: probe_pre /
// probe_post surround every launch of the fast tick and fire for launch
// number probe_kernel only.
void probe_pre(msim_sim* s) {
    if (s->probe_kernel == s->launch_no) (void)hipEventRecord(s->probe_ev[0], s->stream);
}

void probe_post(msim_sim* s) {
    if (s->probe_kernel == s->launch_no) (void)hipEventRecord(s->probe_ev[1], s->stream);
    s->launch_no++;
}

// Segment rebinning: one pass (k_bin_fused) or count + place.
template <uint32_t U, uint32_t T>
void launch_bins(msim_sim* s, const float2* pos, const BinDev& b, size_t lds) {
    if (s->binning == 2) {
        probe_pre(s);
        hipLaunchKernelGGL((k_bin_fused<U, T>), dim3(s->n_seg), dim3(T), lds, s->stream, pos, b);
        probe_post(s);
    } else {
        probe_pre(s);
        hipLaunchKernelGGL((k_bin_count<U, T>), dim3(s->n_seg), dim3(T), lds, s->stream, pos, b);
        probe_post(s);
        probe_pre(s);
        hipLaunchKernelGGL((k_bin_place<U, T>), dim3(s->n_seg), dim3(T), lds, s->stream, pos, b);
        probe_post(s);
    }
}

// Fast move dispatch: streaming advance, route choice for the arrivals, then
// the bucket table rebuild (segment-owned into the other binning buffer, or
// atomic counting + scan + scatter into the current one).  ev (optional)
// receives an event after each launch.
template <bool U, bool W, bool RI>
void launch_advance_ri(msim_sim* s, const QueueDev& q) {
    const AgentsDev a = s->agents(s->cur);
    const GridDev g = s->grid_cur();
    const dim3 nb(agent_blocks(s->n));
    if (s->binning == 1) {
        hipLaunchKernelGGL((k_advance<U, W, true, RI>), nb, dim3(kBlock), 0, s->stream, a, g, q, s->part_stats,
                           s->roads, s->conns);
    } else {
        hipLaunchKernelGGL((k_advance<U, W, false, RI>), nb, dim3(kBlock), 0, s->stream, a, g, q,
                           s->part_stats, s->roads, s->conns);
    }
}

template <bool U, bool W>
void launch_advance_route(msim_sim* s, const QueueDev& q) {
    if (s->route_inline && s->binning != 1) {
        const dim3 nb(std::max(1u, (s->n + kAdv * kBlock - 1) / (kAdv * kBlock)));
        hipLaunchKernelGGL((k_advance4<U, W>), nb, dim3(kBlock), 0, s->stream, s->agents(s->cur),
                           s->grid.inv_cell, s->part_stats, s->roads, s->conns);
    } else if (s->route_inline) {
        launch_advance_ri<U, W, true>(s, q);
    } else {
        launch_advance_ri<U, W, false>(s, q);
    }
}

template <bool U, bool W>
void launch_route(msim_sim* s, const QueueDev& q) {
    hipLaunchKernelGGL((k_route<U, W>), dim3(kQueues * s->q_blocks), dim3(kBlock), 0, s->stream,
                       s->agents(s->cur), q, s->roads, s->conns);
}

int enqueue_move_fast(msim_sim* s, bool write_dir, hipEvent_t* ev) {
    const uint32_t qw = kQueues * kQueueStride;
    QueueDev q{s->q_items, s->q_count + s->qpar * qw, s->q_count + (1 - s->qpar) * qw, s->q_cap};
    s->launch_no = 0;
    probe_pre(s);
    if (s->uniform) {
        if (write_dir) launch_advance_route<true, true>(s, q);
        else launch_advance_route<true, false>(s, q);
    } else {
        if (write_dir) launch_advance_route<false, true>(s, q);
        else launch_advance_route<false, false>(s, q);
    }
    probe_post(s);
    if (!s->route_inline) {
        probe_pre(s);
        if (s->uniform) {
            if (write_dir) launch_route<true, true>(s, q);
            else launch_route<true, false>(s, q);
        } else {
            if (write_dir) launch_route<false, true>(s, q);
            else launch_route<false, false>(s, q);
        }
        probe_post(s);
    }
    s->qpar = 1 - s->qpar;
    const float2* pos = s->agents(s->cur).pos;
    if (s->binning != 1) {
        const int nb = 1 - s->bcur;
        BinDev b{};
        b.old_start = s->bstart[s->bcur];
        b.old_bid = s->bbid[s->bcur];
        b.new_start = s->bstart[nb];
        b.new_bpos = s->bbpos[nb];
        b.new_bid = s->bbid[nb];
        b.seg_total = s->seg_total;
        b.ticket = s->bin_ticket;
        b.status = s->seg_status;
        b.ctr = s->ctr;
        b.inv_cell = s->grid.inv_cell;
        b.lw = s->grid.lw;
        b.wmask = s->grid.wmask;
        b.hmask = s->grid.hmask;
        b.m = s->grid.m;
        b.seg_log = s->seg_log;
        b.n_seg = s->n_seg;
        b.margin = s->margin;
        b.n = s->n;
        const size_t lds = sizeof(uint32_t) << s->seg_log;
        const bool wide = bin_threads_for(s->seg_log) == 1024u;
        if (s->bin_unroll == 8) {
            if (wide) launch_bins<8, 1024>(s, pos, b, lds);
            else launch_bins<8, 512>(s, pos, b, lds);
        } else {
            if (wide) launch_bins<4, 1024>(s, pos, b, lds);
            else launch_bins<4, 512>(s, pos, b, lds);
        }
        s->bcur = nb;
    } else {
        const GridDev g = s->grid_cur();
        const uint32_t n_part = g.m / kScanTile;
        probe_pre(s);
        hipLaunchKernelGGL(k_scan_reduce, dim3(n_part), dim3(kBlock), 0, s->stream, g.cnt, g.part);
        probe_post(s);
        probe_pre(s);
        hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kPartialsThreads), 0, s->stream, g.part, n_part,
                           g.start + g.m);
        probe_post(s);
        probe_pre(s);
        hipLaunchKernelGGL(k_scan_down, dim3(n_part), dim3(kBlock), 0, s->stream, g.cnt, g.part, g.start);
        probe_post(s);
        probe_pre(s);
        hipLaunchKernelGGL(k_scatter, dim3(agent_blocks(s->n)), dim3(kBlock), 0, s->stream, s->agents(s->cur),
                           g);
        probe_post(s);
    }
    s->stats_pending = true;  // folded by the next collision dispatch's finish pass
    return check_launch(s, "move dispatch");
}

// Folds advance statistics that no collision dispatch has folded yet.
void flush_stats(msim_sim* s) {
    if (s->stats_pending) {
        hipLaunchKernelGGL(k_fold_stats, dim3(1), dim3(64), 0, s->stream, stats_buf(s), s->ctr);
        s->stats_pending = false;
    }
}

// Move dispatch: advance + bucket table rebuild.  ev (optional) receives an
// event record after each of the 5 launches of the atomic path.
int enqueue_move_sorted(msim_sim* s, bool write_dir);

int enqueue_move(msim_sim* s, int mode, bool write_dir, hipEvent_t* ev = nullptr) {
    flush_stats(s);
    if (mode == kMoveAll) {
        if (s->sorted) {
            return enqueue_move_sorted(s, write_dir);
        }
        if (s->binning == 3) {
            return fail(s, MSIM_ERR_STATE, "internal: sorted-state move dispatch before the state was sorted");
        }
        return enqueue_move_fast(s, write_dir, ev);
    }
    switch (mode) {
        case kMoveAll:
            launch_advance_mode<kMoveAll>(s, write_dir);
            break;
        case kMoveInit:
            launch_advance_mode<kMoveInit>(s, write_dir);
            break;
        default:
            launch_advance_mode<kBinOnly>(s, false);
            break;
    }
    if (ev) (void)hipEventRecord(ev[0], s->stream);
    const GridDev g = s->grid_cur();
    const uint32_t n_part = g.m / kScanTile;
    hipLaunchKernelGGL(k_scan_reduce, dim3(n_part), dim3(kBlock), 0, s->stream, g.cnt, g.part);
    if (ev) (void)hipEventRecord(ev[1], s->stream);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kPartialsThreads), 0, s->stream, g.part, n_part,
                       g.start + g.m);
    s->stats_pending = (mode != kBinOnly);
    if (ev) (void)hipEventRecord(ev[2], s->stream);
    hipLaunchKernelGGL(k_scan_down, dim3(n_part), dim3(kBlock), 0, s->stream, g.cnt, g.part, g.start);
    if (ev) (void)hipEventRecord(ev[3], s->stream);
    hipLaunchKernelGGL(k_scatter, dim3(agent_blocks(s->n)), dim3(kBlock), 0, s->stream,
                       s->agents(s->cur), g);
    if (ev) (void)hipEventRecord(ev[4], s->stream);
    return check_launch(s, "move dispatch");
}

int enqueue_collide(msim_sim* s, bool slow, hipEvent_t* ev = nullptr) {
    const AgentsDev a = s->agents(s->cur);
    const dim3 g(agent_blocks(s->n));
    const uint32_t fold = s->stats_pending ? 1u : 0u;
    s->stats_pending = false;
    if (s->sorted) {
        // the state itself is in slot order: positions are read in place and
        // the counts stored by slot
        GridDev gd = s->grid_cur();
        gd.bid = nullptr;
        const float2* bpos = s->buf[s->cur].pos;
        probe_pre(s);
        if (s->collide_mode == 1) {
            hipLaunchKernelGGL((k_collide<false>), g, dim3(kBlock), 0, s->stream, a, gd, bpos, s->part_pairs);
        } else if (s->collide_mode == 2) {
            hipLaunchKernelGGL(k_collide_staged<1>, g, dim3(kBlock), 0, s->stream, a, gd, bpos, s->part_pairs);
        } else {
            const uint32_t per = kBlock * 2u;
            hipLaunchKernelGGL(k_collide_staged<2>, dim3(std::max(1u, (s->n + per - 1) / per)), dim3(kBlock), 0,
                               s->stream, a, gd, bpos, s->part_pairs);
        }
        probe_post(s);
        probe_pre(s);
        hipLaunchKernelGGL(k_collide_finish<false>, dim3(1), dim3(64), 0, s->stream, s->part_pairs, s->ctr,
                           s->hist, s->hist_cap - 1, stats_buf(s), fold);
        probe_post(s);
        s->ncount_at = s->cur;
        return check_launch(s, "collision dispatch");
    }
    if (slow) {
        probe_pre(s);
        hipLaunchKernelGGL((k_collide<true>), g, dim3(kBlock), 0, s->stream, a, s->grid_cur(),
                           static_cast<const float2*>(s->bbpos[s->bcur]), s->part_pairs);
        probe_post(s);
        probe_pre(s);
        hipLaunchKernelGGL(k_collide_finish<true>, dim3(1), dim3(64), 0, s->stream, s->part_pairs, s->ctr,
                           s->hist, s->hist_cap - 1, stats_buf(s), fold);
        hipLaunchKernelGGL(k_set_init, g, dim3(kBlock), 0, s->stream, a.init, s->n);
    } else {
        probe_pre(s);
        if (s->collide_mode == 1) {
            hipLaunchKernelGGL((k_collide<false>), g, dim3(kBlock), 0, s->stream, a, s->grid_cur(),
                               static_cast<const float2*>(s->bbpos[s->bcur]), s->part_pairs);
        } else if (s->collide_mode == 2) {
            hipLaunchKernelGGL(k_collide_staged<1>, g, dim3(kBlock), 0, s->stream, a, s->grid_cur(),
                               static_cast<const float2*>(s->bbpos[s->bcur]), s->part_pairs);
        } else {
            const uint32_t per = kBlock * 2u;
            hipLaunchKernelGGL(k_collide_staged<2>, dim3(std::max(1u, (s->n + per - 1) / per)), dim3(kBlock), 0,
                               s->stream, a, s->grid_cur(), static_cast<const float2*>(s->bbpos[s->bcur]),
                               s->part_pairs);
        }
        probe_post(s);
        probe_pre(s);
        hipLaunchKernelGGL(k_collide_finish<false>, dim3(1), dim3(64), 0, s->stream, s->part_pairs, s->ctr,
                           s->hist, s->hist_cap - 1, stats_buf(s), fold);
    }
    probe_post(s);
    return check_launch(s, "collision dispatch");
}

void drop_graphs(msim_sim* s) {
    for (auto& a : s->gexec) {
        for (auto& b : a) {
            for (auto& g : b) {
                if (g.exec) {
                    (void)hipGraphExecDestroy(g.exec);
                    g.exec = nullptr;
                }
            }
        }
    }
}

int reorder(msim_sim* s) {
    const int nxt = 1 - s->cur;
    launch_reorder(s->agents(s->cur), s->agents(nxt), s->grid_cur().bid, s->grid_cur().bpos, s->n, s->stream,
                   !s->capturing);
    int rc = check_launch(s, "k_reorder");
    if (rc != MSIM_OK) {
        return rc;
    }
    s->cur = nxt;
    s->since_reorder = 0;
    return MSIM_OK;
}

SortedBufs sorted_bufs(const msim_sim* s, int k) {
    const Buffers& b = s->buf[k];
    return SortedBufs{b.pos, s->rec[k], s->nxt[k], b.dir};
}

constexpr uint32_t kSortedThreads = 512;


// Enters the sorted state from the classic layout: renumbers the agents into
// the current binning's order (slot map = identity), moves road and PRNG
// state to caller order and computes every slot's next bucket.  Not inside a
// graph capture.
int enter_sorted(msim_sim* s) {
    int rc = reorder(s);
    if (rc != MSIM_OK) return rc;
    const Buffers& b = s->buf[s->cur];
    const dim3 g(agent_blocks(s->n));
    hipLaunchKernelGGL(k_pack_rec, g, dim3(kBlock), 0, s->stream, b.tgt, s->uniform ? nullptr : b.spd, b.perm,
                       s->rec[s->cur], s->n);
    hipLaunchKernelGGL(k_cold_enter, g, dim3(kBlock), 0, s->stream, b.perm, b.road, b.rng, b.tgt, s->roads, s->cold,
                       s->n);
    const GridDev& gd = s->grid;
    const dim3 gs((s->n_seg + kBlock - 1) / kBlock);
    for (int k = 0; k < 2; k++) {
        hipLaunchKernelGGL(k_fill_u32, gs, dim3(kBlock), 0, s->stream, s->seg_cnt[k], 0u, s->n_seg);
    }
    if (s->uniform) {
        hipLaunchKernelGGL(k_sorted_next<true>, g, dim3(kBlock), 0, s->stream, sorted_bufs(s, s->cur), s->n,
                           s->spd_const, gd.inv_cell, gd.lw, gd.wmask, gd.hmask, s->seg_log, s->seg_cnt[s->cur]);
    } else {
        hipLaunchKernelGGL(k_sorted_next<false>, g, dim3(kBlock), 0, s->stream, sorted_bufs(s, s->cur), s->n,
                           s->spd_const, gd.inv_cell, gd.lw, gd.wmask, gd.hmask, s->seg_log, s->seg_cnt[s->cur]);
    }
    s->ncount_at = s->cur;
    s->sorted = true;
    return check_launch(s, "sorted state set-up");
}

int ensure_sorted(msim_sim* s) {
    if (s->binning == 3 && !s->sorted && s->n_uninit == 0) {
        return enter_sorted(s);
    }
    return MSIM_OK;
}

constexpr uint32_t kScanThreads = 1024;

template <uint32_t U, bool UN, bool W>
void launch_sorted(msim_sim* s, const SortedMove& mv, const BinDev& b) {
    probe_pre(s);
    hipLaunchKernelGGL(k_seg_scan<kScanThreads>, dim3(1), dim3(kScanThreads), 0, s->stream, s->seg_cnt[s->cur],
                       s->seg_start, s->n_seg, s->n, &s->ctr->faults);
    probe_post(s);
    probe_pre(s);
    hipLaunchKernelGGL((k_move_sorted<U, kSortedThreads, UN, W>), dim3(s->n_seg), dim3(kSortedThreads),
                       sorted_lds_bytes(s->seg_log), s->stream, mv, b);
    probe_post(s);
    probe_pre(s);
    hipLaunchKernelGGL((k_route_sorted<UN, W>), dim3(s->n_seg), dim3(kBlock), 0, s->stream, mv, b);
    probe_post(s);
}

template <uint32_t U>
void launch_sorted_u(msim_sim* s, const SortedMove& mv, const BinDev& b, bool write_dir) {
    if (s->uniform) {
        if (write_dir) launch_sorted<U, true, true>(s, mv, b);
        else launch_sorted<U, true, false>(s, mv, b);
    } else {
        if (write_dir) launch_sorted<U, false, true>(s, mv, b);
        else launch_sorted<U, false, false>(s, mv, b);
    }
}

// Sorted move dispatch: buf[cur] / binning bcur -> buf[1-cur] / binning
// 1-bcur in one launch.
int enqueue_move_sorted(msim_sim* s, bool write_dir) {
    const int nb = 1 - s->cur;
    const dim3 g(agent_blocks(s->n));
    if (s->ncount_at == nb) {  // the move overwrites the slot order the last counts follow
        hipLaunchKernelGGL(k_scatter_rec<uint32_t>, g, dim3(kBlock), 0, s->stream, s->rec[nb], s->buf[nb].ncount,
                           s->ncount_caller, s->n);
        s->ncount_at = -1;
    }
    SortedMove mv{};
    mv.src = sorted_bufs(s, s->cur);
    mv.dst = sorted_bufs(s, nb);
    mv.cold = s->cold;
    mv.roads = s->roads;
    mv.conns = s->conns;
    mv.spd_const = s->spd_const;
    s->stats_par ^= 1;  // the next tick's collision dispatch folds this half
    mv.part_stats = stats_buf(s);
    mv.faults = &s->ctr->faults;
    mv.arr_rec = s->arr_rec;
    mv.arr_spd = s->arr_spd;
    mv.seg_start = s->seg_start;
    mv.seg_next = s->seg_cnt[nb];
    mv.seg_arr = s->seg_arr;
    mv.n = s->n;
    BinDev b{};
    b.old_start = s->bstart[s->bcur];
    b.new_start = s->bstart[1 - s->bcur];
    b.ticket = s->bin_ticket;
    b.status = s->seg_status;
    b.ctr = s->ctr;
    b.inv_cell = s->grid.inv_cell;
    b.lw = s->grid.lw;
    b.wmask = s->grid.wmask;
    b.hmask = s->grid.hmask;
    b.m = s->grid.m;
    b.seg_log = s->seg_log;
    b.n_seg = s->n_seg;
    b.margin = s->margin;
    b.n = s->n;
    s->launch_no = 0;
    if (s->bin_unroll == 2) launch_sorted_u<2>(s, mv, b, write_dir);
    else if (s->bin_unroll == 8) launch_sorted_u<8>(s, mv, b, write_dir);
    else launch_sorted_u<4>(s, mv, b, write_dir);
    s->cur = nb;
    s->bcur = 1 - s->bcur;
    s->stats_pending = true;  // folded by the next collision dispatch's finish pass
    return check_launch(s, "sorted move dispatch");
}

// Makes the classic per-slot arrays of buf[cur] (target, speed, caller index,
// road, rng, init, ncount) readable for the readbacks; the sorted state stays
// in force.
int publish_sorted(msim_sim* s) {
    if (!s->sorted) return MSIM_OK;
    const Buffers& b = s->buf[s->cur];
    const dim3 g(agent_blocks(s->n));
    hipLaunchKernelGGL(k_unpack_rec, g, dim3(kBlock), 0, s->stream, s->rec[s->cur], b.tgt,
                       s->uniform ? nullptr : b.spd, b.perm, s->n);
    if (s->ncount_at == 1 - s->cur) {
        const Buffers& o = s->buf[1 - s->cur];
        hipLaunchKernelGGL(k_scatter_rec<uint32_t>, g, dim3(kBlock), 0, s->stream, s->rec[1 - s->cur], o.ncount,
                           s->ncount_caller, s->n);
        s->ncount_at = -1;
    }
    if (s->ncount_at == -1) {
        hipLaunchKernelGGL(k_gather_perm<uint32_t>, g, dim3(kBlock), 0, s->stream, b.perm, s->ncount_caller,
                           b.ncount, s->n);
        s->ncount_at = s->cur;
    }
    hipLaunchKernelGGL(k_cold_publish, g, dim3(kBlock), 0, s->stream, b.perm, s->cold, b.road, b.rng, s->n);
    hipLaunchKernelGGL(k_set_init, g, dim3(kBlock), 0, s->stream, b.init, s->n);
    return check_launch(s, "sorted state readback");
}

int wait_stream(msim_sim* s) {
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        hipError_t e = hipStreamQuery(s->stream);
        if (e == hipSuccess) {
            return MSIM_OK;
        }
        if (e != hipErrorNotReady) {
            return fail(s, MSIM_ERR_HIP, "stream failed: %s", hipGetErrorString(e));
        }
        const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(
                            std::chrono::steady_clock::now() - t0)
                            .count();
        if (ms > static_cast<long long>(s->watchdog_ms)) {
            return fail(s, MSIM_ERR_TIMEOUT, "stream did not drain within %u ms", s->watchdog_ms);
        }
        std::this_thread::sleep_for(std::chrono::microseconds(ms < 10 ? 20 : 500));
    }
}

int after_move(msim_sim* s);

// Captures seg_ticks fused ticks starting from the current (cur, bcur, qpar)
// buffers, right after a reorder: the segment is whole reorder periods, each
// ending with its reorder, so every replay starts on freshly ordered agents.
int build_graph(msim_sim* s) {
    hipGraph_t graph = nullptr;
    const int cur0 = s->cur, bcur0 = s->bcur, qpar0 = s->qpar, ncount_at0 = s->ncount_at, spar0 = s->stats_par;
    const uint32_t since0 = s->since_reorder;
    const uint64_t ticks0 = s->ticks_simulated;
    MSIM_TRY_HIP(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = MSIM_OK;
    s->capturing = true;  // run segments are always followed by the run's last tick
    for (uint32_t t = 0; t < s->seg_ticks && rc == MSIM_OK; t++) {
        rc = enqueue_move(s, kMoveAll, false);
        if (rc == MSIM_OK) rc = enqueue_collide(s, false);
        if (rc == MSIM_OK) rc = after_move(s);
    }
    s->capturing = false;
    hipError_t e = hipStreamEndCapture(s->stream, &graph);
    msim_sim::Segment& g = s->gexec[cur0][bcur0][qpar0];
    g.cur = s->cur;
    g.bcur = s->bcur;
    g.qpar = s->qpar;
    s->cur = cur0;  // capture only recorded the ticks; nothing ran
    s->bcur = bcur0;
    s->qpar = qpar0;
    s->ncount_at = ncount_at0;
    s->stats_par = spar0;
    s->since_reorder = since0;
    s->ticks_simulated = ticks0;
    if (rc != MSIM_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) {
        return fail(s, MSIM_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        g.exec = nullptr;
        return fail(s, MSIM_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    return MSIM_OK;
}

int after_move(msim_sim* s) {
    s->ticks_simulated++;
    s->since_reorder++;
    if (!s->sorted && s->reorder_interval && s->since_reorder >= s->reorder_interval) {
        return reorder(s);
    }
    return MSIM_OK;
}

int set_radius(msim_sim* s, float r) {
    if (!(r > 0.0f) || !std::isfinite(r)) {
        return fail(s, MSIM_ERR_INVALID, "collisionRadius must be positive and finite (got %g)", (double)r);
    }
    if (r > 1.0f / s->grid.inv_cell) {
        return fail(s, MSIM_ERR_INVALID,
                    "collisionRadius %g exceeds the grid cell %g chosen at creation", (double)r,
                    (double)(1.0f / s->grid.inv_cell));
    }
    if (r != s->grid.radius) {
        const float lim = d2_limit_for(r);
        if (!box_tests_implied(r, lim)) {
            return fail(s, MSIM_ERR_INVALID, "collisionRadius %a: box tests not implied by the squared limit",
                        (double)r);
        }
        s->grid.radius = r;
        s->grid.d2_limit = lim;
        drop_graphs(s);
    }
    return MSIM_OK;
}

int single_dispatch(msim_sim* s, uint32_t tick) {
    s->pc.tick = tick;
    int rc;
    if ((tick & 1u) == 0u) {
        rc = ensure_sorted(s);
        if (rc != MSIM_OK) return rc;
        rc = enqueue_move(s, s->n_uninit ? kMoveInit : kMoveAll, true);
        s->n_uninit = 0;
        if (rc == MSIM_OK) rc = after_move(s);
    } else {
        rc = enqueue_collide(s, s->n_uninit != 0);
        s->n_uninit = 0;
        s->any_collide = true;
    }
    return rc;
}

template <class T>
int download(msim_sim* s, std::vector<T>& host, const T* dev, size_t count) {
    host.resize(count);
    if (count == 0) return MSIM_OK;
    MSIM_TRY_HIP(s, hipMemcpyAsync(host.data(), dev, count * sizeof(T), hipMemcpyDeviceToHost, s->stream));
    return MSIM_OK;
}

}  // namespace

// ---------------------------------------------------------------------------

extern "C" int msim_create(const msim_road* roads, uint32_t n_roads, const uint32_t* connections,
                           uint32_t n_connections, const msim_entity* entities, uint32_t n_entities,
                           const float* speed, const msim_push_consts* pc, const msim_options* opt,
                           msim_sim** out) {
    if (!out) {
        return fail(nullptr, MSIM_ERR_INVALID, "out handle pointer is NULL");
    }
    *out = nullptr;
    if (!pc) {
        return fail(nullptr, MSIM_ERR_INVALID, "push constants are NULL");
    }
    if ((n_roads && !roads) || (n_connections && !connections) || (n_entities && !entities)) {
        return fail(nullptr, MSIM_ERR_INVALID, "NULL array with non-zero length");
    }
    std::string why;
    if (!validate_map(roads, n_roads, connections, n_connections, &why)) {
        return fail(nullptr, MSIM_ERR_INVALID, "%s", why.c_str());
    }
    if (!validate_entities(entities, n_entities, n_roads, speed, &why)) {
        return fail(nullptr, MSIM_ERR_INVALID, "%s", why.c_str());
    }
    msim_options o{};
    if (opt) {
        if (opt->struct_size != sizeof(msim_options)) {
            return fail(nullptr, MSIM_ERR_INVALID, "msim_options.struct_size must be %zu",
                        sizeof(msim_options));
        }
        o = *opt;
    } else {
        o.device = -1;
    }
    const float radius = pc->collisionRadius;
    if (!(radius > 0.0f) || !std::isfinite(radius) || !(radius * radius < 3.0e38f)) {
        return fail(nullptr, MSIM_ERR_INVALID, "collisionRadius must be positive and finite");
    }
    if (!box_tests_implied(radius, d2_limit_for(radius))) {
        return fail(nullptr, MSIM_ERR_INVALID, "collisionRadius %a: box tests not implied by the squared limit",
                    (double)radius);
    }
    float cell = o.cell_size;
    if (cell == 0.0f) {
        cell = pow2_cell(radius);
    } else if (!is_pow2_float(cell) || cell < radius) {
        return fail(nullptr, MSIM_ERR_INVALID, "cell_size must be a power of two >= collisionRadius");
    }
    const uint32_t block = o.block_size ? o.block_size : kBlock;
    if (block != kBlock) {
        return fail(nullptr, MSIM_ERR_INVALID, "block_size must be %u in this build", kBlock);
    }
    uint32_t tlog = o.table_log2 ? o.table_log2 : auto_table_log2(n_entities);
    if (tlog < 12 || tlog > 28) {
        return fail(nullptr, MSIM_ERR_INVALID, "table_log2 must be in [12, 28]");
    }
    const uint32_t rows_log = extent_rows_log2(roads, n_roads, entities, n_entities, cell);
    const uint32_t lw = o.table_width_log2 ? o.table_width_log2 : auto_width_log2(tlog, rows_log);
    if (lw < 2 || lw + 2 > tlog) {
        return fail(nullptr, MSIM_ERR_INVALID, "table_width_log2 must be in [2, table_log2 - 2]");
    }
    const uint32_t seg_log = o.segment_log2 ? o.segment_log2 : auto_segment_log2(tlog, lw);
    if (seg_log < 9 || seg_log > 14 || seg_log > tlog) {
        return fail(nullptr, MSIM_ERR_INVALID, "segment_log2 must be in [9, 14] and <= table_log2");
    }
    if (o.collide_mode > 2) {
        return fail(nullptr, MSIM_ERR_INVALID, "collide_mode must be 0, 1 or 2");
    }
    if (o.binning > 3) {
        return fail(nullptr, MSIM_ERR_INVALID,
                    "binning must be 0 (segment), 1 (atomic), 2 (segment, one pass) or 3 (sorted state)");
    }
    if (o.bin_unroll != 0 && o.bin_unroll != 4 && o.bin_unroll != 8 && !(o.binning == 3 && o.bin_unroll == 2)) {
        return fail(nullptr, MSIM_ERR_INVALID, "bin_unroll must be 0, 4 or 8 (or 2 with binning 3)");
    }
    if (!validate_extent(roads, n_roads, entities, n_entities, 1.0f / cell, &why)) {
        return fail(nullptr, MSIM_ERR_INVALID, "%s", why.c_str());
    }

    msim_sim* s = new msim_sim();
    s->n = n_entities;
    s->n_roads = n_roads;
    s->n_conns = n_connections;
    s->pc = *pc;
    // the sorted state is in bucket order after every move: no renumbering
    s->reorder_interval = o.binning == 3 ? 0u : o.reorder_interval;
    s->graph_ticks = o.graph_ticks ? o.graph_ticks : 32u;
    s->seg_ticks = segment_ticks(s->graph_ticks, s->reorder_interval);
    s->watchdog_ms = o.watchdog_ms ? o.watchdog_ms : 600000;
    s->hist_cap = next_pow2_u32(o.history_capacity ? o.history_capacity : 4096);
    s->uniform = (speed == nullptr);

    int rc = MSIM_OK;
    if (o.device >= 0) {
        hipError_t e = hipSetDevice(o.device);
        if (e != hipSuccess) {
            delete s;
            return fail(nullptr, MSIM_ERR_HIP, "hipSetDevice(%d): %s", o.device, hipGetErrorString(e));
        }
    }
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return fail(nullptr, MSIM_ERR_HIP, "no HIP device available");
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return fail(nullptr, MSIM_ERR_HIP, "hipStreamCreate failed");
    }

    // LDS per workgroup bounds the segment size (4 bytes per bucket counter).
    int lds_max = 65536;
    (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device);
    const uint32_t sorted_unroll = o.bin_unroll ? o.bin_unroll : 4u;
    auto lds_need = [&](uint32_t l) {
        return (o.binning == 3 ? sorted_lds_bytes(l) : (static_cast<size_t>(4) << l)) + 1024;
    };
    uint32_t seg_use = seg_log;
    while (seg_use > 9 && lds_need(seg_use) > static_cast<size_t>(lds_max)) seg_use--;
    if (o.segment_log2 && seg_use != seg_log) {
        msim_destroy(s);
        return fail(nullptr, MSIM_ERR_INVALID, "segment_log2 %u needs more LDS than the device's %d bytes",
                    seg_log, lds_max);
    }

    const uint32_t n = n_entities;
    GridDev& g = s->grid;
    g.inv_cell = 1.0f / cell;
    g.radius = radius;
    g.d2_limit = d2_limit_for(radius);
    g.m = 1u << tlog;
    g.lw = lw;
    g.wmask = (1u << g.lw) - 1u;
    g.hmask = (1u << (tlog - g.lw)) - 1u;
    s->binning = static_cast<int>(o.binning);
    s->route_inline = o.route_queue == 0;
    s->collide_mode = static_cast<int>(o.collide_mode);
    s->seg_log = seg_use;
    s->n_seg = g.m >> seg_use;
    s->bin_unroll = o.binning == 3 ? sorted_unroll : (o.bin_unroll ? o.bin_unroll : auto_bin_unroll(seg_use));
    // An agent moves at most max_speed per axis per tick (the heading has
    // |dir| = speed up to rounding), i.e. at most D cells; in linear bucket
    // order that is at most (D + 1) rows plus D columns.
    float smax = kShaderSpeed;
    if (speed) {
        smax = 0.0f;
        for (uint32_t i = 0; i < n_entities; i++) smax = std::max(smax, speed[i]);
    }
    const uint32_t D = static_cast<uint32_t>(std::floor(smax * (1.0f + 1.0f / 1048576.0f) / cell)) + 1u;
    s->margin = (D + 1u) * (g.wmask + 1u) + D;

    const size_t n_adv = (static_cast<size_t>(n) + kAdv - 1) / kAdv * kAdv;
#define ALLOC(p, cnt)                                   \
    if (rc == MSIM_OK) rc = dalloc_n(s, &(p), (cnt));
    ALLOC(s->roads, n_roads);
    ALLOC(s->conns, n_connections);
    for (int k = 0; k < 2; k++) {
        Buffers& b = s->buf[k];
        // k_advance4 moves kAdv agents per thread; the neighbour pass's
        // 16-byte staged reads of the sorted state may touch slots n, n + 1
        ALLOC(b.pos, n_adv + 4);
        ALLOC(b.tgt, n_adv);
        ALLOC(b.dir, n_adv);
        if (!s->uniform) ALLOC(b.spd, n_adv);
        ALLOC(b.road, n);
        ALLOC(b.rng, n);
        ALLOC(b.init, n);
        ALLOC(b.ncount, n);
        ALLOC(b.perm, n);
    }
    ALLOC(g.cnt, g.m);
    ALLOC(g.part, g.m / kScanTile);
    ALLOC(g.rank, n);
    for (int k = 0; k < 2; k++) {
        ALLOC(s->bstart[k], g.m + 1);
        ALLOC(s->bbpos[k], n + 2);  // 16-byte staged reads may touch one slot past n
        ALLOC(s->bbid[k], n);
    }
    if (s->binning == 3) {
        ALLOC(s->nxt[0], n);
        ALLOC(s->nxt[1], n);
        ALLOC(s->rec[0], n);
        ALLOC(s->rec[1], n);
        ALLOC(s->cold, 2 * static_cast<size_t>(n));
        ALLOC(s->ncount_caller, n);
        ALLOC(s->arr_rec, n);
        if (!s->uniform) ALLOC(s->arr_spd, n);
        ALLOC(s->seg_start, s->n_seg + 1);
        ALLOC(s->seg_cnt[0], s->n_seg);
        ALLOC(s->seg_cnt[1], s->n_seg);
        ALLOC(s->seg_arr, s->n_seg);
    }
    ALLOC(s->seg_total, s->n_seg);
    ALLOC(s->bin_ticket, 1);
    ALLOC(s->seg_status, s->n_seg);
    s->q_cap = ceil_div(agent_blocks(n), kQueues) * kBlock;
    s->q_blocks = std::min(64u, std::max(1u, ceil_div(n, kQueues * kBlock * 16u)));
    ALLOC(s->q_items, static_cast<size_t>(kQueues) * s->q_cap);
    ALLOC(s->q_count, 2 * kQueues * kQueueStride);
    ALLOC(s->ctr, 1);
    ALLOC(s->part_stats, 2 * kSlots * kSlotStride);
    ALLOC(s->part_pairs, kSlots * kSlotStride);
    ALLOC(s->hist, s->hist_cap);
    ALLOC(s->color0, n);
    ALLOC(s->ent_stage, n);
#undef ALLOC
    if (rc != MSIM_OK) {
        std::string e = s->err;
        msim_destroy(s);
        return fail(nullptr, rc, "%s", e.c_str());
    }

    // The Entity records go up as they are; k_unpack_entities splits them.
    for (uint32_t i = 0; i < n; i++) s->n_uninit += entities[i].initialized ? 0u : 1u;
    auto up = [&](void* dst, const void* src, size_t bytes) {
        if (rc == MSIM_OK && bytes) {
            hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = fail(s, MSIM_ERR_HIP, "upload failed: %s", hipGetErrorString(e));
        }
    };
    Buffers& b0 = s->buf[0];
    up(s->roads, roads, sizeof(msim_road) * n_roads);
    up(s->conns, connections, sizeof(uint32_t) * n_connections);
    up(s->ent_stage, entities, sizeof(msim_entity) * n);
    if (!s->uniform) up(b0.spd, speed, sizeof(float) * n);
    if (rc == MSIM_OK) {
        hipError_t e = hipMemset(g.cnt, 0, sizeof(uint32_t) * g.m);
        if (e == hipSuccess) e = hipMemset(s->ctr, 0, sizeof(Counters));
        if (e == hipSuccess) e = hipMemset(s->bin_ticket, 0, sizeof(unsigned long long));
        if (e == hipSuccess) e = hipMemset(s->seg_status, 0, sizeof(unsigned long long) * s->n_seg);
        if (e == hipSuccess) e = hipMemset(s->q_count, 0, sizeof(uint32_t) * 2 * kQueues * kQueueStride);
        if (e == hipSuccess) e = hipMemset(s->part_stats, 0, sizeof(unsigned long long) * 2 * kSlots * kSlotStride);
        if (e == hipSuccess) e = hipMemset(s->part_pairs, 0, sizeof(unsigned long long) * kSlots * kSlotStride);
        // every binned id slot always holds a valid agent index
        for (int k = 0; k < 2 && e == hipSuccess; k++) e = hipMemset(s->bbid[k], 0, sizeof(uint32_t) * n);
        for (int k = 0; k < 2 && e == hipSuccess; k++) e = hipMemset(s->bstart[k], 0, sizeof(uint32_t) * (g.m + 1));
        if (e == hipSuccess) e = hipMemset(s->hist, 0, sizeof(uint32_t) * s->hist_cap);
        if (e != hipSuccess) rc = fail(s, MSIM_ERR_HIP, "hipMemset failed: %s", hipGetErrorString(e));
    }
    // The uploads and fills above run on the null stream, which the handle's
    // non-blocking stream does not wait for: drain the device before the
    // first launch reads them.
    if (rc == MSIM_OK) {
        hipError_t e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) rc = fail(s, MSIM_ERR_HIP, "upload failed: %s", hipGetErrorString(e));
    }
    if (rc == MSIM_OK && n) {
        hipLaunchKernelGGL(k_unpack_entities, dim3(agent_blocks(n)), dim3(kBlock), 0, s->stream,
                           reinterpret_cast<const uint4*>(s->ent_stage), s->agents(0), s->color0);
        rc = check_launch(s, "k_unpack_entities");
    }
    // Initial binning, then renumber agents into bucket order for locality.
    if (rc == MSIM_OK) rc = enqueue_move(s, kBinOnly, false);
    if (rc == MSIM_OK) rc = reorder(s);
    if (rc == MSIM_OK) rc = wait_stream(s);
    if (rc != MSIM_OK) {
        std::string e = s->err;
        msim_destroy(s);
        return fail(nullptr, rc, "%s", e.c_str());
    }
    *out = s;
    return MSIM_OK;
}

extern "C" int msim_dispatch(msim_sim* s, const msim_push_consts* pc) {
    if (!s || !pc) return fail(s, MSIM_ERR_INVALID, "NULL argument");
    int rc = set_radius(s, pc->collisionRadius);
    if (rc != MSIM_OK) return rc;
    s->pc = *pc;
    return single_dispatch(s, pc->tick);
}

extern "C" int msim_run_ticks(msim_sim* s, uint32_t n_ticks) {
    if (!s) return fail(s, MSIM_ERR_INVALID, "NULL handle");
    uint32_t done = 0;
    int rc = MSIM_OK;
    while (done < n_ticks && rc == MSIM_OK) {
        const uint32_t t1 = s->pc.tick + 1;
        if (s->n_uninit != 0 || (t1 & 1u) != 0u) {
            rc = single_dispatch(s, t1);
            if (rc == MSIM_OK) rc = single_dispatch(s, t1 + 1);
            done++;
            continue;
        }
        rc = ensure_sorted(s);
        if (rc != MSIM_OK) break;
        const uint32_t left = n_ticks - done;
        if (left - 1 >= s->seg_ticks && (s->reorder_interval == 0 || s->since_reorder == 0) &&
            (!s->sorted || s->ncount_at == s->cur)) {
            msim_sim::Segment& g = s->gexec[s->cur][s->bcur][s->qpar];
            if (!g.exec) {
                rc = build_graph(s);
                if (rc != MSIM_OK) break;
            }
            hipError_t e = hipGraphLaunch(g.exec, s->stream);
            if (e != hipSuccess) return fail(s, MSIM_ERR_HIP, "hipGraphLaunch: %s", hipGetErrorString(e));
            s->cur = g.cur;
            s->bcur = g.bcur;
            s->qpar = g.qpar;
            if (s->sorted) {
                s->ncount_at = s->cur;  // every captured tick ends with its collision dispatch
                s->stats_par ^= static_cast<int>(s->seg_ticks & 1u);  // the halves the segment's moves used
            }
            s->pc.tick += 2 * s->seg_ticks;
            s->ticks_simulated += s->seg_ticks;
            s->any_collide = true;
            done += s->seg_ticks;
        } else {
            // Direction is output-only state: write it on the final tick.
            rc = enqueue_move(s, kMoveAll, left == 1);
            if (rc == MSIM_OK) rc = enqueue_collide(s, false);
            s->pc.tick += 2;
            s->any_collide = true;
            done++;
            if (rc == MSIM_OK) rc = after_move(s);
        }
    }
    return rc;
}

int check_faults(msim_sim* s) {
    Counters c{};
    MSIM_TRY_HIP(s, hipMemcpyAsync(&c, s->ctr, sizeof(c), hipMemcpyDeviceToHost, s->stream));
    int rc = wait_stream(s);
    if (rc != MSIM_OK) return rc;
    if (c.faults) {
        return fail(s, MSIM_ERR_STATE, "device-side check failed (fault bits 0x%x%s%s%s)", c.faults,
                    (c.faults & kFaultBinCount) ? ": binned agent count != agent count" : "",
                    (c.faults & kFaultLookback) ? ": rebinning look-back timed out" : "",
                    (c.faults & kFaultNext) ? ": an agent's position left its counted bucket" : "");
    }
    return MSIM_OK;
}

extern "C" int msim_sync(msim_sim* s) {
    if (!s) return fail(s, MSIM_ERR_INVALID, "NULL handle");
    return check_faults(s);
}

namespace {

struct HostState {
    std::vector<float2> pos, tgt, dir;
    std::vector<uint32_t> road, init, ncount, perm;
    std::vector<uint4> rng;
};

int fetch_state(msim_sim* s, HostState& h, bool need_dir, bool need_rng) {
    const Buffers& b = s->buf[s->cur];
    int rc = publish_sorted(s);
    if (rc == MSIM_OK) rc = download(s, h.pos, b.pos, s->n);
    if (rc == MSIM_OK) rc = download(s, h.tgt, b.tgt, s->n);
    if (rc == MSIM_OK && need_dir) rc = download(s, h.dir, b.dir, s->n);
    if (rc == MSIM_OK) rc = download(s, h.road, b.road, s->n);
    if (rc == MSIM_OK && need_rng) rc = download(s, h.rng, b.rng, s->n);
    if (rc == MSIM_OK) rc = download(s, h.init, b.init, s->n);
    if (rc == MSIM_OK) rc = download(s, h.ncount, b.ncount, s->n);
    if (rc == MSIM_OK) rc = download(s, h.perm, b.perm, s->n);
    if (rc == MSIM_OK) rc = check_faults(s);
    return rc;
}

}  // namespace

extern "C" int msim_read_entities(msim_sim* s, msim_entity* out, uint32_t first, uint32_t count) {
    if (!s || (count && !out)) return fail(s, MSIM_ERR_INVALID, "NULL argument");
    if (static_cast<uint64_t>(first) + count > s->n) {
        return fail(s, MSIM_ERR_INVALID, "range [%u, %u) exceeds %u agents", first, first + count, s->n);
    }
    if (count == 0) return msim_sync(s);
    flush_stats(s);
    int rc0 = publish_sorted(s);
    if (rc0 != MSIM_OK) return rc0;
    hipLaunchKernelGGL(k_pack_entities, dim3(agent_blocks(s->n)), dim3(kBlock), 0, s->stream, s->agents(s->cur),
                       s->color0, s->any_collide ? 1u : 0u, first, count, reinterpret_cast<uint4*>(s->ent_stage));
    int rc = check_launch(s, "k_pack_entities");
    if (rc != MSIM_OK) return rc;
    MSIM_TRY_HIP(s, hipMemcpyAsync(out, s->ent_stage, sizeof(msim_entity) * count, hipMemcpyDeviceToHost,
                                   s->stream));
    rc = check_faults(s);
    return rc;
}

extern "C" int msim_read_state(msim_sim* s, float* pos, float* target, float* direction,
                               uint32_t* road_index, uint32_t* rand_state, uint32_t* neighbour_count) {
    if (!s) return fail(s, MSIM_ERR_INVALID, "NULL handle");
    HostState h;
    int rc = fetch_state(s, h, direction != nullptr, rand_state != nullptr);
    if (rc != MSIM_OK) return rc;
    for (uint32_t k = 0; k < s->n; k++) {
        const uint32_t i = h.perm[k];
        if (pos) {
            pos[2 * i] = h.pos[k].x;
            pos[2 * i + 1] = h.pos[k].y;
        }
        if (target) {
            target[2 * i] = h.tgt[k].x;
            target[2 * i + 1] = h.tgt[k].y;
        }
        if (direction) {
            direction[2 * i] = h.dir[k].x;
            direction[2 * i + 1] = h.dir[k].y;
        }
        if (road_index) road_index[i] = h.road[k];
        if (rand_state) {
            rand_state[4 * i] = h.rng[k].x;
            rand_state[4 * i + 1] = h.rng[k].y;
            rand_state[4 * i + 2] = h.rng[k].z;
            rand_state[4 * i + 3] = h.rng[k].w;
        }
        if (neighbour_count) neighbour_count[i] = h.ncount[k];
    }
    return MSIM_OK;
}

extern "C" int msim_read_debug(msim_sim* s, uint32_t out[10]) {
    if (!s || !out) return fail(s, MSIM_ERR_INVALID, "NULL argument");
    flush_stats(s);
    Counters c{};
    MSIM_TRY_HIP(s, hipMemcpyAsync(&c, s->ctr, sizeof(c), hipMemcpyDeviceToHost, s->stream));
    int rc = wait_stream(s);
    if (rc != MSIM_OK) return rc;
    std::memset(out, 0, sizeof(uint32_t) * 10);
    out[0] = static_cast<uint32_t>(c.inserts);
    out[1] = static_cast<uint32_t>(c.pairs_total);
    return MSIM_OK;
}

extern "C" int msim_read_pairs(msim_sim* s, uint64_t* total, uint32_t* per_tick, uint32_t n,
                               uint32_t* n_out) {
    if (!s) return fail(s, MSIM_ERR_INVALID, "NULL handle");
    Counters c{};
    std::vector<uint32_t> hist(s->hist_cap);
    MSIM_TRY_HIP(s, hipMemcpyAsync(&c, s->ctr, sizeof(c), hipMemcpyDeviceToHost, s->stream));
    MSIM_TRY_HIP(s, hipMemcpyAsync(hist.data(), s->hist, sizeof(uint32_t) * s->hist_cap,
                                   hipMemcpyDeviceToHost, s->stream));
    int rc = wait_stream(s);
    if (rc != MSIM_OK) return rc;
    if (total) *total = c.pairs_total;
    const uint64_t avail = std::min<uint64_t>(c.collides, s->hist_cap);
    const uint32_t k = static_cast<uint32_t>(std::min<uint64_t>(avail, per_tick ? n : 0));
    for (uint32_t q = 0; q < k; q++) {
        const uint64_t idx = c.collides - k + q;
        per_tick[q] = hist[idx & (s->hist_cap - 1)];
    }
    if (n_out) *n_out = k;
    return MSIM_OK;
}

namespace {

// Every agent's cell into out[caller index]; in the sorted state the caller
// index comes from the slots' records.
void launch_export_cells(msim_sim* s, int2* out) {
    const dim3 g(agent_blocks(s->n));
    if (s->sorted) {
        hipLaunchKernelGGL(k_export_cells_rec, g, dim3(kBlock), 0, s->stream, s->buf[s->cur].pos, s->rec[s->cur],
                           s->n, s->grid.inv_cell, out);
    } else {
        hipLaunchKernelGGL(k_export_cells, g, dim3(kBlock), 0, s->stream, s->agents(s->cur), s->grid.inv_cell, out);
    }
}

}  // namespace

extern "C" int msim_read_cells(msim_sim* s, int32_t* cell_xy) {
    if (!s || (s->n && !cell_xy)) return fail(s, MSIM_ERR_INVALID, "NULL argument");
    if (s->n == 0) return MSIM_OK;
    int2* d = nullptr;
    MSIM_TRY_HIP(s, hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(int2) * s->n, s->stream));
    launch_export_cells(s, d);
    int rc = check_launch(s, "k_export_cells");
    if (rc == MSIM_OK) {
        hipError_t e = hipMemcpyAsync(cell_xy, d, sizeof(int2) * s->n, hipMemcpyDeviceToHost, s->stream);
        if (e != hipSuccess) rc = fail(s, MSIM_ERR_HIP, "cell readback: %s", hipGetErrorString(e));
    }
    (void)hipFreeAsync(d, s->stream);
    if (rc == MSIM_OK) rc = wait_stream(s);
    return rc;
}

namespace {

// Staging of msim_read_cells_noalloc: one array per GPU, in the code object,
// so that the readback needs no device allocation.  Host code never passes
// the symbol itself to a kernel or a copy: it stands for a host-side shadow,
// and the device address comes from hipGetSymbolAddress, per device.
constexpr uint32_t kCellsScratchAgents = 1u << 21;  // 16 MiB of int2
__device__ __align__(256) int2 cells_scratch[kCellsScratchAgents];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
int2* g_scratch_ptr[kMaxDevices] = {};

// Device address of cells_scratch on the handle's device (looked up once per
// device, with that device current).
int cells_scratch_ptr(msim_sim* s, int2** out) {
    if (s->device < 0 || s->device >= kMaxDevices) {
        return fail(s, MSIM_ERR_INVALID, "device %d out of range for the allocation-free readback", s->device);
    }
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_ptr[s->device]) {
        int prev = 0;
        MSIM_TRY_HIP(s, hipGetDevice(&prev));
        MSIM_TRY_HIP(s, hipSetDevice(s->device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(cells_scratch));
        (void)hipSetDevice(prev);
        if (e != hipSuccess || p == nullptr) {
            return fail(s, MSIM_ERR_HIP, "hipGetSymbolAddress(cells_scratch) failed: %s", hipGetErrorString(e));
        }
        g_scratch_ptr[s->device] = static_cast<int2*>(p);
    }
    *out = g_scratch_ptr[s->device];
    return MSIM_OK;
}

}  // namespace

extern "C" int msim_read_cells_noalloc(msim_sim* s, int32_t* cell_xy) {
    if (!s || (s->n && !cell_xy)) return fail(s, MSIM_ERR_INVALID, "NULL argument");
    if (s->n > kCellsScratchAgents) {
        return fail(s, MSIM_ERR_INVALID,
                    "msim_read_cells_noalloc handles at most %u agents (this handle has %u); use msim_read_cells",
                    kCellsScratchAgents, s->n);
    }
    if (s->n == 0) return MSIM_OK;
    int2* d = nullptr;
    int rc = cells_scratch_ptr(s, &d);
    if (rc != MSIM_OK) return rc;
    launch_export_cells(s, d);
    rc = check_launch(s, "k_export_cells");
    if (rc == MSIM_OK) {
        const hipError_t e = hipMemcpyAsync(cell_xy, d, sizeof(int2) * s->n, hipMemcpyDeviceToHost, s->stream);
        if (e != hipSuccess) rc = fail(s, MSIM_ERR_HIP, "cell readback: %s", hipGetErrorString(e));
    }
    if (rc == MSIM_OK) rc = wait_stream(s);
    return rc;
}

extern "C" int msim_read_grid_occupancy(msim_sim* s, int32_t* cell_xy, uint32_t* counts, uint32_t capacity,
                                        uint32_t* n_cells) {
    if (!s || !n_cells || (capacity && (!cell_xy || !counts))) return fail(s, MSIM_ERR_INVALID, "NULL argument");
    *n_cells = 0;
    if (s->n == 0) return MSIM_OK;
    const uint32_t cap = std::min(capacity, s->n);
    const size_t cell_bytes = sizeof(int2) * std::max(cap, 1u), cnt_bytes = sizeof(uint32_t) * std::max(cap, 1u);
    char* d = nullptr;
    MSIM_TRY_HIP(s, hipMallocAsync(reinterpret_cast<void**>(&d), cell_bytes + cnt_bytes + 256, s->stream));
    int2* d_cell = reinterpret_cast<int2*>(d);
    uint32_t* d_cnt = reinterpret_cast<uint32_t*>(d + cell_bytes);
    uint32_t* d_total = reinterpret_cast<uint32_t*>(d + cell_bytes + cnt_bytes);
    int rc = MSIM_OK;
    hipError_t e = hipMemsetAsync(d_total, 0, sizeof(uint32_t), s->stream);
    if (e != hipSuccess) rc = fail(s, MSIM_ERR_HIP, "hipMemsetAsync: %s", hipGetErrorString(e));
    if (rc == MSIM_OK) {
        // the current binning: the sorted state itself, or the binned copy
        const float2* bpos = s->sorted ? s->buf[s->cur].pos : s->bbpos[s->bcur];
        hipLaunchKernelGGL(k_occupancy, dim3(agent_blocks(s->n)), dim3(kBlock), 0, s->stream, bpos, s->bstart[s->bcur],
                           s->n, s->grid.inv_cell, s->grid.lw, s->grid.wmask, s->grid.hmask, d_cell, d_cnt, d_total,
                           cap);
        rc = check_launch(s, "k_occupancy");
    }
    uint32_t total = 0;
    if (rc == MSIM_OK) {
        e = hipMemcpyAsync(&total, d_total, sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) rc = fail(s, MSIM_ERR_HIP, "occupancy readback: %s", hipGetErrorString(e));
    }
    const uint32_t k = std::min(total, cap);
    if (rc == MSIM_OK && k) {
        e = hipMemcpyAsync(cell_xy, d_cell, sizeof(int2) * k, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(counts, d_cnt, sizeof(uint32_t) * k, hipMemcpyDeviceToHost, s->stream);
        if (e != hipSuccess) rc = fail(s, MSIM_ERR_HIP, "occupancy readback: %s", hipGetErrorString(e));
    }
    (void)hipFreeAsync(d, s->stream);
    if (rc == MSIM_OK) rc = wait_stream(s);
    if (rc == MSIM_OK) *n_cells = total;
    return rc;
}

extern "C" int msim_get_info(msim_sim* s, msim_info* out) {
    if (!s || !out) return fail(s, MSIM_ERR_INVALID, "NULL argument");
    out->n_agents = s->n;
    out->n_roads = s->n_roads;
    out->n_connections = s->n_conns;
    out->tick = s->pc.tick;
    out->ticks_simulated = s->ticks_simulated;
    out->cell_size = 1.0f / s->grid.inv_cell;
    out->collision_radius = s->grid.radius;
    out->table_width = s->grid.wmask + 1;
    out->table_height = s->grid.hmask + 1;
    out->uniform_speed = s->uniform ? 1u : 0u;
    out->n_uninitialized = s->n_uninit;
    out->device_bytes = s->bytes;
    out->kernels_per_tick = kernels_per_tick(s);
    out->binning = static_cast<uint32_t>(s->binning);
    return MSIM_OK;
}

extern "C" const char* msim_kernel_name(const msim_sim* s, uint32_t k) {
    if (!s || k >= kernels_per_tick(s)) return nullptr;
    return tick_kernel_name(s, k);
}

extern "C" int msim_time_ticks(msim_sim* s, uint32_t n_ticks, float* ms_total, float* kernel_ms,
                               uint32_t n_kernel_ms) {
    if (!s) return fail(s, MSIM_ERR_INVALID, "NULL handle");
    if (s->n_uninit != 0) return fail(s, MSIM_ERR_STATE, "agents are not initialised yet; dispatch first");
    if ((s->pc.tick + 1) & 1u) return fail(s, MSIM_ERR_STATE, "next dispatch is not a move dispatch");
    {
        const int rc0 = ensure_sorted(s);
        if (rc0 != MSIM_OK) return rc0;
    }
    constexpr int kMaxK = 8;
    const int nk = static_cast<int>(kernels_per_tick(s));
    hipEvent_t ev[kMaxK + 1];
    for (auto& e : ev) MSIM_TRY_HIP(s, hipEventCreate(&e));
    int rc = MSIM_OK;
    if (ms_total) {
        (void)hipEventRecord(ev[0], s->stream);
        rc = msim_run_ticks(s, n_ticks);
        (void)hipEventRecord(ev[1], s->stream);
        if (rc == MSIM_OK) rc = wait_stream(s);
        if (rc == MSIM_OK) (void)hipEventElapsedTime(ms_total, ev[0], ev[1]);
    }
    if (rc == MSIM_OK && kernel_ms && n_kernel_ms) {
        // Kernel k of a tick is timed alone: an event right before and right
        // after its launch, the tick's other kernels running untimed.
        double acc[kMaxK] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t t = 0; t < n_ticks && rc == MSIM_OK; t++) {
            const int k = static_cast<int>(t % static_cast<uint32_t>(nk));
            s->probe_kernel = k;
            s->probe_ev[0] = ev[0];
            s->probe_ev[1] = ev[1];
            rc = enqueue_move(s, kMoveAll, t + 1 == n_ticks);
            if (rc == MSIM_OK) rc = enqueue_collide(s, false);
            s->probe_kernel = -1;
            s->pc.tick += 2;
            s->any_collide = true;
            if (rc == MSIM_OK) rc = after_move(s);
            if (rc == MSIM_OK) rc = wait_stream(s);
            if (rc == MSIM_OK) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) acc[k] += ms * nk;
            }
        }
        for (uint32_t k = 0; k < n_kernel_ms && k < static_cast<uint32_t>(nk); k++) {
            kernel_ms[k] = static_cast<float>(acc[k]);
        }
    }
    (void)hipGetLastError();
    for (auto& e : ev) (void)hipEventDestroy(e);
    return rc;
}

extern "C" int msim_get_stream(msim_sim* s, void** stream) {
    if (!s || !stream) return fail(s, MSIM_ERR_INVALID, "NULL argument");
    *stream = s->stream;
    return MSIM_OK;
}

extern "C" int msim_destroy(msim_sim* s) {
    if (!s) return MSIM_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    drop_graphs(s);
    for (void* p : s->allocs) (void)hipFree(p);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return MSIM_OK;
}

extern "C" const char* msim_last_error(const msim_sim* s) { return s ? s->err.c_str() : g_last_error.c_str(); }

extern "C" const char* msim_get_last_error(void) { return g_last_error.c_str(); }

#ifndef MSIM_FP_CONTRACT
#define MSIM_FP_CONTRACT "off"
#endif
extern "C" const char* msim_version(void) { return "msim 0.1.0 gfx950 fp-contract=" MSIM_FP_CONTRACT; }
