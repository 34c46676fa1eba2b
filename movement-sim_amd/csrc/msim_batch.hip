// msim_batch.hip — the batched form: many independent scenarios per launch.
//
// Scenarios are packed: agents, roads and connections of scenario s occupy
// [off_s, off_{s+1}) of the flattened arrays, with road indices and
// connection indices rebased to the packed arrays on upload, so the advance
// kernel of the plain path runs on the flattened state unchanged.  Each
// scenario owns a region of the bucket table (its own torus) and the
// contiguous binned slots [agent_off_s, agent_off_s + n_s), so the rebinning
// and neighbour kernels only need the scenario's table parameters, found from
// the segment (rebinning) or by binary search over agent_off (neighbours).
// A tick is 5 launches for all scenarios; ticks run as captured hipGraph
// segments.  Per-scenario results are bit-identical to plain runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "msim/msim.h"
#include "msim_bin.hpp"
#include "msim_host.hpp"
#include "msim_kernels.hpp"

using namespace msim;

namespace msim {

struct ScenDev {
    uint32_t agent_off, n;
    uint32_t start_base;  // offset of this scenario's start[m + 1] in the start arrays
    uint32_t lw, wmask, hmask, m;
    uint32_t seg_log, seg_first, n_seg, margin;
    uint32_t pad;
};

constexpr uint32_t kScenSlots = 16;     // pair partial slots per scenario
constexpr uint32_t kScenSlotStride = 8;  // 64-bit words between slots (64 B)

struct BatchBin {
    const ScenDev* scen;
    const uint32_t* seg_scen;  // global segment -> scenario
    const uint32_t* old_start;
    const uint32_t* old_bid;
    uint32_t* new_start;
    float2* new_bpos;
    uint32_t* new_bid;
    uint32_t* seg_total;
    unsigned int* faults;
    float inv_cell;
};

__device__ __forceinline__ TableView scen_view(const ScenDev& sc, const BatchBin& b) {
    return TableView{b.old_start + sc.start_base, b.old_bid, b.inv_cell, sc.lw, sc.wmask, sc.hmask, sc.m,
                     sc.seg_log, sc.margin};
}

__global__ __launch_bounds__(kBinThreads) void k_bbin_count(const float2* __restrict__ pos, BatchBin b) {
    extern __shared__ uint32_t hist[];
    __shared__ uint32_t lds[kBinThreads / 64];
    const uint32_t g = blockIdx.x;
    const ScenDev sc = b.scen[b.seg_scen[g]];
    const uint32_t s0 = (g - sc.seg_first) << sc.seg_log;
    const uint32_t total = bin_count_segment(pos, scen_view(sc, b), s0, hist, lds, b.new_start + sc.start_base);
    if (threadIdx.x == 0) {
        b.seg_total[g] = total;
    }
}

__global__ __launch_bounds__(kBinThreads) void k_bbin_place(const float2* __restrict__ pos, BatchBin b) {
    extern __shared__ uint32_t cursor[];
    __shared__ uint32_t lds[kBinThreads / 64];
    const uint32_t g = blockIdx.x;
    const ScenDev sc = b.scen[b.seg_scen[g]];
    uint32_t before = 0;
    for (uint32_t k = sc.seg_first + threadIdx.x; k < g; k += kBinThreads) {
        before += b.seg_total[k];
    }
    const uint32_t base = sc.agent_off + block_sum<kBinThreads>(before, lds);
    if (g == sc.seg_first + sc.n_seg - 1 && threadIdx.x == 0) {
        const uint32_t end = base + b.seg_total[g];
        b.new_start[sc.start_base + sc.m] = end;
        if (end != sc.agent_off + sc.n) {
            atomicOr(b.faults, kFaultBinCount);
        }
    }
    bin_place_segment(pos, scen_view(sc, b), (g - sc.seg_first) << sc.seg_log, base, cursor,
                      b.new_start + sc.start_base, b.new_bpos, b.new_bid);
}

// Table parameters of the scenario holding a block's first slot, and the
// slot where that scenario ends (read with scalar loads: block-uniform).
struct BlockScen {
    uint32_t sid, end, start_base, lw, wmask, hmask, pad0, pad1;
};

struct BatchCollide {
    const BlockScen* blk;
    const ScenDev* scen;
    const uint32_t* agent_off;
    uint32_t n;
    const uint32_t* start;
    const float2* bpos;
    const uint32_t* bid;
    uint32_t* ncount;
    float inv_cell;
    float d2_limit;
    unsigned long long* part;  // [n_scen][kScenSlots * kScenSlotStride]
};

// Neighbour pass over every binned slot of every scenario, each slot on its
// own scenario's bucket table (count_in_cells, as in the plain pass).
template <bool STAGED>
__global__ __launch_bounds__(kBlock) void k_bcollide(BatchCollide c) {
    __shared__ float2 stage_all[STAGED ? kBlock / 64 * kStage : 1];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t s = blk * kBlock + threadIdx.x;
    const bool valid = s < c.n;
    const BlockScen bs = c.blk[blk];
    uint32_t cnt = 0;
    uint32_t sid = bs.sid;
    const float2 me = valid ? c.bpos[s] : make_float2(0.0f, 0.0f);
    const int32_t cx = cell_coord(me.x, c.inv_cell);
    const int32_t cy = cell_coord(me.y, c.inv_cell);
    if (valid && s >= bs.end) {  // a later scenario starts inside this block
        sid++;
        while (s >= c.agent_off[sid + 1]) sid++;
    }
    const uint32_t w0 = s - __lane_id();
    bool done = false;
    if (STAGED && w0 < c.n) {
        const uint32_t sid0 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(sid)));
        if (__ballot(valid && sid != sid0) == 0ull) {  // the wave lies in one scenario
            uint32_t base = bs.start_base, lw = bs.lw, wmask = bs.wmask, hmask = bs.hmask;
            if (sid0 != bs.sid) {
                const ScenDev& sc = c.scen[sid0];
                base = sc.start_base;
                lw = sc.lw;
                wmask = sc.wmask;
                hmask = sc.hmask;
            }
            cnt = wave_count_in_cells(me, valid, min(63u, c.n - 1u - w0), cx, cy, c.start + base, lw, wmask,
                                      hmask, c.bpos, c.d2_limit, stage_all + (threadIdx.x >> 6) * kStage);
            done = true;
        }
    }
    if (valid && !done) {
        const ScenDev& sc = c.scen[sid];
        cnt = count_in_cells(me, s, cx, cy, c.start + sc.start_base, sc.lw, sc.wmask, sc.hmask, c.bpos,
                             c.d2_limit);
    }
    if (valid) {
        cnt -= 1u;  // own slot
        c.ncount[c.bid[s]] = cnt;
    }
    // per-scenario partial sums: one atomic per wave when the wave is one
    // scenario, per lane at scenario boundaries
    const unsigned long long act = __ballot(valid);
    const int first = act ? __ffsll(static_cast<long long>(act)) - 1 : 0;
    const uint32_t s0 = __shfl(sid, first, 64);
    const bool one = __ballot(valid && sid != s0) == 0ull;
    const uint32_t wave = (blk * kBlock + threadIdx.x) >> 6;
    const uint32_t slot = wave % kScenSlots;
    if (one) {
        unsigned long long v = cnt;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            v += __shfl_xor(v, o, 64);
        }
        if (act && __lane_id() == 0 && v) {
            atomicAdd(&c.part[(static_cast<size_t>(s0) * kScenSlots + slot) * kScenSlotStride], v);
        }
    } else if (valid && cnt) {
        atomicAdd(&c.part[(static_cast<size_t>(sid) * kScenSlots + slot) * kScenSlotStride],
                  static_cast<unsigned long long>(cnt));
    }
}

// One wave per scenario: fold the slots into the scenario's pair totals.
__global__ __launch_bounds__(64) void k_bfinish(unsigned long long* __restrict__ part,
                                                unsigned long long* __restrict__ totals,
                                                uint32_t* __restrict__ last) {
    const uint32_t sidx = blockIdx.x;
    unsigned long long v = 0;
    if (threadIdx.x < kScenSlots) {
        unsigned long long* p = &part[(static_cast<size_t>(sidx) * kScenSlots + threadIdx.x) * kScenSlotStride];
        v = *p;
        *p = 0ull;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_xor(v, o, 64);
    }
    if (threadIdx.x == 0) {
        const unsigned long long pairs = v / 2;
        totals[sidx] += pairs;
        last[sidx] = static_cast<uint32_t>(pairs);
    }
}

}  // namespace msim

// ---------------------------------------------------------------------------

namespace {

thread_local std::string g_batch_error;

struct BBuffers {
    float2 *pos = nullptr, *tgt = nullptr, *dir = nullptr;
    float* spd = nullptr;
    uint32_t *road = nullptr, *init = nullptr, *ncount = nullptr, *perm = nullptr;
    uint4* rng = nullptr;
};

}  // namespace

struct msim_batch {
    hipStream_t stream = nullptr;
    uint32_t n_scen = 0, n = 0, n_roads = 0, n_conns = 0, n_seg = 0, start_len = 0, max_seg_log = 9;
    std::vector<uint32_t> agent_off_h, road_off_h;
    msim_road* roads = nullptr;
    uint32_t* conns = nullptr;
    BBuffers buf[2];
    int cur = 0;
    bool uniform = true;
    float inv_cell = 1.0f, d2_limit = 1.0f;
    ScenDev* scen = nullptr;
    uint32_t* agent_off = nullptr;
    BlockScen* blk_scen = nullptr;
    uint32_t* seg_scen = nullptr;
    uint32_t* seg_total = nullptr;
    uint32_t* bstart[2] = {nullptr, nullptr};
    float2* bbpos[2] = {nullptr, nullptr};
    uint32_t* bbid[2] = {nullptr, nullptr};
    int bcur = 0;
    unsigned long long* part = nullptr;
    unsigned long long* part_stats = nullptr;
    unsigned long long* totals = nullptr;
    uint32_t* last = nullptr;
    unsigned int* faults = nullptr;
    uint32_t* qdummy = nullptr;
    bool first_tick_insert_only = false;
    bool collide_per_lane = false;
    uint32_t tick = 1;
    uint32_t graph_ticks = 100, seg_ticks = 100;
    uint32_t reorder_interval = 20, since_reorder = 0;
    struct Segment {
        hipGraphExec_t exec = nullptr;
        int cur = 0, bcur = 0;
    };
    Segment gexec[2][2] = {};  // [cur][bcur] at the segment's start
    uint32_t watchdog_ms = 600000;
    // msim_batch_time_ticks probe: launch number probe_kernel of a tick is
    // bracketed by probe_ev[0] / probe_ev[1]; -1 = no probe
    int probe_kernel = -1, launch_no = 0;
    hipEvent_t probe_ev[2] = {nullptr, nullptr};
    std::vector<void*> allocs;
    std::string err;

    AgentsDev agents(int k) const {
        const BBuffers& b = buf[k];
        return AgentsDev{n, b.pos, b.tgt, b.dir, uniform ? nullptr : b.spd, kShaderSpeed,
                         b.road, b.rng, b.init, b.ncount, b.perm};
    }
};

namespace {

int bfail(msim_batch* b, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    g_batch_error = buf;
    return code;
}

int balloc(msim_batch* b, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return bfail(b, MSIM_ERR_OOM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    b->allocs.push_back(*p);
    return MSIM_OK;
}

template <class T>
int balloc_n(msim_batch* b, T** p, size_t count) {
    return balloc(b, reinterpret_cast<void**>(p), count * sizeof(T));
}

uint32_t bblocks(uint32_t n) { return std::max(1u, (n + kBlock - 1) / kBlock); }

int bcheck(msim_batch* b, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, MSIM_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return MSIM_OK;
}

// Launch order of a batched tick, for msim_batch_time_ticks.
const char* const kBatchKernels[] = {"k_advance", "k_bbin_count", "k_bbin_place", "k_bcollide", "k_bfinish"};
constexpr int kBatchKernelCount = 5;

void bprobe_pre(msim_batch* b) {
    if (b->probe_kernel == b->launch_no) (void)hipEventRecord(b->probe_ev[0], b->stream);
}

void bprobe_post(msim_batch* b) {
    if (b->probe_kernel == b->launch_no) (void)hipEventRecord(b->probe_ev[1], b->stream);
    b->launch_no++;
}

int bwait(msim_batch* b) {
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        hipError_t e = hipStreamQuery(b->stream);
        if (e == hipSuccess) return MSIM_OK;
        if (e != hipErrorNotReady) return bfail(b, MSIM_ERR_HIP, "stream failed: %s", hipGetErrorString(e));
        const auto ms =
            std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
        if (ms > static_cast<long long>(b->watchdog_ms)) {
            return bfail(b, MSIM_ERR_TIMEOUT, "stream did not drain within %u ms", b->watchdog_ms);
        }
        std::this_thread::sleep_for(std::chrono::microseconds(ms < 10 ? 20 : 500));
    }
}

GridDev bgrid(const msim_batch* b) {
    GridDev g{};
    g.inv_cell = b->inv_cell;
    g.d2_limit = b->d2_limit;
    g.bid = b->bbid[b->bcur];
    return g;
}

int benqueue_collide(msim_batch* b) {
    BatchCollide c{};
    c.blk = b->blk_scen;
    c.scen = b->scen;
    c.agent_off = b->agent_off;
    c.n = b->n;
    c.start = b->bstart[b->bcur];
    c.bpos = b->bbpos[b->bcur];
    c.bid = b->bbid[b->bcur];
    c.ncount = b->buf[b->cur].ncount;
    c.inv_cell = b->inv_cell;
    c.d2_limit = b->d2_limit;
    c.part = b->part;
    bprobe_pre(b);
    if (b->collide_per_lane) {
        hipLaunchKernelGGL(k_bcollide<false>, dim3(bblocks(b->n)), dim3(kBlock), 0, b->stream, c);
    } else {
        hipLaunchKernelGGL(k_bcollide<true>, dim3(bblocks(b->n)), dim3(kBlock), 0, b->stream, c);
    }
    bprobe_post(b);
    bprobe_pre(b);
    hipLaunchKernelGGL(k_bfinish, dim3(b->n_scen), dim3(64), 0, b->stream, b->part, b->totals, b->last);
    bprobe_post(b);
    return bcheck(b, "batched collision dispatch");
}

template <bool U, bool W>
void badvance(msim_batch* b) {
    const dim3 nb(std::max(1u, (b->n + kAdv * kBlock - 1) / (kAdv * kBlock)));
    hipLaunchKernelGGL((k_advance4<U, W>), nb, dim3(kBlock), 0, b->stream, b->agents(b->cur), b->inv_cell,
                       b->part_stats, b->roads, b->conns);
}

int benqueue_move(msim_batch* b, bool write_dir) {
    b->launch_no = 0;
    bprobe_pre(b);
    if (b->uniform) {
        if (write_dir) badvance<true, true>(b);
        else badvance<true, false>(b);
    } else {
        if (write_dir) badvance<false, true>(b);
        else badvance<false, false>(b);
    }
    bprobe_post(b);
    const int nb = 1 - b->bcur;
    BatchBin bb{};
    bb.scen = b->scen;
    bb.seg_scen = b->seg_scen;
    bb.old_start = b->bstart[b->bcur];
    bb.old_bid = b->bbid[b->bcur];
    bb.new_start = b->bstart[nb];
    bb.new_bpos = b->bbpos[nb];
    bb.new_bid = b->bbid[nb];
    bb.seg_total = b->seg_total;
    bb.faults = b->faults;
    bb.inv_cell = b->inv_cell;
    const size_t lds = sizeof(uint32_t) << b->max_seg_log;
    const float2* pos = b->agents(b->cur).pos;
    bprobe_pre(b);
    hipLaunchKernelGGL(k_bbin_count, dim3(b->n_seg), dim3(kBinThreads), lds, b->stream, pos, bb);
    bprobe_post(b);
    bprobe_pre(b);
    hipLaunchKernelGGL(k_bbin_place, dim3(b->n_seg), dim3(kBinThreads), lds, b->stream, pos, bb);
    bprobe_post(b);
    b->bcur = nb;
    return bcheck(b, "batched move dispatch");
}

int breorder(msim_batch* b) {
    const int nxt = 1 - b->cur;
    hipLaunchKernelGGL(k_reorder, dim3(bblocks(b->n)), dim3(kBlock), 0, b->stream, b->agents(b->cur),
                       b->agents(nxt), bgrid(b));
    int rc = bcheck(b, "k_reorder");
    if (rc != MSIM_OK) return rc;
    b->cur = nxt;
    b->since_reorder = 0;
    return MSIM_OK;
}

int btick(msim_batch* b, bool write_dir) {
    int rc = benqueue_move(b, write_dir);
    if (rc == MSIM_OK) rc = benqueue_collide(b);
    b->tick += 2;
    b->since_reorder++;
    if (rc == MSIM_OK && b->reorder_interval && b->since_reorder >= b->reorder_interval) rc = breorder(b);
    return rc;
}

// Captures seg_ticks ticks (whole reorder periods, each ending with its
// reorder) from the current buffers, right after a reorder.
int bbuild_graph(msim_batch* b) {
    hipGraph_t graph = nullptr;
    const int cur0 = b->cur, bcur0 = b->bcur;
    const uint32_t tick0 = b->tick, since0 = b->since_reorder;
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bfail(b, MSIM_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    int rc = MSIM_OK;
    for (uint32_t t = 0; t < b->seg_ticks && rc == MSIM_OK; t++) {
        rc = btick(b, false);
    }
    e = hipStreamEndCapture(b->stream, &graph);
    msim_batch::Segment& g = b->gexec[cur0][bcur0];
    g.cur = b->cur;
    g.bcur = b->bcur;
    b->cur = cur0;  // capture only recorded the ticks; nothing ran
    b->bcur = bcur0;
    b->tick = tick0;
    b->since_reorder = since0;
    if (rc != MSIM_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, MSIM_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        g.exec = nullptr;
        return bfail(b, MSIM_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    return MSIM_OK;
}

// Initial binning on the host: counting sort of each scenario's agents by
// bucket (positions as uploaded, internal order = packed order).
void host_bin(const std::vector<ScenDev>& sc, const msim_entity* ents, float inv_cell, uint32_t start_len,
              std::vector<uint32_t>& start, std::vector<float2>& bpos, std::vector<uint32_t>& bid) {
    const size_t n = sc.empty() ? 0 : sc.back().agent_off + sc.back().n;
    start.assign(start_len, 0u);
    bpos.resize(n);
    bid.resize(n);
    std::vector<uint32_t> bucket(n);
    for (const ScenDev& s : sc) {
        uint32_t* st = start.data() + s.start_base;
        for (uint32_t i = s.agent_off; i < s.agent_off + s.n; i++) {
            const float x = ents[i].pos.x, y = ents[i].pos.y;
            const int32_t cx = static_cast<int32_t>(std::floor(x * inv_cell));
            const int32_t cy = static_cast<int32_t>(std::floor(y * inv_cell));
            bucket[i] = ((static_cast<uint32_t>(cy) & s.hmask) << s.lw) | (static_cast<uint32_t>(cx) & s.wmask);
            st[bucket[i] + 1]++;
        }
        st[0] = s.agent_off;
        for (uint32_t k = 0; k < s.m; k++) st[k + 1] += st[k];
        std::vector<uint32_t> cursor(st, st + s.m);
        for (uint32_t i = s.agent_off; i < s.agent_off + s.n; i++) {
            const uint32_t slot = cursor[bucket[i]]++;
            bpos[slot] = make_float2(ents[i].pos.x, ents[i].pos.y);
            bid[slot] = i;
        }
    }
}

}  // namespace

extern "C" int msim_batch_create(uint32_t n_scen, const uint32_t* road_off, const msim_road* roads,
                                 const uint32_t* conn_off, const uint32_t* connections, const uint32_t* agent_off,
                                 const msim_entity* entities, const float* speed, const msim_push_consts* pc,
                                 const msim_options* opt, msim_batch** out) {
    if (!out || !pc || !road_off || !conn_off || !agent_off || n_scen == 0) {
        return bfail(nullptr, MSIM_ERR_INVALID, "NULL argument or zero scenarios");
    }
    *out = nullptr;
    if (road_off[0] != 0 || conn_off[0] != 0 || agent_off[0] != 0) {
        return bfail(nullptr, MSIM_ERR_INVALID, "offset arrays must start at 0");
    }
    for (uint32_t s = 0; s < n_scen; s++) {
        if (road_off[s + 1] < road_off[s] || conn_off[s + 1] < conn_off[s] || agent_off[s + 1] < agent_off[s]) {
            return bfail(nullptr, MSIM_ERR_INVALID, "offsets of scenario %u decrease", s);
        }
    }
    const uint32_t n_roads = road_off[n_scen], n_conns = conn_off[n_scen], n = agent_off[n_scen];
    if ((n_roads && !roads) || (n_conns && !connections) || (n && !entities)) {
        return bfail(nullptr, MSIM_ERR_INVALID, "NULL array with non-zero length");
    }
    msim_options o{};
    if (opt) {
        if (opt->struct_size != sizeof(msim_options)) {
            return bfail(nullptr, MSIM_ERR_INVALID, "msim_options.struct_size must be %zu", sizeof(msim_options));
        }
        o = *opt;
    } else {
        o.device = -1;
    }
    if ((pc->tick & 1u) == 0u) {
        return bfail(nullptr, MSIM_ERR_INVALID,
                     "PushConsts.tick must be odd: each batched tick is a move dispatch (even) followed by "
                     "a collision dispatch (odd)");
    }
    const float radius = pc->collisionRadius;
    if (!(radius > 0.0f) || !std::isfinite(radius) || !(radius * radius < 3.0e38f)) {
        return bfail(nullptr, MSIM_ERR_INVALID, "collisionRadius must be positive and finite");
    }
    const float cell = o.cell_size ? o.cell_size : pow2_cell(radius);
    if (!is_pow2_float(cell) || cell < radius) {
        return bfail(nullptr, MSIM_ERR_INVALID, "cell_size must be a power of two >= collisionRadius");
    }
    const float lim = d2_limit_for(radius);
    if (!box_tests_implied(radius, lim)) {
        return bfail(nullptr, MSIM_ERR_INVALID, "collisionRadius: box tests not implied by the squared limit");
    }
    // per-scenario validation, on each scenario's own slice
    std::string why;
    uint32_t init0 = n ? (entities[0].initialized ? 1u : 0u) : 1u;
    for (uint32_t s = 0; s < n_scen; s++) {
        const uint32_t nr = road_off[s + 1] - road_off[s], nc = conn_off[s + 1] - conn_off[s];
        const uint32_t na = agent_off[s + 1] - agent_off[s];
        if (!validate_map(roads + road_off[s], nr, connections + conn_off[s], nc, &why) ||
            !validate_entities(entities + agent_off[s], na, nr, speed ? speed + agent_off[s] : nullptr, &why) ||
            !validate_extent(roads + road_off[s], nr, entities + agent_off[s], na, 1.0f / cell, &why)) {
            return bfail(nullptr, MSIM_ERR_INVALID, "scenario %u: %s", s, why.c_str());
        }
        for (uint32_t i = agent_off[s]; i < agent_off[s + 1]; i++) {
            if ((entities[i].initialized ? 1u : 0u) != init0) {
                return bfail(nullptr, MSIM_ERR_INVALID, "all agents must share one Entity.initialized value");
            }
        }
    }

    msim_batch* b = new msim_batch();
    b->n_scen = n_scen;
    b->n = n;
    b->n_roads = n_roads;
    b->n_conns = n_conns;
    b->uniform = speed == nullptr;
    b->inv_cell = 1.0f / cell;
    b->d2_limit = lim;
    b->first_tick_insert_only = (init0 == 0);
    b->tick = pc->tick;
    b->graph_ticks = o.graph_ticks ? o.graph_ticks : 100u;
    b->reorder_interval = opt ? o.reorder_interval : 20u;
    b->seg_ticks = segment_ticks(b->graph_ticks, b->reorder_interval);
    b->collide_per_lane = o.collide_mode != 0;
    b->watchdog_ms = o.watchdog_ms ? o.watchdog_ms : 600000;
    b->agent_off_h.assign(agent_off, agent_off + n_scen + 1);
    b->road_off_h.assign(road_off, road_off + n_scen + 1);
    if (o.device >= 0 && hipSetDevice(o.device) != hipSuccess) {
        delete b;
        return bfail(nullptr, MSIM_ERR_HIP, "hipSetDevice(%d) failed", o.device);
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return bfail(nullptr, MSIM_ERR_HIP, "hipStreamCreate failed");
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    int lds_max = 65536;
    (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);

    // per-scenario table layout
    std::vector<ScenDev> sc(n_scen);
    std::vector<uint32_t> seg_scen;
    uint32_t start_len = 0;
    float smax = kShaderSpeed;
    if (speed) {
        smax = 0.0f;
        for (uint32_t i = 0; i < n; i++) smax = std::max(smax, speed[i]);
    }
    const uint32_t D = static_cast<uint32_t>(std::floor(smax * (1.0f + 1.0f / 1048576.0f) / cell)) + 1u;
    for (uint32_t s = 0; s < n_scen; s++) {
        ScenDev& d = sc[s];
        d.agent_off = agent_off[s];
        d.n = agent_off[s + 1] - agent_off[s];
        const uint32_t tlog = auto_table_log2(d.n);
        const uint32_t rows_log = extent_rows_log2(roads + road_off[s], road_off[s + 1] - road_off[s],
                                                   entities + agent_off[s], d.n, cell);
        d.lw = auto_width_log2(tlog, rows_log);
        uint32_t seg_log = auto_segment_log2(tlog, d.lw);
        if (o.segment_log2) seg_log = std::min(std::max(o.segment_log2, 9u), tlog);
        while (seg_log > 9 && (static_cast<size_t>(4) << seg_log) + 1024 > static_cast<size_t>(lds_max)) seg_log--;
        d.seg_log = seg_log;
        d.m = 1u << tlog;
        d.wmask = (1u << d.lw) - 1u;
        d.hmask = (1u << (tlog - d.lw)) - 1u;
        d.margin = (D + 1u) * (d.wmask + 1u) + D;
        d.start_base = start_len;
        start_len += d.m + 1u;
        d.seg_first = static_cast<uint32_t>(seg_scen.size());
        d.n_seg = d.m >> seg_log;
        for (uint32_t k = 0; k < d.n_seg; k++) seg_scen.push_back(s);
        b->max_seg_log = std::max(b->max_seg_log, seg_log);
    }
    b->n_seg = static_cast<uint32_t>(seg_scen.size());
    b->start_len = start_len;

    int rc = MSIM_OK;
    const size_t n_adv = (static_cast<size_t>(n) + kAdv - 1) / kAdv * kAdv;
#define BALLOC(p, cnt) \
    if (rc == MSIM_OK) rc = balloc_n(b, &(p), (cnt));
    BALLOC(b->roads, n_roads);
    BALLOC(b->conns, n_conns);
    for (int k = 0; k < 2; k++) {
        BBuffers& bb = b->buf[k];
        BALLOC(bb.pos, n_adv);  // k_advance4 moves kAdv agents per thread
        BALLOC(bb.tgt, n_adv);
        BALLOC(bb.dir, n_adv);
        if (!b->uniform) BALLOC(bb.spd, n_adv);
        BALLOC(bb.road, n);
        BALLOC(bb.rng, n);
        BALLOC(bb.init, n);
        BALLOC(bb.ncount, n);
        BALLOC(bb.perm, n);
        BALLOC(b->bstart[k], start_len);
        BALLOC(b->bbpos[k], n + 2);  // 16-byte staged reads may touch one slot past n
        BALLOC(b->bbid[k], n);
    }
    BALLOC(b->scen, n_scen);
    BALLOC(b->agent_off, n_scen + 1);
    BALLOC(b->blk_scen, bblocks(n));
    BALLOC(b->seg_scen, b->n_seg);
    BALLOC(b->seg_total, b->n_seg);
    BALLOC(b->part, static_cast<size_t>(n_scen) * kScenSlots * kScenSlotStride);
    BALLOC(b->part_stats, kSlots * kSlotStride);
    BALLOC(b->totals, n_scen);
    BALLOC(b->last, n_scen);
    BALLOC(b->faults, 1);
    BALLOC(b->qdummy, kQueues * kQueueStride);
#undef BALLOC
    if (rc != MSIM_OK) {
        std::string e = b->err;
        msim_batch_destroy(b);
        return bfail(nullptr, rc, "%s", e.c_str());
    }

    // rebase to the packed arrays
    std::vector<msim_road> r(roads, roads + n_roads);
    std::vector<uint32_t> cn(connections, connections + n_conns);
    for (uint32_t s = 0; s < n_scen; s++) {
        for (uint32_t k = road_off[s]; k < road_off[s + 1]; k++) {
            r[k].start.connectedIndex += conn_off[s];
            r[k].end.connectedIndex += conn_off[s];
        }
        for (uint32_t k = conn_off[s]; k < conn_off[s + 1]; k++) cn[k] += road_off[s];
    }
    std::vector<float2> pos(n), tgt(n), dir(n);
    std::vector<uint32_t> road(n), init(n), perm(n), zeros(n, 0u);
    std::vector<uint4> rng(n);
    for (uint32_t s = 0; s < n_scen; s++) {
        for (uint32_t i = agent_off[s]; i < agent_off[s + 1]; i++) {
            const msim_entity& e = entities[i];
            pos[i] = make_float2(e.pos.x, e.pos.y);
            tgt[i] = make_float2(e.target.x, e.target.y);
            dir[i] = make_float2(e.direction.x, e.direction.y);
            road[i] = e.roadIndex + road_off[s];
            rng[i] = make_uint4(e.randomState.x, e.randomState.y, e.randomState.z, e.randomState.w);
            init[i] = 1u;  // the insert-only first tick is handled by the handle
            perm[i] = i;
        }
    }
    std::vector<uint32_t> start_h, bid_h;
    std::vector<float2> bpos_h;
    host_bin(sc, entities, b->inv_cell, start_len, start_h, bpos_h, bid_h);
    auto up = [&](void* dst, const void* src, size_t bytes) {
        if (rc == MSIM_OK && bytes) {
            hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = bfail(b, MSIM_ERR_HIP, "upload failed: %s", hipGetErrorString(e));
        }
    };
    BBuffers& b0 = b->buf[0];
    up(b->roads, r.data(), sizeof(msim_road) * n_roads);
    up(b->conns, cn.data(), sizeof(uint32_t) * n_conns);
    up(b0.pos, pos.data(), sizeof(float2) * n);
    up(b0.tgt, tgt.data(), sizeof(float2) * n);
    up(b0.dir, dir.data(), sizeof(float2) * n);
    if (!b->uniform) up(b0.spd, speed, sizeof(float) * n);
    up(b0.road, road.data(), sizeof(uint32_t) * n);
    up(b0.rng, rng.data(), sizeof(uint4) * n);
    up(b0.init, init.data(), sizeof(uint32_t) * n);
    up(b0.ncount, zeros.data(), sizeof(uint32_t) * n);
    up(b0.perm, perm.data(), sizeof(uint32_t) * n);
    up(b->scen, sc.data(), sizeof(ScenDev) * n_scen);
    up(b->agent_off, agent_off, sizeof(uint32_t) * (n_scen + 1));
    std::vector<BlockScen> blk(bblocks(n));
    for (uint32_t k = 0, sid = 0; k < blk.size(); k++) {
        const uint32_t s0 = k * kBlock;
        while (sid + 1 < n_scen && agent_off[sid + 1] <= s0) sid++;  // skips empty scenarios
        const ScenDev& d = sc[sid];
        blk[k] = BlockScen{sid, agent_off[sid + 1], d.start_base, d.lw, d.wmask, d.hmask, 0u, 0u};
    }
    up(b->blk_scen, blk.data(), sizeof(BlockScen) * blk.size());
    up(b->seg_scen, seg_scen.data(), sizeof(uint32_t) * b->n_seg);
    up(b->bstart[0], start_h.data(), sizeof(uint32_t) * start_len);
    up(b->bbpos[0], bpos_h.data(), sizeof(float2) * n);
    up(b->bbid[0], bid_h.data(), sizeof(uint32_t) * n);
    up(b->bbid[1], bid_h.data(), sizeof(uint32_t) * n);
    if (rc == MSIM_OK) {
        hipError_t e = hipMemset(b->part, 0, sizeof(unsigned long long) * n_scen * kScenSlots * kScenSlotStride);
        if (e == hipSuccess) e = hipMemset(b->part_stats, 0, sizeof(unsigned long long) * kSlots * kSlotStride);
        if (e == hipSuccess) e = hipMemset(b->totals, 0, sizeof(unsigned long long) * n_scen);
        if (e == hipSuccess) e = hipMemset(b->last, 0, sizeof(uint32_t) * n_scen);
        if (e == hipSuccess) e = hipMemset(b->faults, 0, sizeof(unsigned int));
        if (e == hipSuccess) e = hipMemset(b->qdummy, 0, sizeof(uint32_t) * kQueues * kQueueStride);
        if (e == hipSuccess) e = hipMemset(b->bstart[1], 0, sizeof(uint32_t) * start_len);
        if (e != hipSuccess) rc = bfail(b, MSIM_ERR_HIP, "hipMemset failed: %s", hipGetErrorString(e));
    }
    // the uploads and fills above run on the null stream, which the handle's
    // non-blocking stream does not wait for
    if (rc == MSIM_OK) {
        hipError_t e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) rc = bfail(b, MSIM_ERR_HIP, "upload failed: %s", hipGetErrorString(e));
    }
    if (rc == MSIM_OK) rc = breorder(b);
    if (rc == MSIM_OK) rc = bwait(b);
    if (rc != MSIM_OK) {
        std::string e = b->err;
        msim_batch_destroy(b);
        return bfail(nullptr, rc, "%s", e.c_str());
    }
    *out = b;
    return MSIM_OK;
}

extern "C" int msim_batch_run_ticks(msim_batch* b, uint32_t n_ticks) {
    if (!b) return bfail(b, MSIM_ERR_INVALID, "NULL handle");
    int rc = MSIM_OK;
    uint32_t done = 0;
    if (n_ticks && b->first_tick_insert_only) {
        // the reference's first sim_tick: insert-only move dispatch (no motion),
        // then a collision dispatch on the initial positions
        rc = benqueue_collide(b);
        b->tick += 2;
        b->first_tick_insert_only = false;
        done = 1;
    }
    while (done < n_ticks && rc == MSIM_OK) {
        const uint32_t left = n_ticks - done;
        if (left - 1 >= b->seg_ticks && (b->reorder_interval == 0 || b->since_reorder == 0)) {
            msim_batch::Segment& g = b->gexec[b->cur][b->bcur];
            if (!g.exec) {
                rc = bbuild_graph(b);
                if (rc != MSIM_OK) break;
            }
            hipError_t e = hipGraphLaunch(g.exec, b->stream);
            if (e != hipSuccess) return bfail(b, MSIM_ERR_HIP, "hipGraphLaunch: %s", hipGetErrorString(e));
            b->cur = g.cur;
            b->bcur = g.bcur;
            b->tick += 2 * b->seg_ticks;
            done += b->seg_ticks;
        } else {
            // direction is output-only state: written on the final tick
            rc = btick(b, left == 1);
            done++;
        }
    }
    return rc;
}

extern "C" int msim_batch_sync(msim_batch* b) {
    if (!b) return bfail(b, MSIM_ERR_INVALID, "NULL handle");
    unsigned int f = 0;
    hipError_t e = hipMemcpyAsync(&f, b->faults, sizeof(f), hipMemcpyDeviceToHost, b->stream);
    if (e != hipSuccess) return bfail(b, MSIM_ERR_HIP, "fault readback: %s", hipGetErrorString(e));
    int rc = bwait(b);
    if (rc != MSIM_OK) return rc;
    if (f) return bfail(b, MSIM_ERR_STATE, "device-side check failed (fault bits 0x%x)", f);
    return MSIM_OK;
}

extern "C" int msim_batch_read_state(msim_batch* b, uint32_t scen, float* pos, float* target, float* direction,
                                     uint32_t* road_index, uint32_t* rand_state, uint32_t* neighbour_count) {
    if (!b || scen >= b->n_scen) return bfail(b, MSIM_ERR_INVALID, "bad handle or scenario index");
    int rc = msim_batch_sync(b);
    if (rc != MSIM_OK) return rc;
    const uint32_t off = b->agent_off_h[scen], na = b->agent_off_h[scen + 1] - off;
    const BBuffers& bb = b->buf[b->cur];
    std::vector<float2> p(na), t(na), d(na);
    std::vector<uint32_t> r(na), c(na), pm(na);
    std::vector<uint4> g(na);
    auto dn = [&](void* dst, const void* src, size_t bytes) {
        if (rc == MSIM_OK && bytes) {
            hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = bfail(b, MSIM_ERR_HIP, "readback failed: %s", hipGetErrorString(e));
        }
    };
    // a scenario's agents stay inside its slot range through every reorder
    dn(p.data(), bb.pos + off, sizeof(float2) * na);
    dn(t.data(), bb.tgt + off, sizeof(float2) * na);
    dn(d.data(), bb.dir + off, sizeof(float2) * na);
    dn(r.data(), bb.road + off, sizeof(uint32_t) * na);
    dn(g.data(), bb.rng + off, sizeof(uint4) * na);
    dn(c.data(), bb.ncount + off, sizeof(uint32_t) * na);
    dn(pm.data(), bb.perm + off, sizeof(uint32_t) * na);
    if (rc != MSIM_OK) return rc;
    const uint32_t road_base = b->road_off_h[scen];
    for (uint32_t k = 0; k < na; k++) {
        const uint32_t i = pm[k] - off;
        if (pos) {
            pos[2 * i] = p[k].x;
            pos[2 * i + 1] = p[k].y;
        }
        if (target) {
            target[2 * i] = t[k].x;
            target[2 * i + 1] = t[k].y;
        }
        if (direction) {
            direction[2 * i] = d[k].x;
            direction[2 * i + 1] = d[k].y;
        }
        if (road_index) road_index[i] = r[k] - road_base;
        if (rand_state) {
            rand_state[4 * i] = g[k].x;
            rand_state[4 * i + 1] = g[k].y;
            rand_state[4 * i + 2] = g[k].z;
            rand_state[4 * i + 3] = g[k].w;
        }
        if (neighbour_count) neighbour_count[i] = c[k];
    }
    return MSIM_OK;
}

extern "C" int msim_batch_read_pairs(msim_batch* b, uint64_t* totals, uint32_t* last) {
    if (!b) return bfail(b, MSIM_ERR_INVALID, "NULL handle");
    int rc = msim_batch_sync(b);
    if (rc != MSIM_OK) return rc;
    if (totals &&
        hipMemcpy(totals, b->totals, sizeof(uint64_t) * b->n_scen, hipMemcpyDeviceToHost) != hipSuccess) {
        return bfail(b, MSIM_ERR_HIP, "pair readback failed");
    }
    if (last && hipMemcpy(last, b->last, sizeof(uint32_t) * b->n_scen, hipMemcpyDeviceToHost) != hipSuccess) {
        return bfail(b, MSIM_ERR_HIP, "pair readback failed");
    }
    return MSIM_OK;
}

extern "C" int msim_batch_time_ticks(msim_batch* b, uint32_t n_ticks, float* ms_total, float* kernel_ms,
                                     uint32_t n_kernel_ms) {
    if (!b) return bfail(b, MSIM_ERR_INVALID, "NULL handle");
    hipEvent_t ev[2];
    if (hipEventCreate(&ev[0]) != hipSuccess) return bfail(b, MSIM_ERR_HIP, "hipEventCreate failed");
    if (hipEventCreate(&ev[1]) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return bfail(b, MSIM_ERR_HIP, "hipEventCreate failed");
    }
    int rc = MSIM_OK;
    if (ms_total) {
        (void)hipEventRecord(ev[0], b->stream);
        rc = msim_batch_run_ticks(b, n_ticks);
        (void)hipEventRecord(ev[1], b->stream);
        if (rc == MSIM_OK) rc = msim_batch_sync(b);
        if (rc == MSIM_OK) (void)hipEventElapsedTime(ms_total, ev[0], ev[1]);
    }
    if (rc == MSIM_OK && kernel_ms && n_kernel_ms) {
        uint32_t t0 = 0;
        if (n_ticks && b->first_tick_insert_only) {  // no move kernels in that tick
            rc = msim_batch_run_ticks(b, 1);
            t0 = 1;
        }
        // kernel k of a tick timed alone, the tick's other launches untimed
        double acc[kBatchKernelCount] = {0, 0, 0, 0, 0};
        for (uint32_t t = t0; t < n_ticks && rc == MSIM_OK; t++) {
            const int k = static_cast<int>((t - t0) % kBatchKernelCount);
            b->probe_kernel = k;
            b->probe_ev[0] = ev[0];
            b->probe_ev[1] = ev[1];
            rc = btick(b, t + 1 == n_ticks);
            b->probe_kernel = -1;
            if (rc == MSIM_OK) rc = bwait(b);
            float ms = 0.f;
            if (rc == MSIM_OK && hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) {
                acc[k] += static_cast<double>(ms) * kBatchKernelCount;
            }
        }
        for (uint32_t k = 0; k < n_kernel_ms && k < static_cast<uint32_t>(kBatchKernelCount); k++) {
            kernel_ms[k] = static_cast<float>(acc[k]);
        }
    }
    (void)hipGetLastError();
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return rc;
}

extern "C" const char* msim_batch_kernel_name(uint32_t k) {
    return k < static_cast<uint32_t>(kBatchKernelCount) ? kBatchKernels[k] : nullptr;
}

extern "C" int msim_batch_destroy(msim_batch* b) {
    if (!b) return MSIM_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto& row : b->gexec)
        for (auto& g : row)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return MSIM_OK;
}

extern "C" const char* msim_batch_last_error(const msim_batch* b) {
    return b ? b->err.c_str() : g_batch_error.c_str();
}
