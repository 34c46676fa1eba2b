// msim_batch.hpp — internals of the batched form shared by msim_batch.hip
// (the tick pipeline and its C-ABI) and msim_diff.hip (the differentiable
// tick and its adjoint on the same packed state).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "msim/msim.h"
#include "msim_bin.hpp"
#include "msim_host.hpp"
#include "msim_kernels.hpp"
#include "msim_sorted.hpp"

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
    unsigned long long* ticket;  // k_bbin_fused segment tickets (never reset)
    unsigned long long* status;  // [n_seg] k_bbin_fused published words
    unsigned int* faults;
    float inv_cell;
    uint32_t n_seg;
};

__device__ __forceinline__ TableView scen_view(const ScenDev& sc, const BatchBin& b) {
    return TableView{b.old_start + sc.start_base, b.old_bid, b.inv_cell, sc.lw, sc.wmask, sc.hmask, sc.m,
                     sc.seg_log, sc.margin};
}

template <uint32_t U, uint32_t T>
static __global__ __launch_bounds__(T) void k_bbin_count(const float2* __restrict__ pos, BatchBin b) {
    extern __shared__ uint32_t hist[];
    __shared__ uint32_t lds[T / 64];
    const uint32_t g = blockIdx.x;
    const ScenDev sc = b.scen[b.seg_scen[g]];
    const uint32_t s0 = (g - sc.seg_first) << sc.seg_log;
    const uint32_t total = bin_count_segment<U, T>(pos, scen_view(sc, b), s0, hist, lds, b.new_start + sc.start_base);
    if (threadIdx.x == 0) {
        b.seg_total[g] = total;
    }
}

template <uint32_t U, uint32_t T>
static __global__ __launch_bounds__(T, 8) void k_bbin_place(const float2* __restrict__ pos, BatchBin b) {
    extern __shared__ uint32_t cursor[];
    __shared__ uint32_t lds[T / 64];
    const uint32_t g = blockIdx.x;
    const ScenDev sc = b.scen[b.seg_scen[g]];
    uint32_t before = 0;
    for (uint32_t k = sc.seg_first + threadIdx.x; k < g; k += T) {
        before += b.seg_total[k];
    }
    const uint32_t base = sc.agent_off + block_sum<T>(before, lds);
    if (g == sc.seg_first + sc.n_seg - 1 && threadIdx.x == 0) {
        const uint32_t end = base + b.seg_total[g];
        b.new_start[sc.start_base + sc.m] = end;
        if (end != sc.agent_off + sc.n) {
            atomicOr(b.faults, kFaultBinCount);
        }
    }
    bin_place_segment<U, T>(pos, scen_view(sc, b), (g - sc.seg_first) << sc.seg_log, base, cursor,
                      b.new_start + sc.start_base, b.new_bpos, b.new_bid);
}

// Single-pass rebinning of every scenario's table: segments are handed out
// by ticket, and each scenario's look-back chain starts at its first segment
// (prefix = the scenario's first slot).
template <uint32_t U, uint32_t T>
static __global__ __launch_bounds__(T, 8) void k_bbin_fused(const float2* __restrict__ pos, BatchBin b) {
    extern __shared__ uint32_t hist[];
    __shared__ uint32_t lds[T / 64];
    __shared__ uint32_t sh[3];
    take_segment(b.ticket, b.n_seg, sh);
    const uint32_t g = sh[0];
    const ScenDev sc = b.scen[b.seg_scen[g]];
    uint32_t base, total;
    bin_fused_segment<U, T>(pos, scen_view(sc, b), (g - sc.seg_first) << sc.seg_log, hist, lds, &sh[2],
                      b.new_start + sc.start_base, b.new_bpos, b.new_bid, b.status, g, sc.seg_first, sc.agent_off,
                      sh[1], b.faults, &base, &total);
    if (g == sc.seg_first + sc.n_seg - 1 && threadIdx.x == 0) {
        const uint32_t end = base + total;
        b.new_start[sc.start_base + sc.m] = end;
        if (end != sc.agent_off + sc.n) {
            atomicOr(b.faults, kFaultBinCount);
        }
    }
}

// Sorted move of every scenario's table (see msim_sorted.hpp): segments are
// handed out by ticket, each scenario's look-back chain starts at its first
// segment (prefix = the scenario's first slot).  Dynamic LDS: counters for
// the largest segment.
template <uint32_t U, uint32_t T, bool UNIFORM, bool WRITE_DIR>
static __global__ __launch_bounds__(T) void k_bmove_sorted(SortedMove mv, BatchBin b) {
    extern __shared__ uint32_t hist[];
    __shared__ uint32_t lds[T / 64];
    __shared__ uint32_t nh[2 * kNextReach + 1];
    __shared__ uint32_t sh[1];
    const uint32_t g = xcd_block(blockIdx.x, gridDim.x);
    const ScenDev sc = b.scen[b.seg_scen[g]];
    const uint32_t gl = g - sc.seg_first;
    const NextCount nc{nh, mv.seg_next, gl, sc.n_seg, sc.seg_first, sc.seg_log};
    uint32_t base, total;
    sorted_segment<U, T, UNIFORM, WRITE_DIR>(mv, scen_view(sc, b), nc, gl << sc.seg_log, g, hist, lds, sh,
                                             b.new_start + sc.start_base, &base, &total);
    if (gl == sc.n_seg - 1 && threadIdx.x == 0) {
        const uint32_t end = base + total;
        b.new_start[sc.start_base + sc.m] = end;
        if (end != sc.agent_off + sc.n) {
            atomicOr(b.faults, kFaultBinCount);
        }
    }
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

// Route choice of the arrivals the batched sorted move listed, one
// workgroup per segment, on the segment's own scenario table.
template <bool UNIFORM, bool WRITE_DIR>
static __global__ __launch_bounds__(kBlock) void k_broute_sorted(SortedMove mv, BatchBin b) {
    __shared__ uint32_t nh[2 * kNextReach + 1];
    const uint32_t g = xcd_block(blockIdx.x, gridDim.x);
    const ScenDev sc = b.scen[b.seg_scen[g]];
    const NextCount nc{nh, mv.seg_next, g - sc.seg_first, sc.n_seg, sc.seg_first, sc.seg_log};
    route_segment_list<UNIFORM, WRITE_DIR>(mv, scen_view(sc, b), nc, g);
}

// nxt for every slot of every scenario: the bucket of the agent's next
// position in its own scenario's table; seg_next (cleared before) counts them
// per (global) segment.
template <bool UNIFORM>
static __global__ __launch_bounds__(kBlock) void k_bsorted_next(SortedBufs st, const BlockScen* __restrict__ blk,
                                                                const ScenDev* __restrict__ scen,
                                                                const uint32_t* __restrict__ agent_off, uint32_t n,
                                                                float spd_const, float inv_cell,
                                                                uint32_t* __restrict__ seg_next) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t gseg = 0;
    if (i < n) {
        const BlockScen bs = blk[blockIdx.x];
        uint32_t sid = bs.sid;
        if (i >= bs.end) {
            sid++;
            while (i >= agent_off[sid + 1]) sid++;
        }
        const ScenDev sc = scen[sid];
        const uint4 r = st.rec[i];
        const float2 np = next_position(st.pos[i], make_float2(__uint_as_float(r.x), __uint_as_float(r.y)),
                                        UNIFORM ? spd_const : __uint_as_float(r.z));
        const uint32_t b = bucket_of(cell_coord(np.x, inv_cell), cell_coord(np.y, inv_cell), sc.lw, sc.wmask,
                                     sc.hmask);
        st.nxt[i] = b;
        gseg = sc.seg_first + (b >> sc.seg_log);
    }
    count_next_segment(seg_next, gseg, i < n);
}

// Per-scenario partial sums of a wave's counts: one atomic per wave when
// all its valid slots lie in one scenario, one per slot otherwise.
__device__ __forceinline__ void bpart_add(const BatchCollide& c, uint32_t wave, bool valid, uint32_t sid,
                                          uint32_t cnt) {
    const unsigned long long act = __ballot(valid);
    const int first = act ? __ffsll(static_cast<long long>(act)) - 1 : 0;
    const uint32_t s0 = __shfl(sid, first, 64);
    const bool one = __ballot(valid && sid != s0) == 0ull;
    const uint32_t slot = wave % kScenSlots;
    if (one) {
        unsigned long long v = valid ? cnt : 0u;
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

// The scenario of slot s, given the BlockScen of the 256-slot block holding
// it (a later scenario may start inside the block).
__device__ __forceinline__ uint32_t slot_scen(const BatchCollide& c, const BlockScen& bs, uint32_t s) {
    uint32_t sid = bs.sid;
    if (s >= bs.end) {
        sid++;
        while (s >= c.agent_off[sid + 1]) sid++;
    }
    return sid;
}

// Neighbour pass over every binned slot of every scenario, each slot on its
// own scenario's bucket table, one slot per lane (count_in_cells, as in the
// plain pass).
static __global__ __launch_bounds__(kBlock) void k_bcollide(BatchCollide c) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t s = blk * kBlock + threadIdx.x;
    const bool valid = s < c.n;
    const BlockScen bs = c.blk[blk];
    uint32_t cnt = 0;
    const float2 me = valid ? c.bpos[s] : make_float2(0.0f, 0.0f);
    const int32_t cx = cell_coord(me.x, c.inv_cell);
    const int32_t cy = cell_coord(me.y, c.inv_cell);
    const uint32_t sid = valid ? slot_scen(c, bs, s) : bs.sid;
    if (valid) {
        const ScenDev& sc = c.scen[sid];
        cnt = count_in_cells(me, s, cx, cy, c.start + sc.start_base, sc.lw, sc.wmask, sc.hmask, c.bpos,
                             c.d2_limit) - 1u;  // own slot
        c.ncount[c.bid ? c.bid[s] : s] = cnt;  // no slot map: the state is in slot order
    }
    bpart_add(c, (blk * kBlock + threadIdx.x) >> 6, valid, sid, cnt);
}

// count_in_cells on a slot's own scenario table (the few waves that
// straddle a scenario boundary).
__device__ __forceinline__ uint32_t bcount_slot(const BatchCollide& c, uint32_t sid, float2 me, uint32_t s, int32_t cx,
                                             int32_t cy) {
    const ScenDev& sc = c.scen[sid];
    return count_in_cells(me, s, cx, cy, c.start + sc.start_base, sc.lw, sc.wmask, sc.hmask, c.bpos, c.d2_limit);
}

// Staged neighbour pass: each wave takes 64 P consecutive slots, P per
// lane; a wave whose slots all lie in one scenario runs wave_count_in_cells
// on that scenario's table, a wave across a scenario boundary walks per slot
// (grid: ceil(n / (kBlock P)) workgroups).
template <int P>
static __global__ __launch_bounds__(kBlock, P == 1 ? 8 : 6) void k_bcollide_staged(BatchCollide c) {
    __shared__ float2 stage_all[kBlock / 64 * kStage];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t wave = blk * (kBlock / 64) + (threadIdx.x >> 6);
    const uint32_t w0 = wave * (64u * P);
    if (w0 >= c.n) {
        return;  // wave-uniform: no barrier follows
    }
    const uint32_t lane = __lane_id();
    // the wave's 64 P slots lie in one 256-slot block, whose entry names the
    // scenario of its first slot and where that scenario ends
    const BlockScen bs = c.blk[w0 / kBlock];
    float2 me[P];
    bool valid[P];
    int32_t cx[P], cy[P];
    uint32_t id[P], cnt[P];
    bool later = false;  // some valid slot lies past scenario bs.sid
#pragma unroll
    for (int p = 0; p < P; p++) {
        const uint32_t s = w0 + 64u * p + lane;
        valid[p] = s < c.n;
        me[p] = valid[p] ? c.bpos[s] : make_float2(0.0f, 0.0f);
        id[p] = valid[p] ? (c.bid ? c.bid[s] : s) : 0u;  // no slot map: the state is in slot order
        cx[p] = cell_coord(me[p].x, c.inv_cell);
        cy[p] = cell_coord(me[p].y, c.inv_cell);
        later = later || (valid[p] && s >= bs.end);
        cnt[p] = 0u;
    }
    const bool one = __ballot(later) == 0ull;
    if (one) {
        const uint32_t last = min(64u * P, c.n - w0) - 1u;
        wave_count_in_cells<P>(me, valid, 0u, 0u, last >> 6, last & 63u, cx, cy, c.start + bs.start_base, bs.lw,
                               bs.wmask, bs.hmask, c.bpos, c.d2_limit, stage_all + (threadIdx.x >> 6) * kStage,
                               cnt);
    }
#pragma unroll
    for (int p = 0; p < P; p++) {
        const uint32_t s = w0 + 64u * p + lane;
        const uint32_t sid = (one || !valid[p]) ? bs.sid : slot_scen(c, bs, s);
        if (valid[p]) {
            if (!one) {
                cnt[p] = bcount_slot(c, sid, me[p], s, cx[p], cy[p]);
            }
            cnt[p] -= 1u;  // own slot
            c.ncount[id[p]] = cnt[p];
        }
        bpart_add(c, wave * P + p, valid[p], sid, cnt[p]);
    }
}

// One wave per scenario: fold the slots into the scenario's pair totals.
static __global__ __launch_bounds__(64) void k_bfinish(unsigned long long* __restrict__ part,
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

using msim::AgentsDev;
using msim::BlockScen;
using msim::ScenDev;
using msim::kShaderSpeed;

struct BBuffers {
    float2 *pos = nullptr, *tgt = nullptr, *dir = nullptr;
    float* spd = nullptr;
    uint32_t *road = nullptr, *init = nullptr, *ncount = nullptr, *perm = nullptr;
    uint4* rng = nullptr;
};

struct DiffState;  // msim_diff.hip: parameters, checkpoints and tape
struct msim_batch;
void diff_release(msim_batch* b);  // msim_diff.hip
namespace msim {
int bleave_sorted(msim_batch* b);  // msim_batch.hip
}  // namespace msim

struct msim_batch {
    hipStream_t stream = nullptr;
    uint32_t n_scen = 0, n = 0, n_roads = 0, n_conns = 0, n_seg = 0, start_len = 0, max_seg_log = 9;
    std::vector<uint32_t> agent_off_h, road_off_h;
    msim_road* roads = nullptr;
    uint32_t* conns = nullptr;
    BBuffers buf[2];
    int cur = 0;
    bool uniform = true;
    float inv_cell = 1.0f, d2_limit = 1.0f, radius = 1.0f;
    ScenDev* scen = nullptr;
    uint32_t* agent_off = nullptr;
    BlockScen* blk_scen = nullptr;
    uint32_t* seg_scen = nullptr;
    uint32_t* seg_total = nullptr;
    unsigned long long* bin_ticket = nullptr;  // [1]
    unsigned long long* seg_status = nullptr;  // [n_seg]
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
    bool capturing = false;  // inside bbuild_graph: renumberings skip the output-only arrays
    int collide_mode = 0;  // 0: staged, 2 slots per lane; 1: per lane; 2: staged, 1 slot per lane
    int binning = 0;  // 0: k_bbin_count + k_bbin_place; 2: k_bbin_fused; 3: sorted state (k_bmove_sorted)
    // Sorted state (binning 3, from creation on): pos, tgt, spd, perm, dir and
    // ncount of buf[cur] follow the slot order of binning bcur, which flips
    // with cur on every move; road and PRNG state stay in packed caller order.
    bool sorted = false;
    uint32_t* nxt[2] = {nullptr, nullptr};  // per buffer: bucket of each slot's next position
    uint4* rec[2] = {nullptr, nullptr};     // per buffer: each slot's target, speed and caller index
    uint4* cold = nullptr;                  // [2n] packed caller order: cold_word, PRNG state
    uint4* arr_rec = nullptr;               // [n] arrival records, listed per segment
    float* arr_spd = nullptr;               // [n] their speeds (per-agent speeds only)
    uint32_t* seg_start = nullptr;          // [n_seg + 1] first slot of each segment (k_seg_scan)
    uint32_t* seg_cnt[2] = {nullptr, nullptr};  // per buffer: agents per segment for the move leaving it
    uint32_t* seg_arr = nullptr;            // [n_seg] arrivals of each segment
    uint32_t bin_unroll = 4;  // window slots per rebinning thread and load round
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
    DiffState* diff = nullptr;

    AgentsDev agents(int k) const {
        const BBuffers& b = buf[k];
        return AgentsDev{n, b.pos, b.tgt, b.dir, uniform ? nullptr : b.spd, kShaderSpeed,
                         b.road, b.rng, b.init, b.ncount, b.perm};
    }
};

namespace msim {

// failure message of the last call on this thread (creation failures)
extern thread_local std::string g_batch_error;

inline int bfail(msim_batch* b, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    g_batch_error = buf;
    return code;
}

inline int balloc(msim_batch* b, void** p, size_t bytes) {
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
inline int balloc_n(msim_batch* b, T** p, size_t count) {
    return balloc(b, reinterpret_cast<void**>(p), count * sizeof(T));
}

inline uint32_t bblocks(uint32_t n) { return std::max(1u, (n + kBlock - 1) / kBlock); }

inline int bcheck(msim_batch* b, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, MSIM_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return MSIM_OK;
}

// Launch order of a batched tick, for msim_batch_time_ticks.
// Launches of a batched tick, in order, by kernel name (template arguments
// dropped).
constexpr int kBatchKernelMax = 5;  // largest batch_kernel_count
inline int batch_kernel_count(const msim_batch* b) { return b->sorted ? 5 : (b->binning == 2 ? 4 : 5); }
inline const char* batch_kernel_name(const msim_batch* b, int k) {
    if (b->sorted) {
        if (k == 0) return "k_seg_scan";
        if (k == 1) return "k_bmove_sorted";
        if (k == 2) return "k_broute_sorted";
        return k == 3 ? (b->collide_mode == 1 ? "k_bcollide" : "k_bcollide_staged") : "k_bfinish";
    }
    if (k == 0) return "k_advance4";
    if (b->binning == 2) {
        if (k == 1) return "k_bbin_fused";
        k++;
    } else if (k <= 2) {
        return k == 1 ? "k_bbin_count" : "k_bbin_place";
    }
    if (k == 3) return b->collide_mode == 1 ? "k_bcollide" : "k_bcollide_staged";
    return "k_bfinish";
}

inline void bprobe_pre(msim_batch* b) {
    if (b->probe_kernel == b->launch_no) (void)hipEventRecord(b->probe_ev[0], b->stream);
}

inline void bprobe_post(msim_batch* b) {
    if (b->probe_kernel == b->launch_no) (void)hipEventRecord(b->probe_ev[1], b->stream);
    b->launch_no++;
}

// Segment rebinning of every scenario: one pass (k_bbin_fused) or count +
// place.
template <uint32_t U, uint32_t T>
inline void blaunch_bins(msim_batch* b, const float2* pos, const BatchBin& bb, size_t lds) {
    if (b->binning == 2) {
        bprobe_pre(b);
        hipLaunchKernelGGL((k_bbin_fused<U, T>), dim3(b->n_seg), dim3(T), lds, b->stream, pos, bb);
        bprobe_post(b);
    } else {
        bprobe_pre(b);
        hipLaunchKernelGGL((k_bbin_count<U, T>), dim3(b->n_seg), dim3(T), lds, b->stream, pos, bb);
        bprobe_post(b);
        bprobe_pre(b);
        hipLaunchKernelGGL((k_bbin_place<U, T>), dim3(b->n_seg), dim3(T), lds, b->stream, pos, bb);
        bprobe_post(b);
    }
}

// The rebinning launch for the handle's load batching.  Workgroups keep
// kBinThreads threads: every scenario's segments (down to 2^9 buckets) run
// in the same launch.
inline void blaunch_bins(msim_batch* b, const float2* pos, const BatchBin& bb, size_t lds) {
    if (b->bin_unroll == 8) {
        blaunch_bins<8, kBinThreads>(b, pos, bb, lds);
    } else {
        blaunch_bins<4, kBinThreads>(b, pos, bb, lds);
    }
}

inline int bwait(msim_batch* b) {
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

inline GridDev bgrid(const msim_batch* b) {
    GridDev g{};
    g.inv_cell = b->inv_cell;
    g.d2_limit = b->d2_limit;
    g.bid = b->bbid[b->bcur];
    return g;
}

inline int benqueue_collide(msim_batch* b) {
    BatchCollide c{};
    c.blk = b->blk_scen;
    c.scen = b->scen;
    c.agent_off = b->agent_off;
    c.n = b->n;
    c.start = b->bstart[b->bcur];
    // sorted state: positions read in place, counts stored by slot
    c.bpos = b->sorted ? b->buf[b->cur].pos : b->bbpos[b->bcur];
    c.bid = b->sorted ? nullptr : b->bbid[b->bcur];
    c.ncount = b->buf[b->cur].ncount;
    c.inv_cell = b->inv_cell;
    c.d2_limit = b->d2_limit;
    c.part = b->part;
    bprobe_pre(b);
    if (b->collide_mode == 1) {
        hipLaunchKernelGGL(k_bcollide, dim3(bblocks(b->n)), dim3(kBlock), 0, b->stream, c);
    } else if (b->collide_mode == 2) {
        hipLaunchKernelGGL(k_bcollide_staged<1>, dim3(bblocks(b->n)), dim3(kBlock), 0, b->stream, c);
    } else {
        const uint32_t per = kBlock * 2u;
        hipLaunchKernelGGL(k_bcollide_staged<2>, dim3(std::max(1u, (b->n + per - 1) / per)), dim3(kBlock), 0,
                           b->stream, c);
    }
    bprobe_post(b);
    bprobe_pre(b);
    hipLaunchKernelGGL(k_bfinish, dim3(b->n_scen), dim3(64), 0, b->stream, b->part, b->totals, b->last);
    bprobe_post(b);
    return bcheck(b, "batched collision dispatch");
}

}  // namespace msim
