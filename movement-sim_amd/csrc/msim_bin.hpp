// msim_bin.hpp — the fast move dispatch.
//
//   k_advance   streams pos/target/speed once.  Agents that keep walking get
//               pos += heading; arriving agents snap to their target
//               (random_move.comp:844-851) and are appended to one of
//               kQueues arrival queues.  With ATOMIC_BIN it also counts the
//               agent into its bucket (global atomics, the atomic path).
//   k_route     route choice (random_move.comp:778-828) and the new heading
//               for the queued arrivals only, so the three dependent gathers
//               of the arrival path no longer stall the streaming kernel.
//   k_bin_count / k_bin_place  (segment path) rebuild the bucket table
//               without global atomics: the toroidal table is cut into n_seg
//               segments of S buckets and workgroup `seg` owns segment seg.
//               An agent moves at most D cells per axis per tick, so in the
//               previous tick's bucket order it sits within `margin` buckets
//               of its new bucket; each workgroup scans that window of the
//               previous binning, counts the agents landing in its segment
//               with LDS atomics and scans the counts (k_bin_count), then
//               adds the segment's global offset (sum of the preceding
//               segments' totals) and places its agents (k_bin_place).
#pragma once

#include "msim_kernels.hpp"

namespace msim {

constexpr uint32_t kBinThreads = 512;

// Rebinning workgroup size: segments of 2^14 buckets hold 64 KB of LDS
// counters, so at most two workgroups share a CU; 1024 threads then keep
// 32 waves per CU resident, as 512-thread groups do at smaller segments.
inline uint32_t bin_threads_for(uint32_t seg_log) { return seg_log >= 14 ? 1024u : kBinThreads; }
constexpr uint32_t kQueues = 64;  // arrival queues, one counter per 128-B line
constexpr uint32_t kQueueStride = 32;  // counter stride in words

struct QueueDev {
    uint32_t* items;       // [kQueues * cap] agent indices
    uint32_t* count;       // this tick's counters, at [q * kQueueStride]
    uint32_t* count_next;  // next tick's counters (zeroed by k_advance)
    uint32_t cap;          // capacity per queue
};

// ROUTE_INLINE: arrivals take the route choice in this kernel (no queue, no
// k_route launch); better when arrivals are cheap to gather (small maps).
template <bool UNIFORM, bool WRITE_DIR, bool ATOMIC_BIN, bool ROUTE_INLINE>
static __global__ __launch_bounds__(kBlock) void k_advance(AgentsDev a, GridDev g, QueueDev q,
                                                    unsigned long long* __restrict__ part_stats,
                                                    const msim_road* __restrict__ roads,
                                                    const uint32_t* __restrict__ conns) {
    __shared__ uint32_t lds[2 * (kBlock / 64)];
    __shared__ uint32_t sh_base;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = blk * kBlock + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x < kQueues) {
        q.count_next[threadIdx.x * kQueueStride] = 0u;
    }
    int inserted = 0;
    int arrived = 0;
    uint32_t bkey = 0;
    if (i < a.n) {
        const float2 p = a.pos[i];
        const float2 t = a.tgt[i];
        const float s = UNIFORM ? a.spd_const : a.spd[i];
        // update_direction (:830-839) and move's distance (:842): the two
        // lengths are bit-identical, (p - t)^2 == (t - p)^2.
        const float dx = t.x - p.x;
        const float dy = t.y - p.y;
        const float len = vlen(dx, dy);
        float2 np;
        if (len > s) {  // len > 0 here, so the heading is the normalised one
            const float2 d = make_float2((dx / len) * s, (dy / len) * s);
            np = make_float2(p.x + d.x, p.y + d.y);
            if (WRITE_DIR) {
                a.dir[i] = d;
            }
        } else {
            np = t;
            arrived = 1;
            if (ROUTE_INLINE) {
                float2 nt = t;
                uint32_t r = a.road[i];
                uint4 st = a.rng[i];
                choose_route(nt, r, st, roads, conns);
                a.tgt[i] = nt;
                a.road[i] = r;
                a.rng[i] = st;
                if (WRITE_DIR) {
                    a.dir[i] = heading(nt, np, s);
                }
            }
        }
        a.pos[i] = np;
        inserted = static_cast<int>(cell_coord(np.x, g.inv_cell) != cell_coord(p.x, g.inv_cell)) |
                   static_cast<int>(cell_coord(np.y, g.inv_cell) != cell_coord(p.y, g.inv_cell));
        if (ATOMIC_BIN) {
            bkey = bucket_of(cell_coord(np.x, g.inv_cell), cell_coord(np.y, g.inv_cell), g.lw, g.wmask,
                             g.hmask);
        }
    }
    if (ATOMIC_BIN) {
        const uint32_t r = wave_rank_add(g.cnt, bkey, i < a.n);
        if (i < a.n) {
            g.rank[i] = r;
        }
    }
    // block-aggregated append of the arrivals to queue blk % kQueues
    const unsigned long long am = __ballot(arrived);
    const unsigned long long im = __ballot(inserted);
    const uint32_t w = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        lds[w] = static_cast<uint32_t>(__popcll(am));
        lds[kBlock / 64 + w] = static_cast<uint32_t>(__popcll(im));
    }
    __syncthreads();
    uint32_t before = 0, tot_a = 0, tot_i = 0;
#pragma unroll
    for (uint32_t k = 0; k < kBlock / 64; k++) {
        before += (k < w) ? lds[k] : 0u;
        tot_a += lds[k];
        tot_i += lds[kBlock / 64 + k];
    }
    const uint32_t qi = blk % kQueues;
    if (threadIdx.x == 0) {
        sh_base = (tot_a && !ROUTE_INLINE) ? atomicAdd(&q.count[qi * kQueueStride], tot_a) : 0u;
        slot_add(part_stats, blk, kWordInserts, tot_i);
        slot_add(part_stats, blk, kWordArrivals, tot_a);
    }
    __syncthreads();
    if (!ROUTE_INLINE && arrived) {
        const unsigned long long lt = (1ull << __lane_id()) - 1ull;
        const uint32_t slot = sh_base + before + static_cast<uint32_t>(__popcll(am & lt));
        q.items[qi * q.cap + slot] = i;
    }
}

// Streaming advance with kAdv agents per thread: 16-byte loads and stores of
// the SoA arrays (the kernels are bound by vector-memory instruction issue,
// not bytes), arrivals compacted per wave in LDS and routed with the wave's
// lanes in parallel.  Per agent the arithmetic is k_advance's exactly.  The
// per-agent arrays are padded to a multiple of kAdv agents; the padding is
// computed on but never routed, counted or read back.
constexpr uint32_t kAdv = 4;

template <bool UNIFORM, bool WRITE_DIR>
static __global__ __launch_bounds__(kBlock) void k_advance4(AgentsDev a, float inv_cell,
                                                     unsigned long long* __restrict__ part_stats,
                                                     const msim_road* __restrict__ roads,
                                                     const uint32_t* __restrict__ conns) {
    __shared__ uint32_t arr_id[kBlock / 64][64 * kAdv];
    __shared__ float2 arr_dir[WRITE_DIR ? kBlock / 64 : 1][WRITE_DIR ? 64 * kAdv : 1];
    __shared__ uint32_t lds[2 * (kBlock / 64)];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t t = blk * kBlock + threadIdx.x;
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t i0 = t * kAdv;
    const bool any = i0 < a.n;
    float2 np[kAdv], d[kAdv];
    bool arr[kAdv];
    uint32_t n_ins = 0;
#pragma unroll
    for (uint32_t u = 0; u < kAdv; u++) {
        arr[u] = false;
        d[u] = make_float2(0.0f, 0.0f);
        np[u] = d[u];
    }
    if (any) {
        const float4* P = reinterpret_cast<const float4*>(a.pos) + 2 * t;
        const float4* G = reinterpret_cast<const float4*>(a.tgt) + 2 * t;
        const float4 p01 = P[0], p23 = P[1], t01 = G[0], t23 = G[1];
        const float4 sv = UNIFORM ? make_float4(a.spd_const, a.spd_const, a.spd_const, a.spd_const)
                                  : reinterpret_cast<const float4*>(a.spd)[t];
        const float2 p[kAdv] = {make_float2(p01.x, p01.y), make_float2(p01.z, p01.w), make_float2(p23.x, p23.y),
                                make_float2(p23.z, p23.w)};
        const float2 tg[kAdv] = {make_float2(t01.x, t01.y), make_float2(t01.z, t01.w),
                                 make_float2(t23.x, t23.y), make_float2(t23.z, t23.w)};
        const float sp[kAdv] = {sv.x, sv.y, sv.z, sv.w};
#pragma unroll
        for (uint32_t u = 0; u < kAdv; u++) {
            const bool valid = i0 + u < a.n;
            const float dx = tg[u].x - p[u].x;
            const float dy = tg[u].y - p[u].y;
            const float len = vlen(dx, dy);
            if (len > sp[u]) {
                d[u] = make_float2((dx / len) * sp[u], (dy / len) * sp[u]);
                np[u] = make_float2(p[u].x + d[u].x, p[u].y + d[u].y);
            } else {
                np[u] = tg[u];
                arr[u] = valid;
            }
            n_ins += (valid && (cell_coord(np[u].x, inv_cell) != cell_coord(p[u].x, inv_cell) ||
                                cell_coord(np[u].y, inv_cell) != cell_coord(p[u].y, inv_cell)))
                         ? 1u
                         : 0u;
        }
    }
    // the wave's arrivals, in (u, lane) order
    const unsigned long long lt = (1ull << __lane_id()) - 1ull;
    uint32_t n_arr = 0;
    uint32_t slot[kAdv];
#pragma unroll
    for (uint32_t u = 0; u < kAdv; u++) {
        const unsigned long long m = __ballot(arr[u]);
        slot[u] = n_arr + static_cast<uint32_t>(__popcll(m & lt));
        n_arr += static_cast<uint32_t>(__popcll(m));
        if (arr[u]) {
            arr_id[w][slot[u]] = i0 + u;
        }
    }
    wave_lds_sync();
    for (uint32_t j = __lane_id(); j < n_arr; j += 64) {
        const uint32_t i = arr_id[w][j];
        float2 nt = a.tgt[i];  // the old target, where the agent now stands
        const float2 at = nt;
        uint32_t r = a.road[i];
        uint4 st = a.rng[i];
        choose_route(nt, r, st, roads, conns);
        a.tgt[i] = nt;
        a.road[i] = r;
        a.rng[i] = st;
        if (WRITE_DIR) {
            arr_dir[w][j] = heading(nt, at, UNIFORM ? a.spd_const : a.spd[i]);
        }
    }
    if (any) {
        if (WRITE_DIR) {
            wave_lds_sync();
#pragma unroll
            for (uint32_t u = 0; u < kAdv; u++) {
                if (arr[u]) {
                    d[u] = arr_dir[w][slot[u]];
                }
            }
            float4* D = reinterpret_cast<float4*>(a.dir) + 2 * t;
            D[0] = make_float4(d[0].x, d[0].y, d[1].x, d[1].y);
            D[1] = make_float4(d[2].x, d[2].y, d[3].x, d[3].y);
        }
        float4* P = reinterpret_cast<float4*>(a.pos) + 2 * t;
        P[0] = make_float4(np[0].x, np[0].y, np[1].x, np[1].y);
        P[1] = make_float4(np[2].x, np[2].y, np[3].x, np[3].y);
    }
    // block totals of insertions and arrivals, one slot update each
    const uint32_t wi = wave_sum(n_ins);
    if (__lane_id() == 0) {
        lds[w] = wi;
        lds[kBlock / 64 + w] = n_arr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t si = 0, sa = 0;
#pragma unroll
        for (uint32_t k = 0; k < kBlock / 64; k++) {
            si += lds[k];
            sa += lds[kBlock / 64 + k];
        }
        slot_add(part_stats, blk, kWordInserts, si);
        slot_add(part_stats, blk, kWordArrivals, sa);
    }
}

// Route choice for the queued arrivals; grid = kQueues * blocks_per_queue.
template <bool UNIFORM, bool WRITE_DIR>
static __global__ __launch_bounds__(kBlock) void k_route(AgentsDev a, QueueDev q, const msim_road* __restrict__ roads,
                                                  const uint32_t* __restrict__ conns) {
    const uint32_t qi = blockIdx.x % kQueues;
    const uint32_t per_queue = gridDim.x / kQueues;
    const uint32_t n_items = q.count[qi * kQueueStride];
    for (uint32_t j = (blockIdx.x / kQueues) * kBlock + threadIdx.x; j < n_items; j += per_queue * kBlock) {
        const uint32_t i = q.items[qi * q.cap + j];
        float2 t = a.tgt[i];  // the old target, where the agent now stands
        const float2 np = t;
        uint32_t r = a.road[i];
        uint4 st = a.rng[i];
        choose_route(t, r, st, roads, conns);
        a.tgt[i] = t;
        a.road[i] = r;
        a.rng[i] = st;
        if (WRITE_DIR) {
            const float s = UNIFORM ? a.spd_const : a.spd[i];
            a.dir[i] = heading(t, np, s);
        }
    }
}

struct BinDev {
    const uint32_t* old_start;  // [m+1] previous binning
    const uint32_t* old_bid;    // [n]
    uint32_t* new_start;        // [m+1]
    float2* new_bpos;           // [n]
    uint32_t* new_bid;          // [n]
    uint32_t* seg_total;        // [n_seg]
    unsigned long long* ticket;  // [1] single-pass rebinning: segments handed out in start order
    unsigned long long* status;  // [n_seg] single-pass rebinning: published totals / prefixes
    Counters* ctr;
    float inv_cell;
    uint32_t lw, wmask, hmask, m;
    uint32_t seg_log, n_seg, margin;
    uint32_t n;
};

// One bucket table region and the previous binning of its agents.
struct TableView {
    const uint32_t* old_start;  // [m+1] slot offsets (absolute)
    const uint32_t* old_bid;    // slot -> internal agent
    float inv_cell;
    uint32_t lw, wmask, hmask, m, seg_log, margin;
};

__device__ __forceinline__ TableView table_view(const BinDev& b) {
    return TableView{b.old_start, b.old_bid, b.inv_cell, b.lw, b.wmask, b.hmask, b.m, b.seg_log, b.margin};
}

// Old-slot ranges [a0, a1) and [c0, c1) holding every agent whose previous
// bucket lies in the cyclic window [s0 - margin, s0 + S + margin).
__device__ __forceinline__ void bin_window(const TableView& t, uint32_t s0, uint32_t& a0, uint32_t& a1,
                                           uint32_t& c0, uint32_t& c1) {
    const uint32_t S = 1u << t.seg_log;
    c0 = 0;
    c1 = 0;
    if (2ull * t.margin + S >= t.m) {
        a0 = t.old_start[0];
        a1 = t.old_start[t.m];
        return;
    }
    const uint32_t lo = (s0 + t.m - t.margin) & (t.m - 1);
    const uint32_t hi = (s0 + S + t.margin) & (t.m - 1);
    const bool wrap = lo >= hi;  // the window wraps around the table's end
    const uint32_t v_lo = t.old_start[lo], v_hi = t.old_start[hi], v_end = t.old_start[t.m];
    const uint32_t v_0 = t.old_start[0];
    a0 = v_lo;
    a1 = wrap ? v_end : v_hi;
    c0 = wrap ? v_0 : 0u;
    c1 = wrap ? v_hi : 0u;
}

__device__ __forceinline__ uint32_t bucket_of_pos(float2 p, const TableView& t) {
    return bucket_of(cell_coord(p.x, t.inv_cell), cell_coord(p.y, t.inv_cell), t.lw, t.wmask, t.hmask);
}


// One pass over the window slots [lo, hi): every agent whose new bucket lies
// in the segment bumps its LDS counter (PLACE: and is written to the slot the
// counter hands out).  The trip count is block-uniform, so the wave-level rank
// helper always runs with the whole wave; U independent loads per
// thread keep enough requests in flight.
template <bool PLACE, uint32_t U, uint32_t T>
__device__ __forceinline__ void bin_pass(const float2* __restrict__ pos, const TableView& t, uint32_t s0,
                                         uint32_t lo, uint32_t hi, uint32_t* table, float2* __restrict__ out_pos,
                                         uint32_t* __restrict__ out_bid) {
    const uint32_t S = 1u << t.seg_log;
    for (uint32_t base = lo; base < hi; base += U * T) {
        uint32_t id[U];
        float2 p[U];
        // unconditional loads (past hi they re-read slot hi - 1), so that all
        // of them are in flight before the first is used
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t k = base + u * T + threadIdx.x;
            id[u] = t.old_bid[min(k, hi - 1u)];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            p[u] = pos[id[u]];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t k = base + u * T + threadIdx.x;
            const uint32_t off = bucket_of_pos(p[u], t) - s0;
            const bool mine = (k < hi) && (off < S);
            const uint32_t slot = wave_rank_add(table, off, mine);
            if (PLACE && mine) {
                out_pos[slot] = p[u];
                out_bid[slot] = id[u];
            }
        }
    }
}

// Segment count: window agents landing in [s0, s0 + S) counted in LDS, local
// exclusive offsets written to start_out[s0 + j]; returns the segment total.
template <uint32_t U, uint32_t T>
__device__ __forceinline__ uint32_t bin_count_segment(const float2* __restrict__ pos, const TableView& t,
                                                      uint32_t s0, uint32_t* hist, uint32_t* lds,
                                                      uint32_t* __restrict__ start_out) {
    const uint32_t tid = threadIdx.x;
    const uint32_t S = 1u << t.seg_log;
    for (uint32_t j = tid; j < S; j += T) {
        hist[j] = 0u;
    }
    __syncthreads();
    uint32_t a0, a1, c0, c1;
    bin_window(t, s0, a0, a1, c0, c1);
    bin_pass<false, U, T>(pos, t, s0, a0, a1, hist, nullptr, nullptr);
    bin_pass<false, U, T>(pos, t, s0, c0, c1, hist, nullptr, nullptr);
    __syncthreads();
    const uint32_t per = S / T;
    uint32_t run = 0;
    for (uint32_t q = 0; q < per; q++) {
        run += hist[tid * per + q];
    }
    uint32_t total = 0;
    uint32_t excl = block_excl_scan<T>(run, lds, total);
    for (uint32_t q = 0; q < per; q++) {  // counts -> segment-local exclusive offsets
        const uint32_t v = hist[tid * per + q];
        hist[tid * per + q] = excl;
        excl += v;
    }
    __syncthreads();
    for (uint32_t j = tid; j < S; j += T) {  // stored in bucket order (whole lines per wave)
        start_out[s0 + j] = hist[j];
    }
    return total;
}

// Segment place: start_out[s0 + j] += base (absolute slots), agents placed.
template <uint32_t U, uint32_t T>
__device__ __forceinline__ void bin_place_segment(const float2* __restrict__ pos, const TableView& t, uint32_t s0,
                                                  uint32_t base, uint32_t* cursor, uint32_t* __restrict__ start_out,
                                                  float2* __restrict__ out_pos, uint32_t* __restrict__ out_bid) {
    const uint32_t S = 1u << t.seg_log;
    for (uint32_t j = threadIdx.x; j < S; j += T) {
        const uint32_t v = base + start_out[s0 + j];
        cursor[j] = v;
        start_out[s0 + j] = v;
    }
    __syncthreads();
    uint32_t a0, a1, c0, c1;
    bin_window(t, s0, a0, a1, c0, c1);
    bin_pass<true, U, T>(pos, t, s0, a0, a1, cursor, out_pos, out_bid);
    bin_pass<true, U, T>(pos, t, s0, c0, c1, cursor, out_pos, out_bid);
}

// ---- single-pass rebinning: count, decoupled look-back, place -------------
// One workgroup per segment counts its window into LDS, publishes the
// segment total, finds the sum of the preceding segments' totals by looking
// back over their published words, and places its agents with a second
// pass over the same window (which its own count pass has just brought
// into the L2 of its XCD).  Segments are handed out by an atomic ticket in
// workgroup start order, so a predecessor has always started: a workgroup
// only waits for segments that run or have run, never for later ones.
// Tickets never reset: launch L's tickets are [L * n_seg, (L + 1) * n_seg),
// and the launch number (the epoch, 30 bits) tags every published word, so
// a stale word from an earlier launch is never read as current.
// Word: [63:34] epoch, [33:32] 1 = segment total only, 2 = inclusive prefix,
// [31:0] value.
constexpr uint32_t kEpochMask = (1u << 30) - 1u;
constexpr uint32_t kSegAggregate = 1u;
constexpr uint32_t kSegPrefix = 2u;
constexpr uint32_t kLookbackPolls = 1u << 20;  // about a second of polling, then a fault is raised

__device__ __forceinline__ unsigned long long seg_word(uint32_t epoch, uint32_t flag, uint32_t v) {
    return (static_cast<unsigned long long>(epoch) << 34) | (static_cast<unsigned long long>(flag) << 32) | v;
}

__device__ __forceinline__ void seg_publish(unsigned long long* w, unsigned long long v) {
    __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Thread-level: publishes segment g's total and returns `first` plus the
// totals of segments [g_first, g).  Prefix words hold absolute values
// (`first` included); segment g_first publishes its prefix at once, so every
// walk ends there at the latest.
__device__ __forceinline__ uint32_t seg_lookback(unsigned long long* status, uint32_t g, uint32_t g_first,
                                                 uint32_t first, uint32_t total, uint32_t epoch,
                                                 unsigned int* faults) {
    if (g == g_first) {
        seg_publish(&status[g], seg_word(epoch, kSegPrefix, first + total));
        return first;
    }
    seg_publish(&status[g], seg_word(epoch, kSegAggregate, total));
    uint32_t excl = 0;
    uint32_t k = g - 1u;
    uint32_t polls = 0;
    for (;;) {
        const unsigned long long w = __hip_atomic_load(&status[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t flag = static_cast<uint32_t>(w >> 32) & 3u;
        if (static_cast<uint32_t>(w >> 34) == epoch && flag != 0u) {
            excl += static_cast<uint32_t>(w);  // a prefix word is absolute (it includes `first`)
            if (flag == kSegPrefix) {
                break;
            }
            if (k == g_first) {  // cannot happen: g_first publishes its prefix directly
                atomicOr(faults, kFaultLookback);
                break;
            }
            k--;
            continue;
        }
        if (++polls >= kLookbackPolls) {
            atomicOr(faults, kFaultLookback);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    seg_publish(&status[g], seg_word(epoch, kSegPrefix, excl + total));
    return excl;
}

// Workgroup-level: the segment's ticket (thread 0 takes it) as (segment,
// epoch) in sh[0], sh[1].
__device__ __forceinline__ void take_segment(unsigned long long* ticket, uint32_t n_seg, uint32_t* sh) {
    if (threadIdx.x == 0) {
        const unsigned long long t = atomicAdd(ticket, 1ull);
        sh[0] = static_cast<uint32_t>(t % n_seg);
        sh[1] = static_cast<uint32_t>(t / n_seg) & kEpochMask;
    }
    __syncthreads();
}

// Count, look back, place.  Writes the segment's absolute bucket offsets to
// start_out[s0 + j]; *base_out / *total_out receive the segment's first slot
// and agent count (all threads).
template <uint32_t U, uint32_t T>
__device__ __forceinline__ void bin_fused_segment(const float2* __restrict__ pos, const TableView& t, uint32_t s0,
                                                  uint32_t* hist, uint32_t* lds, uint32_t* sh_base,
                                                  uint32_t* __restrict__ start_out, float2* __restrict__ out_pos,
                                                  uint32_t* __restrict__ out_bid, unsigned long long* status,
                                                  uint32_t g, uint32_t g_first, uint32_t first, uint32_t epoch,
                                                  unsigned int* faults, uint32_t* base_out, uint32_t* total_out) {
    const uint32_t tid = threadIdx.x;
    const uint32_t S = 1u << t.seg_log;
    for (uint32_t j = tid; j < S; j += T) {
        hist[j] = 0u;
    }
    __syncthreads();
    uint32_t a0, a1, c0, c1;
    bin_window(t, s0, a0, a1, c0, c1);
    bin_pass<false, U, T>(pos, t, s0, a0, a1, hist, nullptr, nullptr);
    bin_pass<false, U, T>(pos, t, s0, c0, c1, hist, nullptr, nullptr);
    __syncthreads();
    const uint32_t per = S / T;
    uint32_t run = 0;
    for (uint32_t q = 0; q < per; q++) {
        run += hist[tid * per + q];
    }
    uint32_t total = 0;
    uint32_t excl = block_excl_scan<T>(run, lds, total);
    for (uint32_t q = 0; q < per; q++) {  // counts -> segment-local exclusive offsets
        const uint32_t v = hist[tid * per + q];
        hist[tid * per + q] = excl;
        excl += v;
    }
    if (tid == 0) {
        *sh_base = seg_lookback(status, g, g_first, first, total, epoch, faults);
    }
    __syncthreads();
    const uint32_t base = *sh_base;
    for (uint32_t j = tid; j < S; j += T) {
        const uint32_t v = base + hist[j];
        hist[j] = v;
        start_out[s0 + j] = v;
    }
    __syncthreads();
    bin_pass<true, U, T>(pos, t, s0, a0, a1, hist, out_pos, out_bid);
    bin_pass<true, U, T>(pos, t, s0, c0, c1, hist, out_pos, out_bid);
    *base_out = base;
    *total_out = total;
}

// Single-pass rebinning of one table (grid: n_seg workgroups).
template <uint32_t U, uint32_t T>
static __global__ __launch_bounds__(T, 8) void k_bin_fused(const float2* __restrict__ pos, BinDev b) {
    extern __shared__ uint32_t hist[];
    __shared__ uint32_t lds[T / 64];
    __shared__ uint32_t sh[3];
    take_segment(b.ticket, b.n_seg, sh);
    const uint32_t seg = sh[0];
    uint32_t base, total;
    bin_fused_segment<U, T>(pos, table_view(b), seg << b.seg_log, hist, lds, &sh[2], b.new_start, b.new_bpos, b.new_bid,
                      b.status, seg, 0u, 0u, sh[1], &b.ctr->faults, &base, &total);
    if (seg == b.n_seg - 1 && threadIdx.x == 0) {
        b.new_start[b.m] = base + total;
        if (base + total != b.n) {
            atomicOr(&b.ctr->faults, kFaultBinCount);
        }
    }
}

// Counts the window agents that land in segment blockIdx.x, writes the
// segment-local exclusive offsets to new_start and the total to seg_total.
template <uint32_t U, uint32_t T>
static __global__ __launch_bounds__(T) void k_bin_count(const float2* __restrict__ pos, BinDev b) {
    extern __shared__ uint32_t hist[];
    __shared__ uint32_t lds[T / 64];
    const uint32_t seg = blockIdx.x;
    const uint32_t total = bin_count_segment<U, T>(pos, table_view(b), seg << b.seg_log, hist, lds, b.new_start);
    if (threadIdx.x == 0) {
        b.seg_total[seg] = total;
    }
}

// Adds the segment's global offset and places its agents.
template <uint32_t U, uint32_t T>
static __global__ __launch_bounds__(T, 8) void k_bin_place(const float2* __restrict__ pos, BinDev b) {
    extern __shared__ uint32_t cursor[];
    __shared__ uint32_t lds[T / 64];
    const uint32_t seg = blockIdx.x;
    uint32_t before = 0;
    for (uint32_t k = threadIdx.x; k < seg; k += T) {
        before += b.seg_total[k];
    }
    const uint32_t base = block_sum<T>(before, lds);
    if (seg == b.n_seg - 1 && threadIdx.x == 0) {
        const uint32_t total = base + b.seg_total[seg];
        b.new_start[b.m] = total;
        if (total != b.n) {
            atomicOr(&b.ctr->faults, kFaultBinCount);
        }
    }
    bin_place_segment<U, T>(pos, table_view(b), seg << b.seg_log, base, cursor, b.new_start, b.new_bpos, b.new_bid);
}

}  // namespace msim
