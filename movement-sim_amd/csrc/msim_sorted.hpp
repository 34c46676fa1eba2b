// msim_sorted.hpp — the move dispatch on bucket-sorted agent state.
//
// The hot agent state (position; target, speed and caller index in one
// 16-byte record) is kept in the slot order of the current binning, so the
// binning needs no slot -> agent map and positions need no gather.  One launch per move dispatch advances
// every agent (random_move.comp:830-852) and writes it, together with the
// rest of its hot state, straight into the next binning's slot order:
//
//   * the toroidal bucket table is cut into segments; workgroup `seg` owns
//     the slots of the agents whose new bucket lies in its segment (segments
//     dealt to the XCDs in contiguous runs, so that neighbouring windows
//     share an L2);
//   * an agent moves at most D cells per axis, so it lands within `margin`
//     buckets of its previous bucket: the segment's agents all sit in a
//     window of the previous slot order;
//   * count pass: every window slot carries `nxt`, the bucket of the agent's
//     next position, computed when the agent was last placed; counting reads
//     these 4 bytes per slot and no state;
//   * the segment's first slot is known before the launch: the previous
//     move and route choice counted, per segment, the agents whose next
//     bucket lies in it, and k_seg_scan turned those counts into first slots,
//     so no segment waits for another;
//   * rank: the window slots whose `nxt` lies in the segment take their new
//     slots from LDS counters; an LDS map records, per new slot, the old
//     slot's window index and the counted bucket;
//   * store pass: the new slots in order (whole lines written per wave),
//     each agent's position and record gathered from its old slot through
//     the map, moved, checked against its counted bucket and stored with the
//     bucket of its following move as the new `nxt`;
//   * arrivals are stored at once but for what the route choice decides
//     (target, next bucket, heading): each is listed (new slot, caller index,
//     position, speed) in the segment's own slice [base, base + total) of
//     the arrival list, and k_route_sorted, launched next, routes every
//     segment's list (random_move.comp:778-828).  Road and PRNG state live in
//     caller order, 32 bytes per agent with the connection index and count
//     of the endpoint the agent heads to, and are touched only there.
//
// Each agent's arithmetic is the plain tick's exactly; only where the state
// is stored changes.  The agents' order inside a bucket follows the rank
// pass's LDS atomics and carries no meaning: no result depends on it.
#pragma once

#include "msim_bin.hpp"

namespace msim {

// pos + heading of an agent that keeps walking (random_move.comp:830-846).
// The sorted move evaluates it twice for every agent and tick: for the move
// itself and, one tick earlier, to bin the agent's next position.  Both must
// round alike.  With contraction on, the compiler may fuse p + d differently
// at different call sites, so that build keeps a single out-of-line copy.
#ifdef MSIM_CONTRACTED
#define MSIM_WALK_ATTR __device__ __noinline__
#else
#define MSIM_WALK_ATTR __device__ __forceinline__
#endif
MSIM_WALK_ATTR float4 walk_step(float px, float py, float dx, float dy, float len, float s) {
    const float hx = (dx / len) * s;
    const float hy = (dy / len) * s;
    return make_float4(px + hx, py + hy, hx, hy);
}

// Position after the next move from p toward t at speed s: the step, or the
// target itself on arrival (random_move.comp:841-852).
__device__ __forceinline__ float2 next_position(float2 p, float2 t, float s) {
    const float dx = t.x - p.x;
    const float dy = t.y - p.y;
    const float len = vlen(dx, dy);
    if (len > s) {
        const float4 w = walk_step(p.x, p.y, dx, dy, len, s);
        return make_float2(w.x, w.y);
    }
    return t;
}

// Hot agent state in slot order (one copy per binning buffer).  Target,
// speed and caller index share one 16-byte record, so that moving an agent
// takes two gathers (position, record) and three stores (position, record,
// next bucket): a scattered gather costs about the same per instruction
// whatever its width, so fewer, wider ones are cheaper.  Positions and next
// buckets stay separate arrays, which the neighbour pass and the count
// passes stream alone.
struct SortedBufs {
    float2* pos;    // [n + pad]
    uint4* rec;     // [n] {target x, target y, speed (float bits; unused with spd_const), caller index}
    uint32_t* nxt;  // [n] bucket (in the agent's table) of its next position
    float2* dir;    // [n + pad] heading, written on a batch's last tick only
};

__device__ __forceinline__ uint4 hot_record(float2 tgt, float spd, uint32_t caller) {
    return make_uint4(__float_as_uint(tgt.x), __float_as_uint(tgt.y), __float_as_uint(spd), caller);
}

// Caller-order route state of one agent (cold[2 h], cold[2 h + 1]): road,
// connection index and count of the endpoint it heads to (route_endpoint),
// and the PRNG state.
__device__ __forceinline__ uint4 cold_word(uint32_t road, uint2 c) { return make_uint4(road, c.x, c.y, 0u); }

struct SortedMove {
    SortedBufs src, dst;
    uint4* arr_rec;      // [n] arrivals {new slot, caller index, x bits, y bits}, segment g's at
                         // [seg_start[g], + seg_arr[g])
    float* arr_spd;      // [n] their speeds; nullptr with spd_const
    const uint32_t* seg_start;  // [n_seg + 1] first slot of each segment (k_seg_scan)
    uint32_t* seg_next;         // [n_seg] agents whose next bucket lies in each segment, for the next move
    uint32_t* seg_arr;          // [n_seg] arrivals of each segment
    uint4* cold;                // [2 n] caller order: cold_word, PRNG state
    uint32_t n;
    const msim_road* roads;
    const uint32_t* conns;
    float spd_const;
    unsigned long long* part_stats;  // kWordInserts / kWordArrivals slots
    unsigned int* faults;
};

enum : unsigned int { kFaultNext = 4u };  // an agent's position left the bucket its count used

// Per-workgroup counts of next segments.  A workgroup works on segment
// g_local of a table of n_local segments (global index seg_first + g_local);
// its agents' next buckets lie within the window margin of that segment, so
// nh[d + kNextReach] counts those d segments away (cyclically) and farther
// ones go straight to seg_next.  next_flush adds nh to seg_next.
constexpr int kNextReach = 2;
struct NextCount {
    uint32_t* nh;  // LDS [2 kNextReach + 1]
    uint32_t* seg_next;
    uint32_t g_local, n_local, seg_first, seg_log;
};

// Wave-level (every lane of the wave calls it; inactive ones with
// active = false): counts the next buckets of the active lanes.
__device__ __forceinline__ void next_add(const NextCount& nc, uint32_t bucket, bool active) {
    const uint32_t ls = bucket >> nc.seg_log;
    const bool own = active && ls == nc.g_local;
    const unsigned long long om = __ballot(own);
    if (own && static_cast<int>(__lane_id()) == __ffsll(static_cast<long long>(om)) - 1) {
        atomicAdd(&nc.nh[kNextReach], static_cast<uint32_t>(__popcll(om)));
    }
    int d = static_cast<int>(ls) - static_cast<int>(nc.g_local);
    const int nl = static_cast<int>(nc.n_local);
    if (2 * d > nl) d -= nl;
    if (2 * d < -nl) d += nl;
    const bool near = d >= -kNextReach && d <= kNextReach;
    // Two separate atomics of different scopes: merged into one on a selected
    // pointer they would become a flat atomic, whose completion the
    // compiler can only await with every outstanding load.
    if (active && !own && near) {
        __hip_atomic_fetch_add(&nc.nh[d + kNextReach], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    if (active && !own && !near) {
        atomicAdd(&nc.seg_next[nc.seg_first + ls], 1u);
    }
}

// Workgroup-level, after a barrier: nh into seg_next.
__device__ __forceinline__ void next_flush(const NextCount& nc) {
    const int k = static_cast<int>(threadIdx.x);
    if (k < 2 * kNextReach + 1 && nc.nh[k]) {
        const int nl = static_cast<int>(nc.n_local);
        const int l = ((static_cast<int>(nc.g_local) + k - kNextReach) % nl + nl) % nl;
        atomicAdd(&nc.seg_next[nc.seg_first + static_cast<uint32_t>(l)], nc.nh[k]);
    }
}

// Count pass over window slots [lo, hi): agents whose next bucket lies in
// [s0, s0 + S) bump their LDS counter.
template <uint32_t U, uint32_t T>
__device__ __forceinline__ void sorted_count(const uint32_t* __restrict__ nxt, uint32_t s0, uint32_t S, uint32_t lo,
                                             uint32_t hi, uint32_t* hist) {
    for (uint32_t base = lo; base < hi; base += U * T) {
        uint32_t b[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            b[u] = nxt[min(base + u * T + threadIdx.x, hi - 1u)];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t k = base + u * T + threadIdx.x;
            const uint32_t off = b[u] - s0;
            (void)wave_rank_add(hist, off, k < hi && off < S);
        }
    }
}

// Hot state of the agent in one old slot: next bucket, position, target,
// speed and caller index.
struct HotAgent {
    float2 p, tg;
    float sp;
    uint32_t h, b;
};

// The agent of old slot k, whose next bucket (counted for its placement) is
// `bucket`.
template <bool UNIFORM>
__device__ __forceinline__ HotAgent load_agent(const SortedMove& mv, uint32_t k, uint32_t bucket) {
    const SortedBufs& src = mv.src;
    HotAgent a;
    a.b = bucket;
    a.p = src.pos[k];
    const uint4 r = src.rec[k];
    a.tg = make_float2(__uint_as_float(r.x), __uint_as_float(r.y));
    a.sp = UNIFORM ? mv.spd_const : __uint_as_float(r.z);
    a.h = r.w;
    return a;
}

// Steps an agent (random_move.comp:841-852) and stores it at new slot q; for
// an arrival, what the route choice decides (target, next bucket, heading)
// is left to k_route_sorted, and rec / rec_spd receive its arrival record.
// Returns whether the agent arrived; n_ins counts a cell change.
template <bool UNIFORM, bool WRITE_DIR>
__device__ __forceinline__ bool step_agent(const SortedMove& mv, const TableView& t, const HotAgent& a, uint32_t q,
                                           uint32_t& n_ins, uint4& rec, float& rec_spd, uint32_t& next_bucket) {
    const SortedBufs& dst = mv.dst;
    const uint32_t b = a.b;
    const float2 p = a.p;
    const float2 tg = a.tg;
    const float sp = a.sp;
    const uint32_t h = a.h;
    const float dx = tg.x - p.x;
    const float dy = tg.y - p.y;
    const float len = vlen(dx, dy);
    float2 np = tg, d = make_float2(0.0f, 0.0f);
    const bool arr = !(len > sp);
    if (!arr) {
        const float4 w = walk_step(p.x, p.y, dx, dy, len, sp);
        np = make_float2(w.x, w.y);
        d = make_float2(w.z, w.w);
    }
    const int32_t cx = cell_coord(np.x, t.inv_cell), cy = cell_coord(np.y, t.inv_cell);
    if (bucket_of(cx, cy, t.lw, t.wmask, t.hmask) != b) {
        atomicOr(mv.faults, kFaultNext);
    }
    n_ins += (cx != cell_coord(p.x, t.inv_cell) || cy != cell_coord(p.y, t.inv_cell)) ? 1u : 0u;
    dst.pos[q] = np;
    dst.rec[q] = hot_record(tg, sp, h);  // an arrival's target is replaced by k_route_sorted
    if (!arr) {
        next_bucket = bucket_of_pos(next_position(np, tg, sp), t);
        dst.nxt[q] = next_bucket;
        if (WRITE_DIR) dst.dir[q] = d;
    }
    rec = make_uint4(q, h, __float_as_uint(np.x), __float_as_uint(np.y));
    rec_spd = sp;
    return arr;
}

// Moves the agent of old slot k, counted in `bucket`, to new slot q
// (load_agent, then step_agent).
template <bool UNIFORM, bool WRITE_DIR>
__device__ __forceinline__ bool move_agent(const SortedMove& mv, const TableView& t, uint32_t k, uint32_t bucket,
                                           uint32_t q, uint32_t& n_ins, uint4& rec, float& rec_spd,
                                           uint32_t& next_bucket) {
    return step_agent<UNIFORM, WRITE_DIR>(mv, t, load_agent<UNIFORM>(mv, k, bucket), q, n_ins, rec, rec_spd,
                                          next_bucket);
}

// Appends one arrival record at list index i.
template <bool UNIFORM>
__device__ __forceinline__ void put_arrival(const SortedMove& mv, uint32_t i, const uint4& rec, float spd) {
    mv.arr_rec[i] = rec;
    if (!UNIFORM) mv.arr_spd[i] = spd;
}

// The segment's window: old slots [a0, a1) then [c0, c1), numbered from 0.
struct Window {
    uint32_t a0, a1, c0, c1;
    __device__ uint32_t slot(uint32_t r) const { return r < a1 - a0 ? a0 + r : c0 + (r - (a1 - a0)); }
    __device__ uint32_t size() const { return (a1 - a0) + (c1 - c0); }
};

// Rank pass over window slots [lo, hi) (window-relative index rel0 for lo):
// each agent whose next bucket lies in the segment takes cursor[bucket]++,
// its new slot.  The segment's output position j = slot - base records the
// agent's window index in map[j] and its bucket in the segment in mapb[j]
// when j < cap; past cap (a segment more
// crowded than the map holds) the agent is moved and stored at once.  A
// position at or past limit (the slots the segment was given) is a count
// mismatch: the agent is not stored and kFaultBinCount is raised.  Reads
// only the next buckets; the next round's are loaded before this round is
// ranked.  Waves run the rounds independently (the rank helper needs whole
// waves, no workgroup barrier).
template <uint32_t U, uint32_t T, bool UNIFORM, bool WRITE_DIR>
__device__ __forceinline__ void sorted_rank(const SortedMove& mv, const TableView& t, const NextCount& nc,
                                            uint32_t s0, uint32_t lo, uint32_t hi, uint32_t rel0, uint32_t* cursor,
                                            uint16_t* map, uint16_t* mapb, uint32_t base, uint32_t cap,
                                            uint32_t limit,
                                            uint32_t* arr_count, uint32_t arr_base, uint32_t& n_ins) {
    if (lo >= hi) {
        return;
    }
    const uint32_t S = 1u << t.seg_log;
    uint32_t bn[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
        bn[u] = mv.src.nxt[min(lo + u * T + threadIdx.x, hi - 1u)];
    }
    for (uint32_t r0 = lo; r0 < hi; r0 += U * T) {
        uint32_t b[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            b[u] = bn[u];
        }
        if (r0 + U * T < hi) {
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                bn[u] = mv.src.nxt[min(r0 + U * T + u * T + threadIdx.x, hi - 1u)];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t k = r0 + u * T + threadIdx.x;
            const bool mine = k < hi && b[u] - s0 < S;
            const uint32_t j = wave_rank_add(cursor, b[u] - s0, mine) - base;
            bool moved = false, arr = false;
            uint32_t nb = 0;
            if (mine) {
                uint4 rec;
                float rs;
                if (j >= limit) {
                    atomicOr(mv.faults, kFaultBinCount);
                } else if (j < cap) {
                    map[j] = static_cast<uint16_t>(rel0 + (k - lo));
                    mapb[j] = static_cast<uint16_t>(b[u] - s0);
                } else {
                    moved = true;
                    arr = move_agent<UNIFORM, WRITE_DIR>(mv, t, k, b[u], base + j, n_ins, rec, rs, nb);
                    if (arr) put_arrival<UNIFORM>(mv, arr_base + atomicAdd(arr_count, 1u), rec, rs);
                }
            }
            next_add(nc, nb, moved && !arr);
        }
    }
}

// Store pass over the segment's first min(total, cap) new slots, in slot
// order: each agent's state is gathered from its old slot (the window index
// in map; its bucket, s0 + mapb, needs no gather), moved and stored, so every wave writes consecutive slots and each
// output line is written whole.  A round's U gathers are all issued before
// the first agent is stepped (its stores would otherwise hold back the
// next agent's loads).  Arrivals are listed for k_route_sorted (one LDS
// atomic per wave and round).
template <uint32_t U, uint32_t T, bool UNIFORM, bool WRITE_DIR>
__device__ __forceinline__ void sorted_store(const SortedMove& mv, const TableView& t, const NextCount& nc,
                                             const Window& win, const uint16_t* map, const uint16_t* mapb,
                                             uint32_t s0, uint32_t base, uint32_t count,
                                             uint32_t* arr_count, uint32_t arr_base, uint32_t& n_ins) {
    const unsigned long long lt = (1ull << __lane_id()) - 1ull;
    for (uint32_t j0 = 0; j0 < count; j0 += U * T) {
        bool arr[U];
        uint4 rec[U];
        float rs[U];
        HotAgent ag[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {  // j0 < count: slot count - 1 is mapped
            const uint32_t j = min(j0 + u * T + threadIdx.x, count - 1u);
            ag[u] = load_agent<UNIFORM>(mv, win.slot(map[j]), s0 + mapb[j]);
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t j = j0 + u * T + threadIdx.x;
            arr[u] = false;
            uint32_t nb = 0;
            if (j < count) {
                arr[u] = step_agent<UNIFORM, WRITE_DIR>(mv, t, ag[u], base + j, n_ins, rec[u], rs[u], nb);
            }
            next_add(nc, nb, j < count && !arr[u]);
        }
        unsigned long long am[U];
        uint32_t round_arr = 0;
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            am[u] = __ballot(arr[u]);
            round_arr += static_cast<uint32_t>(__popcll(am[u]));
        }
        if (round_arr) {  // wave-uniform
            uint32_t at = 0;
            if (__lane_id() == 0) at = atomicAdd(arr_count, round_arr);
            at = arr_base + static_cast<uint32_t>(__shfl(static_cast<int>(at), 0, 64));
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                if (arr[u]) {
                    put_arrival<UNIFORM>(mv, at + static_cast<uint32_t>(__popcll(am[u] & lt)), rec[u], rs[u]);
                }
                at += static_cast<uint32_t>(__popcll(am[u]));
            }
        }
    }
}

// LDS of the sorted move: the segment's counters (4 bytes per bucket) and,
// for its first sorted_map_cap new slots, each agent's window index and its
// bucket in the segment (2 + 2 bytes).  Segments of up to 2^12 buckets map 2
// slots per bucket (12 bytes per bucket in all), larger ones 1 (8 bytes).
__host__ __device__ constexpr uint32_t sorted_map_cap(uint32_t seg_log) {
    return seg_log <= 12u ? (2u << seg_log) : (1u << seg_log);
}
__host__ __device__ constexpr size_t sorted_lds_bytes(uint32_t seg_log) {
    return (size_t{4} << seg_log) + size_t{4} * sorted_map_cap(seg_log);
}

// Count and place segment g (global index; s0 its first bucket in table t).
// Writes the segment's absolute bucket offsets to start_out[s0 + j];
// *base_out / *total_out receive the segment's first slot and agent count
// (all threads).  The segment's insert and arrival counts go to the
// statistics slots; its agents' next segments are counted for the next move.
template <uint32_t U, uint32_t T, bool UNIFORM, bool WRITE_DIR>
__device__ __forceinline__ void sorted_segment(const SortedMove& mv, const TableView& t, const NextCount& nc,
                                               uint32_t s0, uint32_t g, uint32_t* hist, uint32_t* lds,
                                               uint32_t* sh, uint32_t* __restrict__ start_out, uint32_t* base_out,
                                               uint32_t* total_out) {
    const uint32_t tid = threadIdx.x;
    const uint32_t S = 1u << t.seg_log;
    const uint32_t st0 = mv.seg_start[g], st1 = mv.seg_start[g + 1];  // loaded while the window is counted
    for (uint32_t j = tid; j < S; j += T) {
        hist[j] = 0u;
    }
    if (tid < 2 * kNextReach + 1) {
        nc.nh[tid] = 0u;
    }
    if (tid == 0) {
        sh[0] = 0u;  // the segment's arrivals so far
    }
    __syncthreads();
    uint32_t a0, a1, c0, c1;
    bin_window(t, s0, a0, a1, c0, c1);
    const Window win{a0, a1, c0, c1};
    const uint32_t wsize = win.size();
    // A window of at most R T slots is counted in one round whose LDS
    // atomics also give each agent its rank in its bucket (kept in
    // registers, so no rank pass is needed); larger windows are counted,
    // then ranked, in rounds.
    // (U = 4: 12 slots per thread, 6,144 per round, so that more windows
    // take one round; the registers stay within the same occupancy)
    constexpr uint32_t R = U == 4 ? 12 : 2 * U;
    const bool one_round = wsize <= R * T;  // workgroup-uniform
    // (bucket offset << 16) | rank in the bucket, ~0 when not the segment's
    // agent (segments have at most 2^14 buckets, windows here R T <= 2^13 slots)
    uint32_t pk[R];
    if (one_round) {
        uint32_t b[R];
#pragma unroll
        for (uint32_t r = 0; r < R; r++) {
            const uint32_t idx = r * T + tid;
            b[r] = idx < wsize ? mv.src.nxt[win.slot(idx)] : 0u;
        }
#pragma unroll
        for (uint32_t r = 0; r < R; r++) {
            const uint32_t off = b[r] - s0;
            const bool mine = r * T + tid < wsize && off < S;
            const uint32_t rank = wave_rank_add(hist, off, mine);
            pk[r] = mine ? (off << 16) | rank : ~0u;
        }
    } else {
        sorted_count<2 * U, T>(mv.src.nxt, s0, S, a0, a1, hist);
        sorted_count<2 * U, T>(mv.src.nxt, s0, S, c0, c1, hist);
    }
    __syncthreads();
    const uint32_t per = S / T;
    uint32_t run = 0;
    for (uint32_t q = 0; q < per; q++) {
        run += hist[tid * per + q];
    }
    uint32_t total = 0;
    uint32_t excl = block_excl_scan<T>(run, lds, total);
    // the slots this segment was given, clamped to [0, n): counts that went
    // wrong raise a fault but never place an agent outside the state
    const uint32_t base = min(st0, mv.n);
    const uint32_t end = max(base, min(st1, mv.n));
    const uint32_t limit = end - base;
    if (tid == 0 && total != limit) {
        atomicOr(mv.faults, kFaultBinCount);
    }
    for (uint32_t q = 0; q < per; q++) {  // counts -> absolute bucket offsets
        const uint32_t v = hist[tid * per + q];
        hist[tid * per + q] = base + excl;
        excl += v;
    }
    __syncthreads();
    // the offsets into the new bucket table, clamped to the segment's slots,
    // stored in bucket order (whole lines per wave)
    for (uint32_t j = tid; j < S; j += T) {
        start_out[s0 + j] = min(hist[j], base + limit);
    }
    uint32_t n_ins = 0;
    uint16_t* map = reinterpret_cast<uint16_t*>(hist + S);
    uint16_t* mapb = map + sorted_map_cap(t.seg_log);
    // window indices must fit the map's 16 bits; a window past that moves
    // every agent straight from the rank pass
    const uint32_t cap = wsize <= 65536u ? sorted_map_cap(t.seg_log) : 0u;
    if (one_round) {
#pragma unroll
        for (uint32_t r = 0; r < R; r++) {
            const uint32_t idx = r * T + tid;
            bool moved = false, arr = false;
            uint32_t nb = 0;
            if (pk[r] != ~0u) {
                const uint32_t j = hist[pk[r] >> 16] - base + (pk[r] & 0xffffu);
                if (j >= limit) {
                    atomicOr(mv.faults, kFaultBinCount);
                } else if (j < cap) {
                    map[j] = static_cast<uint16_t>(idx);
                    mapb[j] = static_cast<uint16_t>(pk[r] >> 16);
                } else {
                    uint4 rec;
                    float rs;
                    moved = true;
                    arr = move_agent<UNIFORM, WRITE_DIR>(mv, t, win.slot(idx), s0 + (pk[r] >> 16), base + j, n_ins,
                                                         rec, rs, nb);
                    if (arr) put_arrival<UNIFORM>(mv, base + atomicAdd(&sh[0], 1u), rec, rs);
                }
            }
            next_add(nc, nb, moved && !arr);
        }
    } else {
        __syncthreads();  // the rank pass advances the offsets the stores above read
        sorted_rank<U, T, UNIFORM, WRITE_DIR>(mv, t, nc, s0, a0, a1, 0u, hist, map, mapb, base, cap, limit,
                                              &sh[0], base, n_ins);
        sorted_rank<U, T, UNIFORM, WRITE_DIR>(mv, t, nc, s0, c0, c1, a1 - a0, hist, map, mapb, base, cap, limit,
                                              &sh[0], base, n_ins);
    }
    __syncthreads();
    sorted_store<U, T, UNIFORM, WRITE_DIR>(mv, t, nc, win, map, mapb, s0, base, min(min(total, cap), limit),
                                           &sh[0], base, n_ins);
    __syncthreads();
    next_flush(nc);
    // block totals of insertions and arrivals, one slot update each
    const uint32_t si = block_sum<T>(n_ins, lds);
    if (tid == 0) {
        const uint32_t sa = sh[0];
        slot_add(mv.part_stats, g, kWordInserts, si);
        slot_add(mv.part_stats, g, kWordArrivals, sa);
        mv.seg_arr[g] = sa;
    }
    *base_out = base;
    *total_out = limit;
}


// Sorted move of one table (grid: n_seg workgroups; dynamic LDS:
// sorted_lds_bytes(seg_log)).
template <uint32_t U, uint32_t T, bool UNIFORM, bool WRITE_DIR>
static __global__ __launch_bounds__(T) void k_move_sorted(SortedMove mv, BinDev b) {
    extern __shared__ uint32_t hist[];
    __shared__ uint32_t lds[T / 64];
    __shared__ uint32_t nh[2 * kNextReach + 1];
    __shared__ uint32_t sh[1];
    const uint32_t seg = xcd_block(blockIdx.x, gridDim.x);
    const NextCount nc{nh, mv.seg_next, seg, b.n_seg, 0u, b.seg_log};
    uint32_t base, total;
    sorted_segment<U, T, UNIFORM, WRITE_DIR>(mv, table_view(b), nc, seg << b.seg_log, seg, hist, lds, sh,
                                             b.new_start, &base, &total);
    if (seg == b.n_seg - 1 && threadIdx.x == 0) {
        b.new_start[b.m] = base + total;
    }
}

// First slot of every segment from the per-segment agent counts (the
// segments of all tables in order): seg_start[g] = cnt[0] + ... + cnt[g - 1],
// seg_start[n_seg] = the total, which must be n.  cnt is cleared for its next
// use.  One workgroup: wave w takes the contiguous chunk [w C, (w + 1) C) in
// pieces of 64 K entries, each loaded at once (coalesced, K loads in flight
// per lane) and kept in registers for both the chunk sum and the offsets.
template <uint32_t T>
static __global__ __launch_bounds__(T) void k_seg_scan(uint32_t* __restrict__ cnt, uint32_t* __restrict__ seg_start,
                                                       uint32_t n_seg, uint32_t n, unsigned int* faults) {
    constexpr uint32_t W = T / 64, K = 16;
    __shared__ uint32_t wsum[W];
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    const uint32_t C = (n_seg + W - 1) / W;
    const uint32_t c0 = min(w * C, n_seg), c1 = min(c0 + C, n_seg);
    const bool one_piece = C <= 64u * K;  // workgroup-uniform
    uint32_t v[K];
    uint32_t part = 0;
    for (uint32_t p0 = c0; p0 < c1; p0 += 64u * K) {
#pragma unroll
        for (uint32_t k = 0; k < K; k++) {
            const uint32_t g = p0 + 64u * k + lane;
            v[k] = g < c1 ? cnt[g] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < K; k++) {
            part += v[k];
        }
    }
    part = wave_sum(part);
    if (lane == 0) wsum[w] = part;
    __syncthreads();
    uint32_t run = 0, total = 0;
    for (uint32_t k = 0; k < W; k++) {
        run += (k < w) ? wsum[k] : 0u;
        total += wsum[k];
    }
    for (uint32_t p0 = c0; p0 < c1; p0 += 64u * K) {
        if (!one_piece) {
#pragma unroll
            for (uint32_t k = 0; k < K; k++) {
                const uint32_t g = p0 + 64u * k + lane;
                v[k] = g < c1 ? cnt[g] : 0u;
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < K; k++) {
            const uint32_t g = p0 + 64u * k + lane;
            const uint32_t inc = wave_incl_scan(v[k]);
            if (g < c1) {
                seg_start[g] = run + inc - v[k];
                cnt[g] = 0u;
            }
            run += static_cast<uint32_t>(__shfl(static_cast<int>(inc), 63, 64));
        }
    }
    if (threadIdx.x == 0) {
        seg_start[n_seg] = total;
        if (total != n) {
            atomicOr(faults, kFaultBinCount);
        }
    }
}

// Route choice of the arrivals listed by the sorted move: one workgroup per
// segment (grid n_seg), its list at arr_rec[seg_start[g] ...].  An arrival
// stands on its old target (the position in its record); the route choice
// (random_move.comp:778-828) updates the caller-order route state and gives
// the new target, the bucket of the next move and (last tick of a batch) the
// heading.
template <bool UNIFORM, bool WRITE_DIR>
__device__ __forceinline__ void route_segment_list(const SortedMove& mv, const TableView& t, const NextCount& nc,
                                                   uint32_t g) {
    const SortedBufs& dst = mv.dst;
    if (threadIdx.x < 2 * kNextReach + 1) {
        nc.nh[threadIdx.x] = 0u;
    }
    __syncthreads();
    const uint32_t base = min(mv.seg_start[g], mv.n), cnt = mv.seg_arr[g];
    for (uint32_t j0 = 0; j0 < cnt; j0 += blockDim.x) {  // whole-block rounds: next_add is wave-level
        const uint32_t j = j0 + threadIdx.x;
        uint32_t nb = 0;
        if (j < cnt) {
            const uint4 a = mv.arr_rec[base + j];
            const uint32_t q = a.x, hh = a.y;
            const float2 at = make_float2(__uint_as_float(a.z), __uint_as_float(a.w));
            const float s = UNIFORM ? mv.spd_const : mv.arr_spd[base + j];
            const uint4 w = mv.cold[2u * hh];
            uint4 st = mv.cold[2u * hh + 1u];
            float2 nt = at;
            uint32_t r = w.x;
            uint2 c = make_uint2(w.y, w.z);
            choose_route_at(nt, r, c, st, mv.roads, mv.conns);
            mv.cold[2u * hh] = cold_word(r, c);
            mv.cold[2u * hh + 1u] = st;
            nb = bucket_of_pos(next_position(at, nt, s), t);
            *reinterpret_cast<float2*>(&dst.rec[q]) = nt;  // the record's target
            dst.nxt[q] = nb;
            if (WRITE_DIR) dst.dir[q] = heading(nt, at, s);
        }
        next_add(nc, nb, j < cnt);
    }
    __syncthreads();
    next_flush(nc);
}

template <bool UNIFORM, bool WRITE_DIR>
static __global__ __launch_bounds__(kBlock) void k_route_sorted(SortedMove mv, BinDev b) {
    __shared__ uint32_t nh[2 * kNextReach + 1];
    const uint32_t g = xcd_block(blockIdx.x, gridDim.x);
    route_segment_list<UNIFORM, WRITE_DIR>(mv, table_view(b), NextCount{nh, mv.seg_next, g, b.n_seg, 0u, b.seg_log},
                                           g);
}

// Wave-level: one count per distinct segment among the active lanes' gseg
// (the lanes of a wave hold consecutive slots in bucket order, so there are
// few).
__device__ __forceinline__ void count_next_segment(uint32_t* seg_next, uint32_t gseg, bool active) {
    unsigned long long todo = __ballot(active);
    while (todo) {
        const int leader = __ffsll(static_cast<long long>(todo)) - 1;
        const uint32_t k = static_cast<uint32_t>(__shfl(static_cast<int>(gseg), leader, 64));
        const unsigned long long peers = __ballot(active && gseg == k) & todo;
        if (static_cast<int>(__lane_id()) == leader) {
            atomicAdd(&seg_next[k], static_cast<uint32_t>(__popcll(peers)));
        }
        todo &= ~peers;
    }
}

// nxt for every slot of one table: the bucket of the agent's next position;
// seg_next (cleared before) counts them per segment.
template <bool UNIFORM>
static __global__ __launch_bounds__(kBlock) void k_sorted_next(SortedBufs st, uint32_t n, float spd_const,
                                                               float inv_cell, uint32_t lw, uint32_t wmask,
                                                               uint32_t hmask, uint32_t seg_log,
                                                               uint32_t* __restrict__ seg_next) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t b = 0;
    if (i < n) {
        const uint4 r = st.rec[i];
        const float2 np = next_position(st.pos[i], make_float2(__uint_as_float(r.x), __uint_as_float(r.y)),
                                        UNIFORM ? spd_const : __uint_as_float(r.z));
        b = bucket_of(cell_coord(np.x, inv_cell), cell_coord(np.y, inv_cell), lw, wmask, hmask);
        st.nxt[i] = b;
    }
    count_next_segment(seg_next, b >> seg_log, i < n);
}

// p[0, n) = v (in graphs too, where memset nodes are avoided).
static __global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        p[i] = v;
    }
}

// Route state into caller order: cold[perm[i]] from slot i's road, PRNG
// state and target.
static __global__ __launch_bounds__(kBlock) void k_cold_enter(const uint32_t* __restrict__ perm,
                                                        const uint32_t* __restrict__ road,
                                                        const uint4* __restrict__ rng,
                                                        const float2* __restrict__ tgt,
                                                        const msim_road* __restrict__ roads, uint4* __restrict__ cold,
                                                        uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) {
        return;
    }
    const uint32_t h = perm[i], r = road[i];
    Endpoint s, e;
    load_road(roads, r, s, e);
    cold[2u * h] = cold_word(r, route_endpoint(s, e, tgt[i]));
    cold[2u * h + 1u] = rng[i];
}

// Route state back into slot order: road[i], rng[i] of caller perm[i].
static __global__ __launch_bounds__(kBlock) void k_cold_publish(const uint32_t* __restrict__ perm,
                                                          const uint4* __restrict__ cold,
                                                          uint32_t* __restrict__ road, uint4* __restrict__ rng,
                                                          uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) {
        return;
    }
    const uint32_t h = perm[i];
    road[i] = cold[2u * h].x;
    rng[i] = cold[2u * h + 1u];
}

// dst[i] = src[perm[i]]
template <class V>
static __global__ __launch_bounds__(kBlock) void k_gather_perm(const uint32_t* __restrict__ perm,
                                                               const V* __restrict__ src, V* __restrict__ dst,
                                                               uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        dst[i] = src[perm[i]];
    }
}

// dst[caller index of slot i] = src[i], the caller index from the slot's record.
template <class V>
static __global__ __launch_bounds__(kBlock) void k_scatter_rec(const uint4* __restrict__ rec,
                                                               const V* __restrict__ src, V* __restrict__ dst,
                                                               uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        dst[rec[i].w] = src[i];
    }
}

// Cells of the sorted state's agents in caller order (the caller index from
// each slot's record): k_export_cells for the sorted state.
static __global__ __launch_bounds__(kBlock) void k_export_cells_rec(const float2* __restrict__ pos,
                                                                    const uint4* __restrict__ rec, uint32_t n,
                                                                    float inv_cell, int2* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const float2 p = pos[i];
        out[rec[i].w] = make_int2(cell_coord(p.x, inv_cell), cell_coord(p.y, inv_cell));
    }
}

// Records from the classic per-slot arrays (entering the sorted state); spd
// nullptr: every agent moves at the constant speed.
static __global__ __launch_bounds__(kBlock) void k_pack_rec(const float2* __restrict__ tgt,
                                                            const float* __restrict__ spd,
                                                            const uint32_t* __restrict__ perm,
                                                            uint4* __restrict__ rec, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        rec[i] = hot_record(tgt[i], spd ? spd[i] : 0.0f, perm[i]);
    }
}

// The classic per-slot arrays from the records (readbacks, leaving the
// sorted state).
static __global__ __launch_bounds__(kBlock) void k_unpack_rec(const uint4* __restrict__ rec,
                                                              float2* __restrict__ tgt, float* __restrict__ spd,
                                                              uint32_t* __restrict__ perm, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const uint4 r = rec[i];
        tgt[i] = make_float2(__uint_as_float(r.x), __uint_as_float(r.y));
        if (spd) spd[i] = __uint_as_float(r.z);
        perm[i] = r.w;
    }
}

}  // namespace msim
