// msim_sorted.hpp — the move dispatch on bucket-sorted agent state.
//
// The hot agent state (position, target, speed, caller index) is kept in the
// slot order of the current binning, so the binning needs no slot -> agent
// map and positions need no gather.  One launch per move dispatch advances
// every agent (random_move.comp:830-852) and writes it, together with the
// rest of its hot state, straight into the next binning's slot order:
//
//   * the toroidal bucket table is cut into segments; workgroup `seg` (taken
//     by ticket, as in the one-pass rebinning) owns the slots of the agents
//     whose new bucket lies in its segment;
//   * an agent moves at most D cells per axis, so it lands within `margin`
//     buckets of its previous bucket: the segment's agents all sit in a
//     window of the previous slot order;
//   * count pass: every window slot carries `nxt`, the bucket of the agent's
//     next position, computed when the agent was last placed; counting reads
//     these 4 bytes per slot and no state;
//   * look-back over the earlier segments' totals gives the segment's first
//     slot;
//   * place pass: the window slots whose `nxt` lies in the segment load
//     their state (coalesced: the window is a contiguous slot range), move,
//     take arrivals through the route choice (random_move.comp:778-828;
//     road and PRNG state live in caller order and are touched only here),
//     and are stored at their new slots with the bucket of their following
//     move as the new `nxt`.
//
// Each agent's arithmetic is the plain tick's exactly; only where the state
// is stored changes.  The agents' order inside a bucket follows the place
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

// Hot agent state in slot order (one copy per binning buffer).
struct SortedBufs {
    float2* pos;     // [n + pad]
    float2* tgt;     // [n + pad]
    float* spd;      // [n + pad]; nullptr when every agent moves at spd_const
    uint32_t* perm;  // [n] caller index of the slot's agent
    uint32_t* nxt;   // [n] bucket (in the agent's table) of its next position
    float2* dir;     // [n + pad] heading, written on a batch's last tick only
};

struct SortedMove {
    SortedBufs src, dst;
    uint32_t* road;  // [n] caller order
    uint4* rng;      // [n] caller order
    const msim_road* roads;
    const uint32_t* conns;
    float spd_const;
    unsigned long long* part_stats;  // kWordInserts / kWordArrivals slots
    unsigned int* faults;
};

enum : unsigned int { kFaultNext = 4u };  // an agent's position left the bucket its count used

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

// Routes the arrivals listed in this wave's LDS list (entries: new slot,
// old slot): route choice (random_move.comp:778-828) on the caller-order
// road and PRNG state, then the new target, the bucket of the next move and
// (last tick) the heading.  The wave's lanes take the entries together.
template <bool UNIFORM, bool WRITE_DIR>
__device__ __forceinline__ void sorted_route_list(const SortedMove& mv, const TableView& t, const uint2* alist,
                                                  uint32_t na) {
    const SortedBufs& src = mv.src;
    const SortedBufs& dst = mv.dst;
    wave_lds_sync();
    for (uint32_t j = __lane_id(); j < na; j += 64) {
        const uint2 e = alist[j];
        const float2 at = src.tgt[e.y];  // the old target, where the agent now stands
        const uint32_t hh = src.perm[e.y];
        const float s = UNIFORM ? mv.spd_const : src.spd[e.y];
        float2 nt = at;
        uint32_t r = mv.road[hh];
        uint4 st = mv.rng[hh];
        choose_route(nt, r, st, mv.roads, mv.conns);
        mv.road[hh] = r;
        mv.rng[hh] = st;
        dst.tgt[e.x] = nt;
        dst.nxt[e.x] = bucket_of_pos(next_position(at, nt, s), t);
        if (WRITE_DIR) dst.dir[e.x] = heading(nt, at, s);
    }
    wave_lds_sync();
}

// Place pass over window slots [lo, hi): the segment's agents move and are
// stored at cursor[new bucket - s0]++.  Only slots whose next bucket lies in
// the segment load their state (the margins of the window mostly hold agents
// of the neighbouring segments, and a lane that loads nothing costs no
// traffic); the next round's buckets are loaded before this round's state, so
// each round waits for one memory round trip.  Arrivals are stored at once
// except for what the route choice decides (target, next bucket, heading);
// those are listed in the wave's LDS list and routed when the list could
// overflow and at the end of the segment.  Waves run the rounds
// independently (the rank helper needs whole waves, no workgroup barrier).
template <uint32_t U, uint32_t T, bool UNIFORM, bool WRITE_DIR>
__device__ __forceinline__ void sorted_place(const SortedMove& mv, const TableView& t, uint32_t s0, uint32_t lo,
                                             uint32_t hi, uint32_t* cursor, uint2* alist, uint32_t& na,
                                             uint32_t& n_ins, uint32_t& n_arr) {
    if (lo >= hi) {
        return;
    }
    const uint32_t S = 1u << t.seg_log;
    const unsigned long long lt = (1ull << __lane_id()) - 1ull;
    const SortedBufs& src = mv.src;
    const SortedBufs& dst = mv.dst;
    uint32_t bn[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
        bn[u] = src.nxt[min(lo + u * T + threadIdx.x, hi - 1u)];
    }
    for (uint32_t base = lo; base < hi; base += U * T) {
        uint32_t b[U];
        bool mine[U];
        bool any = false;
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            b[u] = bn[u];
            mine[u] = base + u * T + threadIdx.x < hi && b[u] - s0 < S;
            any = any || mine[u];
        }
        if (base + U * T < hi) {
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                bn[u] = src.nxt[min(base + U * T + u * T + threadIdx.x, hi - 1u)];
            }
        }
        if (__ballot(any) == 0ull) {  // wave-uniform: none of the wave's slots is the segment's
            continue;
        }
        uint32_t h[U];
        float2 p[U], tg[U];
        float sp[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t k = base + u * T + threadIdx.x;
            p[u] = make_float2(0.0f, 0.0f);
            tg[u] = make_float2(1.0f, 1.0f);
            sp[u] = mv.spd_const;
            h[u] = 0u;
            if (mine[u]) {
                p[u] = src.pos[k];
                tg[u] = src.tgt[k];
                if (!UNIFORM) sp[u] = src.spd[k];
                h[u] = src.perm[k];
            }
        }
        bool arr[U];
        float2 np[U], d[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const float dx = tg[u].x - p[u].x;
            const float dy = tg[u].y - p[u].y;
            const float len = vlen(dx, dy);
            if (len > sp[u]) {
                const float4 w = walk_step(p[u].x, p[u].y, dx, dy, len, sp[u]);
                np[u] = make_float2(w.x, w.y);
                d[u] = make_float2(w.z, w.w);
                arr[u] = false;
            } else {
                np[u] = tg[u];
                d[u] = make_float2(0.0f, 0.0f);
                arr[u] = mine[u];
            }
        }
        uint32_t slot[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            slot[u] = wave_rank_add(cursor, b[u] - s0, mine[u]);
        }
        uint32_t round_arr = 0;
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            round_arr += static_cast<uint32_t>(__popcll(__ballot(arr[u])));
        }
        if (na + round_arr > 64u * U) {  // wave-uniform: route the list before it overflows
            sorted_route_list<UNIFORM, WRITE_DIR>(mv, t, alist, na);
            n_arr += __lane_id() == 0 ? na : 0u;
            na = 0;
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            if (mine[u]) {
                const int32_t cx = cell_coord(np[u].x, t.inv_cell), cy = cell_coord(np[u].y, t.inv_cell);
                if (bucket_of(cx, cy, t.lw, t.wmask, t.hmask) != b[u]) {
                    atomicOr(mv.faults, kFaultNext);
                }
                n_ins += (cx != cell_coord(p[u].x, t.inv_cell) || cy != cell_coord(p[u].y, t.inv_cell)) ? 1u : 0u;
                const uint32_t q = slot[u];
                dst.pos[q] = np[u];
                if (!UNIFORM) dst.spd[q] = sp[u];
                dst.perm[q] = h[u];
                if (!arr[u]) {
                    dst.tgt[q] = tg[u];
                    dst.nxt[q] = bucket_of_pos(next_position(np[u], tg[u], sp[u]), t);
                    if (WRITE_DIR) dst.dir[q] = d[u];
                }
            }
            const unsigned long long m = __ballot(arr[u]);
            if (arr[u]) {
                alist[na + static_cast<uint32_t>(__popcll(m & lt))] = make_uint2(slot[u], base + u * T + threadIdx.x);
            }
            na += static_cast<uint32_t>(__popcll(m));
        }
    }
}

// LDS the sorted move needs besides the segment's counters.
template <uint32_t U, uint32_t T>
constexpr size_t sorted_list_bytes() {
    return sizeof(uint2) * (T / 64) * 64 * U;
}

// Count, look back, place.  Writes the segment's absolute bucket offsets to
// start_out[s0 + j]; *base_out / *total_out receive the segment's first slot
// and agent count (all threads).  The segment's insert and arrival counts go
// to the statistics slots.
template <uint32_t U, uint32_t T, bool UNIFORM, bool WRITE_DIR>
__device__ __forceinline__ void sorted_segment(const SortedMove& mv, const TableView& t, uint32_t s0, uint32_t* hist,
                                               uint2* alist_all, uint32_t* lds, uint32_t* sh_base,
                                               uint32_t* __restrict__ start_out, unsigned long long* status,
                                               uint32_t g, uint32_t g_first, uint32_t first, uint32_t epoch,
                                               uint32_t* base_out, uint32_t* total_out) {
    const uint32_t tid = threadIdx.x;
    const uint32_t S = 1u << t.seg_log;
    for (uint32_t j = tid; j < S; j += T) {
        hist[j] = 0u;
    }
    __syncthreads();
    uint32_t a0, a1, c0, c1;
    bin_window(t, s0, a0, a1, c0, c1);
    sorted_count<2 * U, T>(mv.src.nxt, s0, S, a0, a1, hist);
    sorted_count<2 * U, T>(mv.src.nxt, s0, S, c0, c1, hist);
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
        *sh_base = seg_lookback(status, g, g_first, first, total, epoch, mv.faults);
    }
    __syncthreads();
    const uint32_t base = *sh_base;
    for (uint32_t j = tid; j < S; j += T) {
        const uint32_t v = base + hist[j];
        hist[j] = v;
        start_out[s0 + j] = v;
    }
    __syncthreads();
    uint32_t n_ins = 0, n_arr = 0, na = 0;
    uint2* alist = alist_all + (tid >> 6) * (64 * U);
    sorted_place<U, T, UNIFORM, WRITE_DIR>(mv, t, s0, a0, a1, hist, alist, na, n_ins, n_arr);
    sorted_place<U, T, UNIFORM, WRITE_DIR>(mv, t, s0, c0, c1, hist, alist, na, n_ins, n_arr);
    if (na) {  // wave-uniform
        sorted_route_list<UNIFORM, WRITE_DIR>(mv, t, alist, na);
        n_arr += __lane_id() == 0 ? na : 0u;
    }
    // block totals of insertions and arrivals, one slot update each
    const uint32_t si = block_sum<T>(n_ins, lds);
    __syncthreads();
    const uint32_t sa = block_sum<T>(n_arr, lds);
    if (tid == 0) {
        slot_add(mv.part_stats, g, kWordInserts, si);
        slot_add(mv.part_stats, g, kWordArrivals, sa);
    }
    *base_out = base;
    *total_out = total;
}

// Sorted move of one table (grid: n_seg workgroups; dynamic LDS: the
// segment's counters followed by the per-wave arrival lists).
template <uint32_t U, uint32_t T, bool UNIFORM, bool WRITE_DIR>
static __global__ __launch_bounds__(T) void k_move_sorted(SortedMove mv, BinDev b) {
    extern __shared__ uint32_t hist[];
    __shared__ uint32_t lds[T / 64];
    __shared__ uint32_t sh[3];
    take_segment(b.ticket, b.n_seg, sh);
    const uint32_t seg = sh[0];
    uint2* alist = reinterpret_cast<uint2*>(hist + (1u << b.seg_log));
    uint32_t base, total;
    sorted_segment<U, T, UNIFORM, WRITE_DIR>(mv, table_view(b), seg << b.seg_log, hist, alist, lds, &sh[2],
                                             b.new_start, b.status, seg, 0u, 0u, sh[1], &base, &total);
    if (seg == b.n_seg - 1 && threadIdx.x == 0) {
        b.new_start[b.m] = base + total;
        if (base + total != b.n) {
            atomicOr(mv.faults, kFaultBinCount);
        }
    }
}

// nxt for every slot of one table: the bucket of the agent's next position.
template <bool UNIFORM>
static __global__ __launch_bounds__(kBlock) void k_sorted_next(SortedBufs st, uint32_t n, float spd_const,
                                                               float inv_cell, uint32_t lw, uint32_t wmask,
                                                               uint32_t hmask) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) {
        return;
    }
    const float2 np = next_position(st.pos[i], st.tgt[i], UNIFORM ? spd_const : st.spd[i]);
    st.nxt[i] = bucket_of(cell_coord(np.x, inv_cell), cell_coord(np.y, inv_cell), lw, wmask, hmask);
}

// dst[perm[i]] = src[i]
template <class V>
static __global__ __launch_bounds__(kBlock) void k_scatter_perm(const uint32_t* __restrict__ perm,
                                                                const V* __restrict__ src, V* __restrict__ dst,
                                                                uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        dst[perm[i]] = src[i];
    }
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

}  // namespace msim
