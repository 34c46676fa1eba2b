// msim_kernels.hpp — the tick kernels (advance+bin, bucket scan, scatter,
// neighbour pass, spatial reorder) and the small helpers they share.
//
// One tick of Simulator::sim_tick (src/sim/Simulator.cpp:213-279) is the move
// dispatch followed by the collision dispatch of random_move.comp:860-879.
// Here the move dispatch is
//     k_advance_bin  -> advance every agent, bin it into the toroidal bucket
//                       table (per-bucket count + rank of the agent in it)
//     k_scan_reduce / k_scan_partials / k_scan_down
//                    -> exclusive scan of the bucket counts (bucket starts),
//                       clearing the counts for the next tick
//     k_scatter      -> binned copy of positions and agent ids
// and the collision dispatch is k_collide (3x3 cell search per agent).
// The quadtree of random_move.comp:100-539 is replaced by this uniform grid;
// only the neighbour set it yields is observable (pair counts), and that set
// is identical.
#pragma once

#include "msim_device.hpp"

namespace msim {

constexpr uint32_t kBlock = 256;                   // 4 wave64s
constexpr uint32_t kScanItems = 16;                // buckets per thread in the scan
constexpr uint32_t kScanTile = kBlock * kScanItems;  // 4096 buckets per workgroup
constexpr uint32_t kPartialsThreads = 1024;

struct AgentsDev {
    uint32_t n;
    float2* pos;
    float2* tgt;
    float2* dir;
    float* spd;        // nullptr when every agent uses spd_const
    float spd_const;
    uint32_t* road;
    uint4* rng;
    uint32_t* init;    // Entity.initialized
    uint32_t* ncount;  // neighbours found in the last collision dispatch
    uint32_t* perm;    // internal slot -> caller's agent index
};

struct GridDev {
    float inv_cell;
    float radius;
    float d2_limit;
    uint32_t lw, wmask, hmask;
    uint32_t m;        // buckets = (wmask + 1) * (hmask + 1)
    uint32_t* cnt;     // [m]   agents per bucket (zero between ticks)
    uint32_t* start;   // [m+1] exclusive prefix of cnt
    uint32_t* part;    // [m / kScanTile] per-tile sums, then their prefix
    uint32_t* rank;    // [n]   agent's position inside its bucket
    float2* bpos;      // [n]   positions in bucket order
    uint32_t* bid;     // [n]   internal agent index in bucket order
};

// Device-side counters.  Each is updated by exactly one thread of a
// single-workgroup kernel (k_scan_partials, k_collide_finish) from per-block
// partial sums, so no kernel contends on them with atomics.
struct Counters {
    unsigned long long inserts;      // spatial-index insertions (debugData[0])
    unsigned long long pairs_total;  // cumulative pairs (debugData[1], 64-bit)
    unsigned long long arrivals;     // agents that reached their target
    unsigned long long collides;     // collision dispatches completed
    unsigned int faults;             // kFault* bits raised by device-side checks
    unsigned int pad;
};

enum FaultBits : unsigned int {
    kFaultBinCount = 1u  // binned agents != agents (an agent left its rebinning window)
};

enum AdvanceMode : int {
    kMoveAll = 0,   // every agent is initialised: advance all
    kMoveInit = 1,  // per-agent Entity.initialized: 0 -> insert only, set to 1
    kBinOnly = 2    // no motion, just rebuild the bucket table
};

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(v, o, 64);
        if (lane >= o) {
            v += y;
        }
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_xor(v, o, 64);
    }
    return v;
}

// Exclusive scan across a workgroup of NT threads; lds holds NT/64 words.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
    const int lane = __lane_id();
    const int w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) {
        lds[w] = inc;
    }
    __syncthreads();
    uint32_t woff = 0;
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < NT / 64; k++) {
        const uint32_t t = lds[k];
        woff += (k < w) ? t : 0u;
        tot += t;
    }
    total = tot;
    __syncthreads();
    return woff + inc - v;
}

template <int NT>
__device__ __forceinline__ unsigned long long block_sum64(unsigned long long v, unsigned long long* lds) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_xor(v, o, 64);
    }
    if (__lane_id() == 0) {
        lds[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    unsigned long long s = 0;
#pragma unroll
    for (int k = 0; k < NT / 64; k++) {
        s += lds[k];
    }
    return s;
}

template <int NT>
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* lds) {
    v = wave_sum(v);
    if (__lane_id() == 0) {
        lds[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < NT / 64; k++) {
        s += lds[k];
    }
    return s;
}

// ---- move dispatch ------------------------------------------------------

// blk_stats[2 * block] / [2 * block + 1]: insertions / arrivals of the block
// (folded into the counters by k_scan_partials).
template <int MODE, bool UNIFORM, bool WRITE_DIR>
__global__ __launch_bounds__(kBlock) void k_advance_bin(AgentsDev a, GridDev g,
                                                        const msim_road* __restrict__ roads,
                                                        const uint32_t* __restrict__ conns,
                                                        uint32_t* __restrict__ blk_stats) {
    __shared__ uint32_t lds[2 * (kBlock / 64)];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = blk * kBlock + threadIdx.x;
    int inserted = 0;
    int arrived = 0;
    if (i < a.n) {
        const float2 p = a.pos[i];
        float2 np = p;
        bool moves = (MODE == kMoveAll);
        if (MODE == kMoveInit) {
            moves = a.init[i] != 0u;
            if (!moves) {
                a.init[i] = 1u;
                inserted = 1;  // first insert (random_move.comp:863-867)
            }
        }
        if (moves) {
            float2 t = a.tgt[i];
            const float s = UNIFORM ? a.spd_const : a.spd[i];
            float2 d = heading(t, p, s);                 // :870
            const float dist = vlen(p.x - t.x, p.y - t.y);  // :842
            if (dist > s) {
                np = make_float2(p.x + d.x, p.y + d.y);  // :844-846
            } else {
                np = t;  // :848-850 snap, route choice, new heading
                uint32_t r = a.road[i];
                uint4 st = a.rng[i];
                choose_route(t, r, st, roads, conns);
                a.tgt[i] = t;
                a.road[i] = r;
                a.rng[i] = st;
                d = heading(t, np, s);
                arrived = 1;
            }
            a.pos[i] = np;
            if (WRITE_DIR) {
                a.dir[i] = d;
            }
        }
        const int32_t cx = cell_coord(np.x, g.inv_cell);
        const int32_t cy = cell_coord(np.y, g.inv_cell);
        if (MODE != kBinOnly && moves) {
            inserted = static_cast<int>(cx != cell_coord(p.x, g.inv_cell)) |
                       static_cast<int>(cy != cell_coord(p.y, g.inv_cell));
        }
        const uint32_t b = bucket_of(cx, cy, g.lw, g.wmask, g.hmask);
        g.rank[i] = atomicAdd(&g.cnt[b], 1u);
    }
    if (MODE != kBinOnly) {
        const uint32_t wi = static_cast<uint32_t>(__popcll(__ballot(inserted)));
        const uint32_t wa = static_cast<uint32_t>(__popcll(__ballot(arrived)));
        const uint32_t w = threadIdx.x >> 6;
        if (__lane_id() == 0) {
            lds[w] = wi;
            lds[kBlock / 64 + w] = wa;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t si = 0, sa = 0;
#pragma unroll
            for (uint32_t k = 0; k < kBlock / 64; k++) {
                si += lds[k];
                sa += lds[kBlock / 64 + k];
            }
            blk_stats[2 * blk] = si;
            blk_stats[2 * blk + 1] = sa;
        }
    }
}

// Per-tile bucket totals.
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ cnt,
                                                        uint32_t* __restrict__ part) {
    __shared__ uint32_t lds[kBlock / 64];
    const uint4* c4 = reinterpret_cast<const uint4*>(cnt + static_cast<size_t>(blockIdx.x) * kScanTile);
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanItems / 4; k++) {
        const uint4 v = c4[threadIdx.x + k * kBlock];
        s += v.x + v.y + v.z + v.w;
    }
    s = block_sum<kBlock>(s, lds);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s;
    }
}

// Exclusive scan of the tile totals (one workgroup); start[m] = total.  Also
// folds the advance kernel's n_stat per-block statistics into the counters.
__global__ __launch_bounds__(kPartialsThreads) void k_scan_partials(uint32_t* __restrict__ part,
                                                                    uint32_t n_part,
                                                                    uint32_t* __restrict__ start_end,
                                                                    const uint32_t* __restrict__ blk_stats,
                                                                    uint32_t n_stat,
                                                                    Counters* __restrict__ ctr) {
    __shared__ uint32_t lds[kPartialsThreads / 64];
    if (n_stat) {
        unsigned long long si = 0, sa = 0;
        for (uint32_t k = threadIdx.x; k < n_stat; k += kPartialsThreads) {
            si += blk_stats[2 * k];
            sa += blk_stats[2 * k + 1];
        }
        const uint32_t ti = block_sum<kPartialsThreads>(static_cast<uint32_t>(si), lds);
        __syncthreads();
        const uint32_t ta = block_sum<kPartialsThreads>(static_cast<uint32_t>(sa), lds);
        __syncthreads();
        if (threadIdx.x == 0) {
            ctr->inserts += ti;
            ctr->arrivals += ta;
        }
    }
    const uint32_t per = (n_part + kPartialsThreads - 1) / kPartialsThreads;
    const uint32_t lo = threadIdx.x * per;
    uint32_t s = 0;
    for (uint32_t k = 0; k < per; k++) {
        const uint32_t j = lo + k;
        s += (j < n_part) ? part[j] : 0u;
    }
    uint32_t total = 0;
    uint32_t off = block_excl_scan<kPartialsThreads>(s, lds, total);
    for (uint32_t k = 0; k < per; k++) {
        const uint32_t j = lo + k;
        if (j < n_part) {
            const uint32_t v = part[j];
            part[j] = off;
            off += v;
        }
    }
    if (threadIdx.x == 0) {
        *start_end = total;
    }
}

// Bucket starts for one tile; clears the counts for the next tick.
__global__ __launch_bounds__(kBlock) void k_scan_down(uint32_t* __restrict__ cnt,
                                                      const uint32_t* __restrict__ part,
                                                      uint32_t* __restrict__ start) {
    __shared__ uint32_t lds[kBlock / 64];
    const size_t base = static_cast<size_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    uint4* c4 = reinterpret_cast<uint4*>(cnt + base);
    uint4 v[kScanItems / 4];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanItems / 4; k++) {
        v[k] = c4[k];
        s += v[k].x + v[k].y + v[k].z + v[k].w;
    }
    uint32_t total = 0;
    uint32_t off = block_excl_scan<kBlock>(s, lds, total) + part[blockIdx.x];
    uint4* s4 = reinterpret_cast<uint4*>(start + base);
#pragma unroll
    for (uint32_t k = 0; k < kScanItems / 4; k++) {
        uint4 o;
        o.x = off;
        off += v[k].x;
        o.y = off;
        off += v[k].y;
        o.z = off;
        off += v[k].z;
        o.w = off;
        off += v[k].w;
        s4[k] = o;
        c4[k] = make_uint4(0u, 0u, 0u, 0u);
    }
}

__global__ __launch_bounds__(kBlock) void k_scatter(AgentsDev a, GridDev g) {
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= a.n) {
        return;
    }
    const float2 p = a.pos[i];
    const uint32_t b = bucket_of(cell_coord(p.x, g.inv_cell), cell_coord(p.y, g.inv_cell), g.lw,
                                 g.wmask, g.hmask);
    const uint32_t dst = g.start[b] + g.rank[i];
    g.bpos[dst] = p;
    g.bid[dst] = i;
}

// ---- collision dispatch -------------------------------------------------

// Neighbour pass: for the agent in bucket slot s, count every other agent
// whose position satisfies quad_tree_in_range (random_move.comp:551-562)
// among the 3x3 cells around it (the cell edge is at least R, so this covers
// the radius).  With SLOW the pair (i, j), i > j in caller order, counts only
// when agent i searches (Entity.initialized was 1 at dispatch start), which is
// the reference's "j < i" rule at random_move.comp:574.  The last workgroup to
// finish folds the dispatch's pair total into the counters and history.
template <bool SLOW>
__global__ __launch_bounds__(kBlock) void k_collide(AgentsDev a, GridDev g, uint32_t* __restrict__ blk_pairs) {
    __shared__ uint32_t lds[kBlock / 64];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t s = blk * kBlock + threadIdx.x;
    uint32_t cnt = 0;
    uint32_t own = 0;
    if (s < a.n) {
        const float2 me = g.bpos[s];
        const int32_t cx = cell_coord(me.x, g.inv_cell);
        const int32_t cy = cell_coord(me.y, g.inv_cell);
        const uint32_t me_id = g.bid[s];
        uint32_t me_ext = 0;
        bool me_search = true;
        if (SLOW) {
            me_ext = a.perm[me_id];
            me_search = a.init[me_id] != 0u;
        }
        const float R = g.radius;
        const float T = g.d2_limit;
#pragma unroll 1
        for (int dy = -1; dy <= 1; dy++) {
            const uint32_t rowb = (static_cast<uint32_t>(cy + dy) & g.hmask) << g.lw;
            const uint32_t c0 = static_cast<uint32_t>(cx - 1) & g.wmask;
            const uint32_t c2 = static_cast<uint32_t>(cx + 1) & g.wmask;
            uint32_t r0 = g.start[rowb + c0];
            uint32_t r1;
            uint32_t r2 = 0;
            uint32_t r3 = 0;
            if (c0 < c2) {
                r1 = g.start[rowb + c2 + 1];
            } else {  // the 3 columns wrap around the table edge
                r1 = g.start[rowb + g.wmask + 1];
                r2 = g.start[rowb];
                r3 = g.start[rowb + c2 + 1];
            }
#pragma unroll 1
            for (int seg = 0; seg < 2; seg++) {
                const uint32_t lo = seg ? r2 : r0;
                const uint32_t hi = seg ? r3 : r1;
                for (uint32_t k = lo; k < hi; k++) {
                    const float2 q = g.bpos[k];
                    const bool hit = in_range(q, me, R, T) & (k != s);
                    if (!SLOW) {
                        cnt += hit ? 1u : 0u;
                    } else if (hit) {
                        const uint32_t j = g.bid[k];
                        const uint32_t je = a.perm[j];
                        const bool counted = (je < me_ext) ? me_search : (a.init[j] != 0u);
                        if (counted) {
                            cnt++;
                            own += (je < me_ext) ? 1u : 0u;
                        }
                    }
                }
            }
        }
        a.ncount[me_id] = cnt;
    }
    const uint32_t bs = block_sum<kBlock>(SLOW ? own : cnt, lds);
    if (threadIdx.x == 0) {
        blk_pairs[blk] = bs;
    }
}

// Folds k_collide's per-block sums into the dispatch's pair total: every pair
// is found from both ends in the fast path, once in the SLOW path.
// Folds n_stat per-block advance statistics (insertions, arrivals) into the
// counters; one workgroup.
__device__ __forceinline__ void fold_stats(const uint32_t* __restrict__ blk_stats, uint32_t n_stat,
                                           Counters* __restrict__ ctr, unsigned long long* lds) {
    unsigned long long si = 0, sa = 0;
    for (uint32_t k = threadIdx.x; k < n_stat; k += kPartialsThreads) {
        si += blk_stats[2 * k];
        sa += blk_stats[2 * k + 1];
    }
    const unsigned long long ti = block_sum64<kPartialsThreads>(si, lds);
    __syncthreads();
    const unsigned long long ta = block_sum64<kPartialsThreads>(sa, lds);
    __syncthreads();
    if (threadIdx.x == 0) {
        ctr->inserts += ti;
        ctr->arrivals += ta;
    }
}

__global__ __launch_bounds__(kPartialsThreads) void k_fold_stats(const uint32_t* __restrict__ blk_stats,
                                                                 uint32_t n_stat, Counters* __restrict__ ctr) {
    __shared__ unsigned long long lds[kPartialsThreads / 64];
    fold_stats(blk_stats, n_stat, ctr, lds);
}

// Also folds n_stat pending advance statistics (0: none pending).
template <bool SLOW>
__global__ __launch_bounds__(kPartialsThreads) void k_collide_finish(const uint32_t* __restrict__ blk_pairs,
                                                                     uint32_t nb, Counters* __restrict__ ctr,
                                                                     uint32_t* __restrict__ hist,
                                                                     uint32_t hist_mask,
                                                                     const uint32_t* __restrict__ blk_stats,
                                                                     uint32_t n_stat) {
    __shared__ unsigned long long lds[kPartialsThreads / 64];
    if (n_stat) {
        fold_stats(blk_stats, n_stat, ctr, lds);
    }
    unsigned long long s = 0;
    for (uint32_t k = threadIdx.x; k < nb; k += kPartialsThreads) {
        s += blk_pairs[k];
    }
    const unsigned long long tot = block_sum64<kPartialsThreads>(s, lds);
    if (threadIdx.x == 0) {
        const unsigned long long pairs = SLOW ? tot : tot / 2;
        ctr->pairs_total += pairs;
        hist[ctr->collides & hist_mask] = static_cast<uint32_t>(pairs);
        ctr->collides += 1;
    }
}

__global__ __launch_bounds__(kBlock) void k_set_init(uint32_t* init, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        init[i] = 1u;
    }
}

// ---- spatial renumbering -------------------------------------------------

// Gathers every per-agent array into bucket order so that agents that are
// close in space are close in memory; afterwards bid is the identity.
__global__ __launch_bounds__(kBlock) void k_reorder(AgentsDev src, AgentsDev dst, GridDev g) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= src.n) {
        return;
    }
    const uint32_t j = g.bid[s];
    dst.pos[s] = src.pos[j];
    dst.tgt[s] = src.tgt[j];
    dst.dir[s] = src.dir[j];
    if (src.spd) {
        dst.spd[s] = src.spd[j];
    }
    dst.road[s] = src.road[j];
    dst.rng[s] = src.rng[j];
    dst.init[s] = src.init[j];
    dst.ncount[s] = src.ncount[j];
    dst.perm[s] = src.perm[j];
    g.bid[s] = s;
}

// Per-agent cell export in caller order.
__global__ __launch_bounds__(kBlock) void k_export_cells(AgentsDev a, float inv_cell, int2* out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) {
        return;
    }
    const float2 p = a.pos[i];
    out[a.perm[i]] = make_int2(cell_coord(p.x, inv_cell), cell_coord(p.y, inv_cell));
}

}  // namespace msim
