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
    kFaultBinCount = 1u,  // binned agents != agents (an agent left its rebinning window)
    kFaultLookback = 2u   // a segment's look-back gave up waiting for a predecessor
};

// Per-dispatch partial sums are accumulated into kSlots 64-bit slots, one per
// 128-byte line, by one atomic per workgroup on slot (block % kSlots); the
// single-workgroup finish/fold passes read the kSlots values and clear them.
constexpr uint32_t kSlots = 64;
constexpr uint32_t kSlotStride = 16;  // 64-bit words between slots
enum SlotWord : uint32_t { kWordPairs = 0, kWordInserts = 0, kWordArrivals = 1 };

__device__ __forceinline__ void slot_add(unsigned long long* part, uint32_t blk, uint32_t word,
                                         unsigned long long v) {
    if (v) {
        atomicAdd(&part[(blk % kSlots) * kSlotStride + word], v);
    }
}

// Adds 1 to table[key] for every lane with `active` set and returns the lane's
// pre-increment value (its rank within the key).  Lanes of a wave that sit
// in the same heavy cell share keys: when the active keys form at most 8
// runs, the wave issues one atomic per distinct key; otherwise one per lane.
// Must be called by the whole wave (inactive lanes pass active = false).
__device__ __forceinline__ uint32_t wave_rank_add(uint32_t* table, uint32_t key, bool active) {
    const uint32_t lane = __lane_id();
    const unsigned long long act = __ballot(active);
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool prev_active = lane ? ((act >> (lane - 1)) & 1ull) != 0 : false;
    const unsigned long long heads = __ballot(active && (!prev_active || key != prev));
    if (__popcll(heads) > 8) {
        return active ? atomicAdd(&table[key], 1u) : 0u;
    }
    uint32_t rank = 0;
    unsigned long long todo = act;
    while (todo) {
        const int leader = __ffsll(static_cast<long long>(todo)) - 1;
        const uint32_t k = __shfl(key, leader, 64);
        const unsigned long long peers = __ballot(active && key == k) & todo;
        uint32_t base = 0;
        if (static_cast<int>(lane) == leader) {
            base = atomicAdd(&table[k], static_cast<uint32_t>(__popcll(peers)));
        }
        base = __shfl(base, leader, 64);
        if ((peers >> lane) & 1ull) {
            rank = base + static_cast<uint32_t>(__popcll(peers & ((1ull << lane) - 1ull)));
        }
        todo &= ~peers;
    }
    return rank;
}

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

// Insertions / arrivals of the block go to part_stats (kWordInserts,
// kWordArrivals of the block's slot).
template <int MODE, bool UNIFORM, bool WRITE_DIR>
static __global__ __launch_bounds__(kBlock) void k_advance_bin(AgentsDev a, GridDev g,
                                                        const msim_road* __restrict__ roads,
                                                        const uint32_t* __restrict__ conns,
                                                        unsigned long long* __restrict__ part_stats) {
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
            slot_add(part_stats, blk, kWordInserts, si);
            slot_add(part_stats, blk, kWordArrivals, sa);
        }
    }
}

// Per-tile bucket totals.
static __global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ cnt,
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

// Exclusive scan of the tile totals (one workgroup); start[m] = total.
static __global__ __launch_bounds__(kPartialsThreads) void k_scan_partials(uint32_t* __restrict__ part,
                                                                    uint32_t n_part,
                                                                    uint32_t* __restrict__ start_end) {
    __shared__ uint32_t lds[kPartialsThreads / 64];
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
static __global__ __launch_bounds__(kBlock) void k_scan_down(uint32_t* __restrict__ cnt,
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

static __global__ __launch_bounds__(kBlock) void k_scatter(AgentsDev a, GridDev g) {
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
// Slots within the d2 limit of `me` over the 3x3 cells around (cx, cy) of a
// bucket table (the agent's own slot `self` included).  Latency first: the 12
// row bounds are loaded together, then the first kNbBatch candidates of the
// three main row ranges together; longer ranges and the ranges of rows that
// wrap around the table's edge column continue in batches.  Loads past a
// range's end read the own slot instead (always valid) and are not counted,
// so every load is unconditional.
constexpr uint32_t kNbBatch = 4;

__device__ __forceinline__ uint32_t count_in_cells(float2 me, uint32_t self, int32_t cx, int32_t cy,
                                                   const uint32_t* __restrict__ start, uint32_t lw,
                                                   uint32_t wmask, uint32_t hmask,
                                                   const float2* __restrict__ bpos, float T) {
    uint32_t lo[3], hi[3], wlo[3], whi[3];
    const uint32_t c0 = static_cast<uint32_t>(cx - 1) & wmask;
    const uint32_t c2 = static_cast<uint32_t>(cx + 1) & wmask;
    const bool wrap = c0 > c2;  // the 3 columns cross the table's edge
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const uint32_t rowb = (static_cast<uint32_t>(cy + r - 1) & hmask) << lw;
        lo[r] = start[rowb + c0];
        hi[r] = start[rowb + (wrap ? wmask + 1 : c2 + 1)];
        wlo[r] = start[rowb];
        whi[r] = start[rowb + (wrap ? c2 + 1 : 0)];
    }
    uint32_t cnt = 0;
    float2 q[3][kNbBatch];
#pragma unroll
    for (int r = 0; r < 3; r++) {
#pragma unroll
        for (uint32_t u = 0; u < kNbBatch; u++) {
            const uint32_t k = lo[r] + u;
            q[r][u] = bpos[k < hi[r] ? k : self];
        }
    }
#pragma unroll
    for (int r = 0; r < 3; r++) {
#pragma unroll
        for (uint32_t u = 0; u < kNbBatch; u++) {
            cnt += (lo[r] + u < hi[r] && in_range(q[r][u], me, T)) ? 1u : 0u;
        }
    }
#pragma unroll
    for (int g = 0; g < 6; g++) {  // unrolled: the row arrays stay in registers
        const int r = g >> 1;
        uint32_t k = (g & 1) ? wlo[r] : lo[r] + kNbBatch;
        const uint32_t h = (g & 1) ? whi[r] : hi[r];
        for (; k < h; k += kNbBatch) {
            float2 v[kNbBatch];
#pragma unroll
            for (uint32_t u = 0; u < kNbBatch; u++) {
                v[u] = bpos[k + u < h ? k + u : self];
            }
#pragma unroll
            for (uint32_t u = 0; u < kNbBatch; u++) {
                cnt += (k + u < h && in_range(v[u], me, T)) ? 1u : 0u;
            }
        }
    }
    return cnt;
}

// Orders this wave's LDS writes before its later LDS reads (other lanes'
// data), and its reads before the next writes; no other wave takes part.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr uint32_t kStage = 512;  // binned positions staged per wave round (4 x 16 B per lane, 4 KB)
constexpr uint32_t kStageLoads = kStage / 128;

// Staged candidates [j0, j1) of one lane within range of `me`; two LDS reads
// are issued before either is used.  (Four per round measured slower: the
// lanes' remainders diverge.)
__device__ __forceinline__ uint32_t count_staged(const float2* stage, uint32_t j0, uint32_t j1, float2 me,
                                                 float T) {
    // one induction variable (the LDS address) and one compare per round
    uint32_t cnt = 0;
    const float2* q = stage + j0;
    const float2* const pair_end = stage + (j0 + ((j1 - j0) & ~1u));
    for (; q != pair_end; q += 2) {
        const float2 q0 = q[0];
        const float2 q1 = q[1];
        cnt += (in_range(q0, me, T) ? 1u : 0u) + (in_range(q1, me, T) ? 1u : 0u);
    }
    if ((j1 - j0) & 1u) {
        cnt += in_range(*q, me, T) ? 1u : 0u;
    }
    return cnt;
}

// Stages up to kStage binned positions: candidate c (even) of the round
// comes from slot a0 + c (c < off1), a1 + c - off1 (c < off2) or a2 + c -
// off2.  Every lane issues its kStageLoads 16-byte loads and LDS writes
// unconditionally (past U they re-read slot a0 and fill the unused rest of
// the kStage-slot area), all loads before the first write.  A load may read
// the slot after a row's end (at most slot n, inside the n + 2 allocation);
// that candidate is never scanned.
__device__ __forceinline__ void stage_rows(float2* stage, const float2* __restrict__ bpos, uint32_t U, uint32_t a0,
                                           uint32_t a1, uint32_t a2, uint32_t off1, uint32_t off2) {
    const uint32_t lane = __lane_id();
    float4 v[kStageLoads];
#pragma unroll
    for (uint32_t u = 0; u < kStageLoads; u++) {
        const uint32_t c = 2u * (u * 64u + lane);
        const uint32_t cc = c < U ? c : 0u;
        const uint32_t slot = cc < off1 ? a0 + cc : (cc < off2 ? a1 + (cc - off1) : a2 + (cc - off2));
        v[u] = *reinterpret_cast<const float4*>(bpos + slot);
    }
#pragma unroll
    for (uint32_t u = 0; u < kStageLoads; u++) {
        *reinterpret_cast<float4*>(stage + 2u * (u * 64u + lane)) = v[u];
    }
}

// count_in_cells for a whole wave whose lanes hold P slots each: slot
// w0 + 64 p + lane in element p.  The valid slots (consecutive slots of one
// bucket table, the first in element first_p of lane first_lane, the last
// in element last_p of lane last_lane) have nondecreasing buckets; their
// counts are added to cnt.  Each slot's main
// range in a row, its columns clipped to the row ([col-1, col+1] within
// [0, wmask]), then starts and ends no earlier than the previous slot's, so
// the first slot's start and the last slot's end bound the wave's
// candidates in that row.  When the three rows' unions fit kStage slots they
// are loaded in one round (one 16-byte load per lane per 128 slots, all
// issued before any is used) into LDS and every lane scans its slots'
// ranges there; larger unions are staged row by row, kStage slots a round.
// Rows where a slot's neighbour row wraps around the table's top or bottom
// edge, and the wrapped column of slots in the first or last table column,
// are read per lane from global memory.  Each slot's own position is
// included in its count.
template <int P>
__device__ __forceinline__ void wave_count_in_cells(const float2 (&me)[P], const bool (&valid)[P], uint32_t first_p,
                                                    uint32_t first_lane, uint32_t last_p, uint32_t last_lane,
                                                    const int32_t (&cx)[P],
                                                    const int32_t (&cy)[P], const uint32_t* __restrict__ start,
                                                    uint32_t lw, uint32_t wmask, uint32_t hmask,
                                                    const float2* __restrict__ bpos, float T, float2* stage,
                                                    uint32_t (&cnt)[P]) {
    uint32_t col[P], row[P], lo[P][3], hi[P][3];
#pragma unroll
    for (int p = 0; p < P; p++) {
        col[p] = static_cast<uint32_t>(cx[p]) & wmask;
        row[p] = static_cast<uint32_t>(cy[p]) & hmask;
        const uint32_t cl = col[p] ? col[p] - 1u : 0u;
        const uint32_t ch = col[p] < wmask ? col[p] + 1u : wmask;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const uint32_t rowb = ((row[p] + static_cast<uint32_t>(r - 1)) & hmask) << lw;
            lo[p][r] = start[rowb + cl];
            hi[p][r] = start[rowb + ch + 1u];
        }
    }
    // The column across the table's edge (slots in its first or last column,
    // when the table has more than one): bounds loaded with the others.
    uint32_t elo[P][3], ehi[P][3];
    bool ecol[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        ecol[p] = valid[p] && (col[p] == 0u || col[p] == wmask) && wmask != 0u;
        const uint32_t wc = col[p] == 0u ? wmask : 0u;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            elo[p][r] = 0u;
            ehi[p][r] = 0u;
            if (ecol[p]) {
                const uint32_t rowb = ((row[p] + static_cast<uint32_t>(r - 1)) & hmask) << lw;
                elo[p][r] = start[rowb + wc];
                ehi[p][r] = start[rowb + wc + 1u];
            }
        }
    }
    // wave-uniform row blocks: first slot a (even), staged length n (even),
    // last staged slot that is a candidate ub
    uint32_t a[3], n[3], ub[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        bool edge = false;
#pragma unroll
        for (int p = 0; p < P; p++) {
            edge = edge || (valid[p] && ((r == 0 && row[p] == 0u) || (r == 2 && row[p] == hmask)));
        }
        n[r] = 0u;
        a[r] = 0u;
        ub[r] = 0u;
        if (__ballot(edge) != 0ull) {  // the wave's rows are not monotone here
#pragma unroll
            for (int p = 0; p < P; p++) {
                if (valid[p]) {
                    for (uint32_t k = lo[p][r]; k < hi[p][r]; k++) {
                        cnt[p] += in_range(bpos[k], me[p], T) ? 1u : 0u;
                    }
                }
            }
            continue;
        }
        uint32_t lf = lo[0][r], hl = hi[0][r];
#pragma unroll
        for (int p = 1; p < P; p++) {
            lf = (first_p == static_cast<uint32_t>(p)) ? lo[p][r] : lf;
            hl = (last_p == static_cast<uint32_t>(p)) ? hi[p][r] : hl;
        }
        const uint32_t u_lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(lf),
                                                                              static_cast<int>(first_lane)));
        const uint32_t u_hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(hl),
                                                                              static_cast<int>(last_lane)));
        a[r] = u_lo & ~1u;
        n[r] = (u_hi - a[r] + 1u) & ~1u;
        ub[r] = u_hi;
    }
    const uint32_t off[3] = {0u, n[0], n[0] + n[1]};
    const uint32_t U = off[2] + n[2];
    if (U <= kStage) {
        stage_rows(stage, bpos, U, a[0], a[1], a[2], off[1], off[2]);
        wave_lds_sync();
#pragma unroll
        for (int p = 0; p < P; p++) {
            if (valid[p]) {
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    if (n[r]) {
                        cnt[p] += count_staged(stage, off[r] + lo[p][r] - a[r], off[r] + hi[p][r] - a[r], me[p], T);
                    }
                }
            }
        }
        // The edge column's candidates usually lie inside the staged rows
        // (narrow tables: the wave's rows are staged whole); those are
        // counted from LDS and marked done.
#pragma unroll
        for (int p = 0; p < P; p++) {
#pragma unroll
            for (int r = 0; r < 3; r++) {
                if (ecol[p] && n[r] && elo[p][r] >= a[r] && ehi[p][r] <= ub[r]) {
                    cnt[p] += count_staged(stage, off[r] + elo[p][r] - a[r], off[r] + ehi[p][r] - a[r], me[p], T);
                    ehi[p][r] = elo[p][r];
                }
            }
        }
        wave_lds_sync();
    } else {
#pragma unroll
        for (int r = 0; r < 3; r++) {
#pragma unroll 1
            for (uint32_t base = a[r]; base < a[r] + n[r]; base += kStage) {
                const uint32_t m = min(kStage, a[r] + n[r] - base);
                stage_rows(stage, bpos, m, base, base, base, kStage, kStage);
                wave_lds_sync();
#pragma unroll
                for (int p = 0; p < P; p++) {
                    if (valid[p]) {
                        const uint32_t k0 = max(lo[p][r], base);
                        const uint32_t k1 = min(hi[p][r], base + m);
                        if (k0 < k1) {
                            cnt[p] += count_staged(stage, k0 - base, k1 - base, me[p], T);
                        }
                    }
                }
                wave_lds_sync();
            }
        }
    }
#pragma unroll
    for (int p = 0; p < P; p++) {
        if (ecol[p]) {  // the rest of the column across the table's edge, from global memory
#pragma unroll
            for (int r = 0; r < 3; r++) {
                for (uint32_t k = elo[p][r]; k < ehi[p][r]; k++) {
                    cnt[p] += in_range(bpos[k], me[p], T) ? 1u : 0u;
                }
            }
        }
    }
}

// Staged neighbour pass (collision dispatch of initialised agents): each
// wave takes 64 P consecutive slots, P per lane (grid: ceil(n / (kBlock P))
// workgroups).  At most 64 VGPRs (8 waves per SIMD).
template <int P>
static __global__ __launch_bounds__(kBlock, 8) void k_collide_staged(AgentsDev a, GridDev g,
                                                              const float2* __restrict__ bpos,
                                                              unsigned long long* __restrict__ part_pairs) {
    __shared__ uint32_t lds[kBlock / 64];
    __shared__ float2 stage_all[kBlock / 64 * kStage];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t w0 = (blk * (kBlock / 64) + (threadIdx.x >> 6)) * (64u * P);
    const float T = g.d2_limit;
    float2 me[P];
    bool valid[P];
    int32_t cx[P], cy[P];
    uint32_t id[P], cnt[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        const uint32_t s = w0 + 64u * p + __lane_id();
        valid[p] = s < a.n;
        me[p] = valid[p] ? bpos[s] : make_float2(0.0f, 0.0f);
        id[p] = valid[p] ? (g.bid ? g.bid[s] : s) : 0u;  // no slot map: the state is in slot order
        cx[p] = cell_coord(me[p].x, g.inv_cell);
        cy[p] = cell_coord(me[p].y, g.inv_cell);
        cnt[p] = 0u;
    }
    if (w0 < a.n) {
        const uint32_t last = min(64u * P, a.n - w0) - 1u;
        wave_count_in_cells<P>(me, valid, 0u, 0u, last >> 6, last & 63u, cx, cy, g.start, g.lw, g.wmask, g.hmask,
                               bpos, T, stage_all + (threadIdx.x >> 6) * kStage, cnt);
    }
    uint32_t tot = 0;
#pragma unroll
    for (int p = 0; p < P; p++) {
        if (valid[p]) {
            const uint32_t c = cnt[p] - 1u;  // the agent's own slot passed the predicate
            a.ncount[id[p]] = c;
            tot += c;
        }
    }
    const uint32_t bs = block_sum<kBlock>(tot, lds);
    if (threadIdx.x == 0) {
        slot_add(part_pairs, blk, kWordPairs, bs);
    }
}

// Neighbour pass over the binned slots, one slot per lane: count_in_cells.
// SLOW: the initialisation dispatch, where agents with Entity.initialized ==
// 0 do not search (random_move.comp:863-867).
template <bool SLOW>
static __global__ __launch_bounds__(kBlock) void k_collide(AgentsDev a, GridDev g, const float2* __restrict__ bpos,
                                                    unsigned long long* __restrict__ part_pairs) {
    __shared__ uint32_t lds[kBlock / 64];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t s = blk * kBlock + threadIdx.x;
    const bool valid = s < a.n;
    uint32_t cnt = 0;
    uint32_t own = 0;
    const float T = g.d2_limit;
    const float2 me = valid ? bpos[s] : make_float2(0.0f, 0.0f);
    const int32_t cx = cell_coord(me.x, g.inv_cell);
    const int32_t cy = cell_coord(me.y, g.inv_cell);
    if (!SLOW && valid) {
        cnt = count_in_cells(me, s, cx, cy, g.start, g.lw, g.wmask, g.hmask, bpos, T);
    }
    if (SLOW && valid) {
        // initialisation dispatch: a pair counts once, found from the end
        // with the larger caller index if that agent searches, else from the
        // other end if it searches
        const uint32_t me_id = g.bid[s];
        const uint32_t me_ext = a.perm[me_id];
        const bool me_search = a.init[me_id] != 0u;
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
                    const float2 q = bpos[k];
                    if (in_range(q, me, T) && k != s) {
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
    }
    if (!SLOW && valid) {
        cnt -= 1u;  // the agent's own slot passed the predicate
    }
    if (valid) {
        a.ncount[(SLOW || g.bid) ? g.bid[s] : s] = cnt;  // no slot map: the state is in slot order
    }
    const uint32_t bs = block_sum<kBlock>(SLOW ? own : cnt, lds);
    if (threadIdx.x == 0) {
        slot_add(part_pairs, blk, kWordPairs, bs);
    }
}

static __global__ __launch_bounds__(kBlock) void k_set_init(uint32_t* init, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        init[i] = 1u;
    }
}

// Sums word w of the kSlots slots (threads 0..kSlots-1) and clears them;
// the result is valid in thread 0.
__device__ __forceinline__ unsigned long long take_slots(unsigned long long* part, uint32_t w) {
    unsigned long long v = 0;
    if (threadIdx.x < kSlots) {
        unsigned long long* p = &part[threadIdx.x * kSlotStride + w];
        v = *p;
        *p = 0ull;
    }
    if (threadIdx.x < 64) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            v += __shfl_xor(v, o, 64);
        }
    }
    return v;
}

// Folds the pending advance statistics (insertions, arrivals) into the
// counters; launched with one wave.
static __global__ __launch_bounds__(64) void k_fold_stats(unsigned long long* __restrict__ part_stats,
                                                   Counters* __restrict__ ctr) {
    const unsigned long long ins = take_slots(part_stats, kWordInserts);
    const unsigned long long arr = take_slots(part_stats, kWordArrivals);
    if (threadIdx.x == 0) {
        ctr->inserts += ins;
        ctr->arrivals += arr;
    }
}

// Folds k_collide's slots into the dispatch's pair total (every pair is found
// from both ends in the fast path, once in the SLOW path) and, when
// fold_stats is set, the pending advance statistics; launched with one wave.
template <bool SLOW>
static __global__ __launch_bounds__(64) void k_collide_finish(unsigned long long* __restrict__ part_pairs,
                                                       Counters* __restrict__ ctr, uint32_t* __restrict__ hist,
                                                       uint32_t hist_mask,
                                                       unsigned long long* __restrict__ part_stats,
                                                       uint32_t fold_stats) {
    const unsigned long long tot = take_slots(part_pairs, kWordPairs);
    unsigned long long ins = 0, arr = 0;
    if (fold_stats) {
        ins = take_slots(part_stats, kWordInserts);
        arr = take_slots(part_stats, kWordArrivals);
    }
    if (threadIdx.x == 0) {
        const unsigned long long pairs = SLOW ? tot : tot / 2;
        ctr->pairs_total += pairs;
        hist[ctr->collides & hist_mask] = static_cast<uint32_t>(pairs);
        ctr->collides += 1;
        ctr->inserts += ins;
        ctr->arrivals += arr;
    }
}

// ---- spatial renumbering -------------------------------------------------

// Gathers the per-agent arrays into bucket order so that agents that are
// close in space are close in memory.  Workgroup row blockIdx.y handles one
// array or a group of small ones, named by the row's 4-bit code in `codes`
// (kRow*); a row's arrays share its slot-map reads.  Rows are dispatched one
// after the other, so the resident workgroups mostly gather from the same
// arrays and share their cache lines.  Each thread handles
// kReorderU slots kBlock apart and issues all of their slot-map loads, then
// all of their gathers, before the first store: a thread with one slot spends
// two dependent memory round trips on 4 to 16 bytes.  Positions need no
// gather: the binning's bpos[slot] is pos[bid[slot]] already.
// k_bid_identity then resets bid.
constexpr uint32_t kReorderU = 8;
// Row codes: arrays that share a row share its slot-map reads.
constexpr uint32_t kRowPos = 0, kRowTgtRoad = 1, kRowRng = 2, kRowInitPerm = 3, kRowInitPermSpd = 4,
                   kRowDirNcount = 5;

// dst_a[s] = src_a[bid[s]] (and likewise for b, c when given) for this
// thread's kReorderU slots s0 + u * kBlock.
template <class A, class B = uint32_t, class C = uint32_t>
__device__ __forceinline__ void reorder_gather(const uint32_t* __restrict__ bid, uint32_t s0, uint32_t n,
                                               const A* __restrict__ sa, A* __restrict__ da,
                                               const B* __restrict__ sb = nullptr, B* __restrict__ db = nullptr,
                                               const C* __restrict__ sc = nullptr, C* __restrict__ dc = nullptr) {
    uint32_t j[kReorderU];
    A va[kReorderU];
    B vb[kReorderU];
    C vc[kReorderU];
#pragma unroll
    for (uint32_t u = 0; u < kReorderU; u++) {
        j[u] = bid[min(s0 + u * kBlock, n - 1u)];  // s0 < n
    }
#pragma unroll
    for (uint32_t u = 0; u < kReorderU; u++) {
        va[u] = sa[j[u]];
        if (sb) vb[u] = sb[j[u]];
        if (sc) vc[u] = sc[j[u]];
    }
#pragma unroll
    for (uint32_t u = 0; u < kReorderU; u++) {
        if (s0 + u * kBlock < n) {
            da[s0 + u * kBlock] = va[u];
            if (sb) db[s0 + u * kBlock] = vb[u];
            if (sc) dc[s0 + u * kBlock] = vc[u];
        }
    }
}

static __global__ __launch_bounds__(kBlock) void k_reorder(AgentsDev src, AgentsDev dst,
                                                    const uint32_t* __restrict__ bid,
                                                    const float2* __restrict__ bpos, unsigned long long codes) {
    const uint32_t s0 = xcd_block(blockIdx.x, gridDim.x) * (kBlock * kReorderU) + threadIdx.x;
    const uint32_t n = src.n;
    if (s0 >= n) {
        return;
    }
    switch (static_cast<uint32_t>(codes >> (4u * blockIdx.y)) & 15u) {
        case kRowPos:
#pragma unroll
            for (uint32_t u = 0; u < kReorderU; u++) {
                if (s0 + u * kBlock < n) {
                    dst.pos[s0 + u * kBlock] = bpos[s0 + u * kBlock];
                }
            }
            break;
        case kRowTgtRoad: reorder_gather(bid, s0, n, src.tgt, dst.tgt, src.road, dst.road); break;
        case kRowRng: reorder_gather(bid, s0, n, src.rng, dst.rng); break;
        case kRowInitPerm: reorder_gather(bid, s0, n, src.init, dst.init, src.perm, dst.perm); break;
        case kRowInitPermSpd:
            reorder_gather(bid, s0, n, src.init, dst.init, src.perm, dst.perm, src.spd, dst.spd);
            break;
        default: reorder_gather(bid, s0, n, src.dir, dst.dir, src.ncount, dst.ncount); break;
    }
}

// Agent index = slot after a renumbering.
static __global__ __launch_bounds__(kBlock) void k_bid_identity(uint32_t* __restrict__ bid, uint32_t n) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s < n) {
        bid[s] = s;
    }
}

// Launches the renumbering of n agents from src into dst (bid, bpos: the
// current binning's slot -> agent map, reset to the identity, and its
// positions).  live_outputs = false skips dir and ncount: a renumbering
// inside a captured run segment is always followed by the run's last tick,
// whose move rewrites every heading and whose collision dispatch every
// neighbour count.
inline void launch_reorder(const AgentsDev& src, const AgentsDev& dst, uint32_t* bid, const float2* bpos,
                           uint32_t n, hipStream_t stream, bool live_outputs = true) {
    const uint32_t nb = std::max(1u, (n + kBlock * kReorderU - 1) / (kBlock * kReorderU));
    unsigned long long codes = 0;
    uint32_t rows = 0;
    auto row = [&](uint32_t code) { codes |= static_cast<unsigned long long>(code) << (4u * rows++); };
    row(kRowPos);
    row(kRowTgtRoad);
    row(kRowRng);
    row(src.spd ? kRowInitPermSpd : kRowInitPerm);
    if (live_outputs) {
        row(kRowDirNcount);
    }
    hipLaunchKernelGGL(k_reorder, dim3(nb, rows), dim3(kBlock), 0, stream, src, dst,
                       static_cast<const uint32_t*>(bid), bpos, codes);
    hipLaunchKernelGGL(k_bid_identity, dim3(std::max(1u, (n + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream, bid,
                       n);
}

// ---- Entity records (64 B, the layout of sim::Entity, Entity.hpp:33-46) ----
// @0 colour f32x4, @16 randomState u32x4, @32 pos, @40 target, @48 direction
// f32x2 each, @56 roadIndex, @60 initialized: four 16-byte words per record.

// AoS records -> the SoA state (internal order = caller order) and the
// records' colours.
static __global__ __launch_bounds__(kBlock) void k_unpack_entities(const uint4* __restrict__ rec, AgentsDev a,
                                                                   float4* __restrict__ color0) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) {
        return;
    }
    const uint4* r = rec + 4ull * i;
    const uint4 c = r[0], st = r[1], pt = r[2], dr = r[3];
    color0[i] = make_float4(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z), __uint_as_float(c.w));
    a.rng[i] = st;
    a.pos[i] = make_float2(__uint_as_float(pt.x), __uint_as_float(pt.y));
    a.tgt[i] = make_float2(__uint_as_float(pt.z), __uint_as_float(pt.w));
    a.dir[i] = make_float2(__uint_as_float(dr.x), __uint_as_float(dr.y));
    a.road[i] = dr.z;
    a.init[i] = dr.w ? 1u : 0u;
    a.ncount[i] = 0u;
    a.perm[i] = i;
}

// The SoA state -> records of caller indices [first, first + count), at
// out[index - first].  coloured: blue for agents in a pair in the last
// collision dispatch, green otherwise (random_move.comp:545-548, 876);
// otherwise the uploaded colours.
static __global__ __launch_bounds__(kBlock) void k_pack_entities(AgentsDev a, const float4* __restrict__ color0,
                                                                 uint32_t coloured, uint32_t first, uint32_t count,
                                                                 uint4* __restrict__ out) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= a.n) {
        return;
    }
    const uint32_t i = a.perm[k];
    const uint32_t j = i - first;
    if (j >= count) {
        return;
    }
    float4 c;
    if (coloured) {
        c = a.ncount[k] ? make_float4(0.0f, 0.0f, 1.0f, 1.0f) : make_float4(0.0f, 1.0f, 0.0f, 1.0f);
    } else {
        c = color0[i];
    }
    const float2 p = a.pos[k], t = a.tgt[k], d = a.dir[k];
    uint4* r = out + 4ull * j;
    r[0] = make_uint4(__float_as_uint(c.x), __float_as_uint(c.y), __float_as_uint(c.z), __float_as_uint(c.w));
    r[1] = a.rng[k];
    r[2] = make_uint4(__float_as_uint(p.x), __float_as_uint(p.y), __float_as_uint(t.x), __float_as_uint(t.y));
    r[3] = make_uint4(__float_as_uint(d.x), __float_as_uint(d.y), a.road[k], a.init[k]);
}

// Per-agent cell export in caller order.
static __global__ __launch_bounds__(kBlock) void k_export_cells(AgentsDev a, float inv_cell, int2* out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) {
        return;
    }
    const float2 p = a.pos[i];
    out[a.perm[i]] = make_int2(cell_coord(p.x, inv_cell), cell_coord(p.y, inv_cell));
}

// Occupied cells of the current binning, the grid's counterpart of the
// quadtree nodes the UI overlay draws (src/sim/GpuQuadTree.hpp:14-41 via
// Simulator::get_quad_tree_nodes, Simulator.cpp:149-153): one entry per
// distinct cell (cx, cy) holding agents, with its agent count.  A bucket can
// hold several world cells (the table is a torus), so each slot looks through
// its own bucket: the bucket's first slot of a cell writes the cell, counted
// over the bucket.  Entries are appended in no particular order; *total
// receives the number of occupied cells, entries past cap are not written.
static __global__ __launch_bounds__(kBlock) void k_occupancy(const float2* __restrict__ bpos,
                                                             const uint32_t* __restrict__ start, uint32_t n,
                                                             float inv_cell, uint32_t lw, uint32_t wmask,
                                                             uint32_t hmask, int2* __restrict__ out_cell,
                                                             uint32_t* __restrict__ out_count,
                                                             uint32_t* __restrict__ total, uint32_t cap) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) {
        return;
    }
    const float2 p = bpos[s];
    const int32_t cx = cell_coord(p.x, inv_cell), cy = cell_coord(p.y, inv_cell);
    const uint32_t b = bucket_of(cx, cy, lw, wmask, hmask);
    const uint32_t lo = start[b], hi = start[b + 1];
    uint32_t cnt = 0;
    for (uint32_t k = lo; k < hi; k++) {
        const float2 q = bpos[k];
        if (cell_coord(q.x, inv_cell) == cx && cell_coord(q.y, inv_cell) == cy) {
            if (k < s) {
                return;  // an earlier slot writes this cell
            }
            cnt++;
        }
    }
    const uint32_t idx = atomicAdd(total, 1u);
    if (idx < cap) {
        out_cell[idx] = make_int2(cx, cy);
        out_count[idx] = cnt;
    }
}

}  // namespace msim
