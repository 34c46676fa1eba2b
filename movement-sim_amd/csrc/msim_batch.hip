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
#include "msim_batch.hpp"

#include <cmath>
#include <cstring>

using namespace msim;

namespace msim {
thread_local std::string g_batch_error;
}  // namespace msim

namespace {

template <bool U, bool W>
void badvance(msim_batch* b) {
    const dim3 nb(std::max(1u, (b->n + kAdv * kBlock - 1) / (kAdv * kBlock)));
    hipLaunchKernelGGL((k_advance4<U, W>), nb, dim3(kBlock), 0, b->stream, b->agents(b->cur), b->inv_cell,
                       b->part_stats, b->roads, b->conns);
}

constexpr uint32_t kBSortedThreads = 512;

inline size_t bsorted_lds_bytes(uint32_t seg_log) { return sorted_lds_bytes(seg_log); }

SortedBufs bsorted_bufs(const msim_batch* b, int k) {
    const BBuffers& x = b->buf[k];
    return SortedBufs{x.pos, b->rec[k], b->nxt[k], x.dir};
}

constexpr uint32_t kBScanThreads = 1024;

template <uint32_t U, bool UN, bool W>
void blaunch_sorted(msim_batch* b, const SortedMove& mv, const BatchBin& bb) {
    bprobe_pre(b);
    hipLaunchKernelGGL(k_seg_scan<kBScanThreads>, dim3(1), dim3(kBScanThreads), 0, b->stream, b->seg_cnt[b->cur],
                       b->seg_start, b->n_seg, b->n, b->faults);
    bprobe_post(b);
    bprobe_pre(b);
    hipLaunchKernelGGL((k_bmove_sorted<U, kBSortedThreads, UN, W>), dim3(b->n_seg), dim3(kBSortedThreads),
                       bsorted_lds_bytes(b->max_seg_log), b->stream, mv, bb);
    bprobe_post(b);
    bprobe_pre(b);
    hipLaunchKernelGGL((k_broute_sorted<UN, W>), dim3(b->n_seg), dim3(kBlock), 0, b->stream, mv, bb);
    bprobe_post(b);
}

template <uint32_t U>
void blaunch_sorted_u(msim_batch* b, const SortedMove& mv, const BatchBin& bb, bool write_dir) {
    if (b->uniform) {
        if (write_dir) blaunch_sorted<U, true, true>(b, mv, bb);
        else blaunch_sorted<U, true, false>(b, mv, bb);
    } else {
        if (write_dir) blaunch_sorted<U, false, true>(b, mv, bb);
        else blaunch_sorted<U, false, false>(b, mv, bb);
    }
}

// Sorted move of every scenario: buf[cur] / binning bcur -> buf[1-cur] /
// binning 1-bcur in one launch.
int benqueue_move_sorted(msim_batch* b, bool write_dir) {
    b->launch_no = 0;
    const int nb = 1 - b->cur;
    SortedMove mv{};
    mv.src = bsorted_bufs(b, b->cur);
    mv.dst = bsorted_bufs(b, nb);
    mv.cold = b->cold;
    mv.roads = b->roads;
    mv.conns = b->conns;
    mv.spd_const = kShaderSpeed;
    mv.part_stats = b->part_stats;
    mv.faults = b->faults;
    mv.arr_rec = b->arr_rec;
    mv.arr_spd = b->arr_spd;
    mv.seg_start = b->seg_start;
    mv.seg_next = b->seg_cnt[nb];
    mv.seg_arr = b->seg_arr;
    mv.n = b->n;
    BatchBin bb{};
    bb.scen = b->scen;
    bb.seg_scen = b->seg_scen;
    bb.old_start = b->bstart[b->bcur];
    bb.new_start = b->bstart[1 - b->bcur];
    bb.ticket = b->bin_ticket;
    bb.status = b->seg_status;
    bb.faults = b->faults;
    bb.inv_cell = b->inv_cell;
    bb.n_seg = b->n_seg;
    if (b->bin_unroll == 2) blaunch_sorted_u<2>(b, mv, bb, write_dir);
    else if (b->bin_unroll == 8) blaunch_sorted_u<8>(b, mv, bb, write_dir);
    else blaunch_sorted_u<4>(b, mv, bb, write_dir);
    b->cur = nb;
    b->bcur = 1 - b->bcur;
    return bcheck(b, "batched sorted move dispatch");
}

int benqueue_move(msim_batch* b, bool write_dir) {
    if (b->sorted) return benqueue_move_sorted(b, write_dir);
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
    bb.ticket = b->bin_ticket;
    bb.status = b->seg_status;
    bb.faults = b->faults;
    bb.inv_cell = b->inv_cell;
    bb.n_seg = b->n_seg;
    const size_t lds = sizeof(uint32_t) << b->max_seg_log;
    const float2* pos = b->agents(b->cur).pos;
    blaunch_bins(b, pos, bb, lds);
    b->bcur = nb;
    return bcheck(b, "batched move dispatch");
}

int breorder(msim_batch* b) {
    const int nxt = 1 - b->cur;
    launch_reorder(b->agents(b->cur), b->agents(nxt), b->bbid[b->bcur], b->bbpos[b->bcur], b->n, b->stream,
                   !b->capturing);
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
    if (rc == MSIM_OK && !b->sorted && b->reorder_interval && b->since_reorder >= b->reorder_interval) {
        rc = breorder(b);
    }
    return rc;
}

// Enters the sorted state right after the creation's renumbering (buf[cur]
// in the order of binning bcur): road and PRNG state into caller order, the
// next bucket of every slot and the agent count of every segment.
int benter_sorted(msim_batch* b) {
    const dim3 g(bblocks(b->n));
    const BBuffers& x = b->buf[b->cur];
    hipLaunchKernelGGL(k_pack_rec, g, dim3(kBlock), 0, b->stream, x.tgt, b->uniform ? nullptr : x.spd, x.perm,
                       b->rec[b->cur], b->n);
    hipLaunchKernelGGL(k_cold_enter, g, dim3(kBlock), 0, b->stream, x.perm, x.road, x.rng, x.tgt, b->roads, b->cold,
                       b->n);
    const dim3 gs((b->n_seg + kBlock - 1) / kBlock);
    for (int k = 0; k < 2; k++) {
        hipLaunchKernelGGL(k_fill_u32, gs, dim3(kBlock), 0, b->stream, b->seg_cnt[k], 0u, b->n_seg);
    }
    if (b->uniform) {
        hipLaunchKernelGGL(k_bsorted_next<true>, g, dim3(kBlock), 0, b->stream, bsorted_bufs(b, b->cur), b->blk_scen,
                           b->scen, b->agent_off, b->n, kShaderSpeed, b->inv_cell, b->seg_cnt[b->cur]);
    } else {
        hipLaunchKernelGGL(k_bsorted_next<false>, g, dim3(kBlock), 0, b->stream, bsorted_bufs(b, b->cur), b->blk_scen,
                           b->scen, b->agent_off, b->n, kShaderSpeed, b->inv_cell, b->seg_cnt[b->cur]);
    }
    b->sorted = true;
    return bcheck(b, "sorted state set-up");
}


#define BTRY_HIP(call)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return bfail(b, MSIM_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// Captures seg_ticks ticks (whole reorder periods, each ending with its
// reorder) from the current buffers, right after a reorder.
int bbuild_graph(msim_batch* b) {
    hipGraph_t graph = nullptr;
    const int cur0 = b->cur, bcur0 = b->bcur;
    const uint32_t tick0 = b->tick, since0 = b->since_reorder;
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bfail(b, MSIM_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    int rc = MSIM_OK;
    b->capturing = true;  // run segments are always followed by the run's last tick
    for (uint32_t t = 0; t < b->seg_ticks && rc == MSIM_OK; t++) {
        rc = btick(b, false);
    }
    b->capturing = false;
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

namespace msim {

// Leaves the sorted state for the classic layout (the differentiable tick
// runs on it): road and PRNG state gathered into slot order, the binning's
// slot map reset to the identity and its positions copied.
int bleave_sorted(msim_batch* b) {
    if (!b->sorted) return MSIM_OK;
    const BBuffers& x = b->buf[b->cur];
    const dim3 g(bblocks(b->n));
    hipLaunchKernelGGL(k_unpack_rec, g, dim3(kBlock), 0, b->stream, b->rec[b->cur], x.tgt,
                       b->uniform ? nullptr : x.spd, x.perm, b->n);
    hipLaunchKernelGGL(k_cold_publish, g, dim3(kBlock), 0, b->stream, x.perm, b->cold, x.road, x.rng, b->n);
    hipLaunchKernelGGL(k_bid_identity, g, dim3(kBlock), 0, b->stream, b->bbid[b->bcur], b->n);
    int rc = bcheck(b, "sorted state hand-over");
    if (rc != MSIM_OK) return rc;
    if (b->n) {
        hipError_t e = hipMemcpyAsync(b->bbpos[b->bcur], x.pos, sizeof(float2) * b->n, hipMemcpyDeviceToDevice,
                                      b->stream);
        if (e != hipSuccess) return bfail(b, MSIM_ERR_HIP, "position copy: %s", hipGetErrorString(e));
    }
    b->sorted = false;
    for (auto& row : b->gexec)
        for (auto& gx : row)
            if (gx.exec) {
                (void)hipGraphExecDestroy(gx.exec);
                gx.exec = nullptr;
            }
    return MSIM_OK;
}

}  // namespace msim

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
    if (o.collide_mode > 2) {
        return bfail(nullptr, MSIM_ERR_INVALID, "collide_mode must be 0, 1 or 2");
    }
    if (o.binning != 0 && o.binning != 2 && o.binning != 3) {
        return bfail(nullptr, MSIM_ERR_INVALID,
                     "batched binning must be 0 (two passes), 2 (one pass) or 3 (sorted state)");
    }
    if (o.bin_unroll != 0 && o.bin_unroll != 4 && o.bin_unroll != 8 && !(o.binning == 3 && o.bin_unroll == 2)) {
        return bfail(nullptr, MSIM_ERR_INVALID, "bin_unroll must be 0, 4 or 8 (or 2 with binning 3)");
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
    b->radius = radius;
    b->first_tick_insert_only = (init0 == 0);
    b->tick = pc->tick;
    b->graph_ticks = o.graph_ticks ? o.graph_ticks : 100u;
    // the sorted state is in bucket order after every move: no renumbering
    b->reorder_interval = o.binning == 3 ? 0u : (opt ? o.reorder_interval : 20u);
    b->seg_ticks = segment_ticks(b->graph_ticks, b->reorder_interval);
    b->collide_mode = static_cast<int>(o.collide_mode);
    b->binning = static_cast<int>(o.binning);
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
        d.lw = o.table_width_log2 ? std::max(2u, std::min(o.table_width_log2, tlog - 2u))
                                  : auto_width_log2(tlog, rows_log);
        uint32_t seg_log = auto_segment_log2(tlog, d.lw);
        if (o.segment_log2) seg_log = std::min(std::max(o.segment_log2, 9u), tlog);
        auto lds_need = [&](uint32_t l) {
            return (o.binning == 3 ? sorted_lds_bytes(l) : (static_cast<size_t>(4) << l)) + 1024;
        };
        while (seg_log > 9 && lds_need(seg_log) > static_cast<size_t>(lds_max)) seg_log--;
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
    b->bin_unroll = o.bin_unroll ? o.bin_unroll : (o.binning == 3 ? 4u : auto_bin_unroll(b->max_seg_log));
    b->start_len = start_len;

    int rc = MSIM_OK;
    const size_t n_adv = (static_cast<size_t>(n) + kAdv - 1) / kAdv * kAdv;
#define BALLOC(p, cnt) \
    if (rc == MSIM_OK) rc = balloc_n(b, &(p), (cnt));
    BALLOC(b->roads, n_roads);
    BALLOC(b->conns, n_conns);
    for (int k = 0; k < 2; k++) {
        BBuffers& bb = b->buf[k];
        // k_advance4 moves kAdv agents per thread; the neighbour pass's
        // 16-byte staged reads of the sorted state may touch slots n, n + 1
        BALLOC(bb.pos, n_adv + 4);
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
    if (b->binning == 3) {
        BALLOC(b->nxt[0], n);
        BALLOC(b->nxt[1], n);
        BALLOC(b->rec[0], n);
        BALLOC(b->rec[1], n);
        BALLOC(b->cold, 2 * static_cast<size_t>(n));
        BALLOC(b->arr_rec, n);
        if (!b->uniform) BALLOC(b->arr_spd, n);
    }
    BALLOC(b->scen, n_scen);
    BALLOC(b->agent_off, n_scen + 1);
    BALLOC(b->blk_scen, bblocks(n));
    BALLOC(b->seg_scen, b->n_seg);
    BALLOC(b->seg_total, b->n_seg);
    BALLOC(b->bin_ticket, 1);
    BALLOC(b->seg_status, b->n_seg);
    if (b->binning == 3) {
        BALLOC(b->seg_start, b->n_seg + 1);
        BALLOC(b->seg_cnt[0], b->n_seg);
        BALLOC(b->seg_cnt[1], b->n_seg);
        BALLOC(b->seg_arr, b->n_seg);
    }
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
        if (e == hipSuccess) e = hipMemset(b->bin_ticket, 0, sizeof(unsigned long long));
        if (e == hipSuccess) e = hipMemset(b->seg_status, 0, sizeof(unsigned long long) * b->n_seg);
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
    if (rc == MSIM_OK && b->binning == 3) rc = benter_sorted(b);
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
    if (b->diff) {
        return bfail(b, MSIM_ERR_STATE, "this handle runs differentiable ticks (msim_batch_diff_setup); plain ticks "
                                        "need a handle of their own");
    }
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
    if (f) {
        return bfail(b, MSIM_ERR_STATE, "device-side check failed (fault bits 0x%x%s%s%s)", f,
                     (f & kFaultBinCount) ? ": binned agent count != agent count" : "",
                     (f & kFaultLookback) ? ": rebinning look-back timed out" : "",
                     (f & kFaultNext) ? ": an agent's position left its counted bucket" : "");
    }
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
    dn(d.data(), bb.dir + off, sizeof(float2) * na);
    // sorted state: target and caller index in the slots' records, road and
    // PRNG state in caller order
    if (b->sorted) {
        std::vector<uint4> rc4(na);
        dn(rc4.data(), b->rec[b->cur] + off, sizeof(uint4) * na);
        for (uint32_t k = 0; k < na; k++) {
            std::memcpy(&t[k], &rc4[k], sizeof(float2));  // the record's first 8 bytes: the target
            pm[k] = rc4[k].w;
        }
        std::vector<uint4> cw(2 * static_cast<size_t>(na));
        dn(cw.data(), b->cold + 2 * static_cast<size_t>(off), sizeof(uint4) * 2 * na);
        for (uint32_t k = 0; k < na; k++) {
            r[k] = cw[2 * k].x;
            g[k] = cw[2 * k + 1];
        }
    } else {
        dn(t.data(), bb.tgt + off, sizeof(float2) * na);
        dn(pm.data(), bb.perm + off, sizeof(uint32_t) * na);
        dn(r.data(), bb.road + off, sizeof(uint32_t) * na);
        dn(g.data(), bb.rng + off, sizeof(uint4) * na);
    }
    dn(c.data(), bb.ncount + off, sizeof(uint32_t) * na);
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
        const uint32_t c_i = b->sorted ? i : k;  // caller-order arrays are indexed by agent
        if (road_index) road_index[i] = r[c_i] - road_base;
        if (rand_state) {
            rand_state[4 * i] = g[c_i].x;
            rand_state[4 * i + 1] = g[c_i].y;
            rand_state[4 * i + 2] = g[c_i].z;
            rand_state[4 * i + 3] = g[c_i].w;
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
    if (b->diff) return bfail(b, MSIM_ERR_STATE, "this handle runs differentiable ticks");
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
        const int nk = batch_kernel_count(b);
        double acc[kBatchKernelMax] = {0, 0, 0, 0, 0};
        for (uint32_t t = t0; t < n_ticks && rc == MSIM_OK; t++) {
            const int k = static_cast<int>((t - t0) % static_cast<uint32_t>(nk));
            b->probe_kernel = k;
            b->probe_ev[0] = ev[0];
            b->probe_ev[1] = ev[1];
            rc = btick(b, t + 1 == n_ticks);
            b->probe_kernel = -1;
            if (rc == MSIM_OK) rc = bwait(b);
            float ms = 0.f;
            if (rc == MSIM_OK && hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) {
                acc[k] += static_cast<double>(ms) * nk;
            }
        }
        for (uint32_t k = 0; k < n_kernel_ms && k < static_cast<uint32_t>(nk); k++) {
            kernel_ms[k] = static_cast<float>(acc[k]);
        }
    }
    (void)hipGetLastError();
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return rc;
}

extern "C" const char* msim_batch_kernel_name(const msim_batch* b, uint32_t k) {
    if (!b || k >= static_cast<uint32_t>(batch_kernel_count(b))) return nullptr;
    return batch_kernel_name(b, static_cast<int>(k));
}

extern "C" int msim_batch_destroy(msim_batch* b) {
    if (!b) return MSIM_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->diff) diff_release(b);
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
