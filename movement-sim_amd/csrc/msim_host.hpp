// msim_host.hpp — host-side helpers shared by the C-ABI entry points:
// map/agent validation, grid sizing and the exact squared-radius threshold.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "msim/types.h"

namespace msim {

inline uint32_t next_pow2_u32(uint32_t v) {
    if (v <= 1) return 1;
    v--;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    return v + 1;
}

inline uint32_t ceil_log2_u64(uint64_t v) {
    uint32_t l = 0;
    while ((1ull << l) < v) l++;
    return l;
}

// Bucket table size: about one bucket per agent, at least 2^12 (one scan
// tile) and at most 2^26.
inline uint32_t auto_table_log2(uint32_t n_agents) {
    uint32_t l = ceil_log2_u64(n_agents ? n_agents : 1);
    if (l < 12) l = 12;
    if (l > 26) l = 26;
    return l;
}

// Table rows: the table must fold the world vertically (height <= the
// world's rows in cells), otherwise rows the world never reaches stay empty
// and the used buckets alias more.  Beyond that, rows are kept narrow so the
// rebinning window (a segment plus a few rows each side) stays small.
inline uint32_t auto_width_log2(uint32_t tlog, uint32_t rows_log) {
    uint32_t fold = rows_log + 2 > tlog ? 2 : tlog - rows_log;  // height <= 2^rows_log
    uint32_t l = fold < 5 ? 5 : fold;
    if (l + 2 > tlog) l = tlog - 2;
    return l;
}

// Segment owned by one rebinning workgroup: 16 table rows, 512..16384
// buckets (at most 64 KB of LDS counters).
inline uint32_t auto_segment_log2(uint32_t tlog, uint32_t lw) {
    uint32_t l = lw + 4;
    if (l < 9) l = 9;
    if (l > 14) l = 14;
    if (l > tlog) l = tlog;
    return l;
}

// Window slots each rebinning thread loads per round: 8 keeps more loads in
// flight; segments of 2^14 buckets (64 KB of LDS counters, two workgroups
// per CU) measured faster with 4.
inline uint32_t auto_bin_unroll(uint32_t seg_log) { return seg_log <= 13 ? 8u : 4u; }

// floor(log2(rows spanned by the roads and agents, in cells)), at least 2.
inline uint32_t extent_rows_log2(const msim_road* roads, uint32_t n_roads, const msim_entity* e, uint32_t n,
                                 float cell) {
    double lo = 0, hi = 0;
    bool any = false;
    auto add = [&](float y) {
        if (!any) {
            lo = hi = y;
            any = true;
        }
        lo = y < lo ? y : lo;
        hi = y > hi ? y : hi;
    };
    for (uint32_t r = 0; r < n_roads; r++) {
        add(roads[r].start.pos.y);
        add(roads[r].end.pos.y);
    }
    for (uint32_t i = 0; i < n; i++) {
        add(e[i].pos.y);
    }
    const double rows = (hi - lo) / cell + 1.0;
    uint32_t l = 2;
    while (l < 30 && static_cast<double>(1ull << (l + 1)) <= rows) l++;
    return l;
}

// Ticks per captured segment: whole reorder periods (at least one), about
// graph_ticks long; graph_ticks itself when reordering is off.
inline uint32_t segment_ticks(uint32_t graph_ticks, uint32_t reorder_interval) {
    if (reorder_interval == 0) return graph_ticks;
    uint32_t k = (graph_ticks + reorder_interval / 2) / reorder_interval;
    if (k == 0) k = 1;
    return k * reorder_interval;
}

inline bool is_pow2_float(float v) {
    if (!(v > 0.0f) || !std::isfinite(v)) return false;
    int e = 0;
    return std::frexp(v, &e) == 0.5f;
}

// Smallest power of two >= r: with such a cell edge, floor(x / cell) is exact
// and any two points closer than r lie in the same or adjacent cells.
inline float pow2_cell(float r) {
    int e = 0;
    const float m = std::frexp(r, &e);
    if (m == 0.5f) return r;
    return std::ldexp(1.0f, e);
}

inline float bits_to_float(uint32_t b) {
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// Smallest binary32 x >= 0 with sqrtf(x) >= r.  sqrtf is correctly rounded
// and non-decreasing, so sqrtf(d2) < r  <=>  d2 < d2_limit_for(r).
inline float d2_limit_for(float r) {
    uint32_t lo = 0, hi = 0x7f800000u;  // +inf
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        volatile float x = bits_to_float(mid);
        if (std::sqrt(static_cast<float>(x)) >= r) {
            hi = mid;
        } else {
            lo = mid + 1;
        }
    }
    return bits_to_float(lo);
}

// The neighbour predicate drops quad_tree_in_range's box tests, which is exact
// when every |dx| > r already fails d2 < limit, i.e. when fl(r*r) >= limit.
inline bool box_tests_implied(float r, float limit) {
    volatile float rr = r;
    const float sq = rr * rr;
    return sq >= limit;
}

// Every read random_move.comp:778-828 can make must stay inside the arrays:
// an endpoint with count >= 3 may read connections[index + count] (the draw
// can return the top offset), count == 2 reads connections[index + 1].
inline bool validate_map(const msim_road* roads, uint32_t n_roads, const uint32_t* conns,
                         uint32_t n_conns, std::string* why) {
    char buf[256];
    for (uint32_t r = 0; r < n_roads; r++) {
        const msim_coordinate* ends[2] = {&roads[r].start, &roads[r].end};
        for (int k = 0; k < 2; k++) {
            const msim_coordinate& c = *ends[k];
            if (!std::isfinite(c.pos.x) || !std::isfinite(c.pos.y)) {
                snprintf(buf, sizeof(buf), "road %u has a non-finite endpoint", r);
                *why = buf;
                return false;
            }
            if (c.connectedCount >= (1u << 24)) {
                // float(count) must be exact for the draw to stay <= count
                snprintf(buf, sizeof(buf), "road %u %s group count %u exceeds 2^24", r,
                         k ? "end" : "start", c.connectedCount);
                *why = buf;
                return false;
            }
            uint64_t top = 0;
            if (c.connectedCount == 2) {
                top = static_cast<uint64_t>(c.connectedIndex) + 1;
            } else if (c.connectedCount >= 3) {
                top = static_cast<uint64_t>(c.connectedIndex) + c.connectedCount;
            } else {
                continue;
            }
            if (top >= n_conns) {
                snprintf(buf, sizeof(buf),
                         "road %u %s group (index %u, count %u) can read connections[%llu] but only "
                         "%u entries exist; append a guard entry",
                         r, k ? "end" : "start", c.connectedIndex, c.connectedCount,
                         static_cast<unsigned long long>(top), n_conns);
                *why = buf;
                return false;
            }
        }
    }
    for (uint32_t k = 0; k < n_conns; k++) {
        if (conns[k] >= n_roads) {
            snprintf(buf, sizeof(buf), "connections[%u] = %u is not a road index (< %u)", k, conns[k],
                     n_roads);
            *why = buf;
            return false;
        }
    }
    return true;
}

// Cell coordinates must stay far from the +-2^30 clamp of the cell index so
// that every tick's displacement bound holds in cell units.
inline bool validate_extent(const msim_road* roads, uint32_t n_roads, const msim_entity* e, uint32_t n,
                            float inv_cell, std::string* why) {
    const double lim = 536870912.0;  // 2^29 cells
    auto bad = [&](float v) { return !(std::fabs(static_cast<double>(v) * inv_cell) < lim); };
    char buf[160];
    for (uint32_t r = 0; r < n_roads; r++) {
        if (bad(roads[r].start.pos.x) || bad(roads[r].start.pos.y) || bad(roads[r].end.pos.x) ||
            bad(roads[r].end.pos.y)) {
            snprintf(buf, sizeof(buf), "road %u lies beyond 2^29 grid cells from the origin", r);
            *why = buf;
            return false;
        }
    }
    for (uint32_t i = 0; i < n; i++) {
        if (bad(e[i].pos.x) || bad(e[i].pos.y) || bad(e[i].target.x) || bad(e[i].target.y)) {
            snprintf(buf, sizeof(buf), "agent %u lies beyond 2^29 grid cells from the origin", i);
            *why = buf;
            return false;
        }
    }
    return true;
}

inline bool validate_entities(const msim_entity* e, uint32_t n, uint32_t n_roads, const float* speed,
                              std::string* why) {
    char buf[256];
    for (uint32_t i = 0; i < n; i++) {
        if (e[i].roadIndex >= n_roads) {
            snprintf(buf, sizeof(buf), "agent %u has roadIndex %u >= %u roads", i, e[i].roadIndex, n_roads);
            *why = buf;
            return false;
        }
        if (!std::isfinite(e[i].pos.x) || !std::isfinite(e[i].pos.y) || !std::isfinite(e[i].target.x) ||
            !std::isfinite(e[i].target.y)) {
            snprintf(buf, sizeof(buf), "agent %u has a non-finite position or target", i);
            *why = buf;
            return false;
        }
        if (speed && !(speed[i] >= 0.0f && std::isfinite(speed[i]))) {
            snprintf(buf, sizeof(buf), "agent %u has an invalid speed", i);
            *why = buf;
            return false;
        }
    }
    return true;
}

}  // namespace msim
