// msim_device.hpp — per-agent device functions of the movement-sim tick.
//
// Semantics follow src/sim/shader/random_move.comp line by line; the
// arithmetic is binary32 with one rounding per operation.  The library is
// compiled with -ffp-contract=off, and on gfx950 sqrtf and '/' lower to the
// correctly rounded v_sqrt/v_div_scale/v_div_fixup sequences, so every
// operation below matches an IEEE CPU evaluation of the same expression.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "msim/types.h"

namespace msim {

// random_move.comp:750  float SPEED = 1.4;
constexpr float kShaderSpeed = 1.4f;

// random_move.comp:725-736: xorshift128 step on the uvec4 (x, y, z, w).
__device__ __forceinline__ uint32_t xs128_next(uint4& s) {
    uint32_t t = s.w;
    const uint32_t x = s.x;
    s.w = s.z;
    s.z = s.y;
    s.y = x;
    t ^= t << 11;
    t ^= t >> 8;
    s.x = t ^ x ^ (x >> 19);
    return s.x;
}

// The value of x, opaque to the optimiser: no operation is folded across
// it, so a product passed through it stays a separately rounded product in
// every build.  With -ffp-contract=fast the backend fuses a*b+c into an fma
// whatever a scoped "fp contract(off)" says; arithmetic that decides integer
// state therefore passes its products through here.
__device__ __forceinline__ float opaque(float x) {
    asm("" : "+v"(x));
    return x;
}

// random_move.comp:738-746: offset in {lo..hi}; the product and the sum are
// two separate roundings (no fused multiply-add, in every build: the result
// decides integer state).
__device__ __forceinline__ uint32_t xs128_range(uint4& s, uint32_t lo, uint32_t hi) {
    const float f = static_cast<float>(xs128_next(s)) * 2.3283064365386963e-10f;
    const float prod = opaque(f * static_cast<float>(hi - lo + 1u));
    const float sum = static_cast<float>(lo) + prod;
    return static_cast<uint32_t>(ceilf(sum)) - 1u;
}

// GLSL length(v) = sqrt(x*x + y*y), one rounding per operation in every
// build (the move compares it with the speed: an integer-deciding branch).
__device__ __forceinline__ float vlen(float x, float y) {
    const float xx = opaque(x * x);
    const float yy = opaque(y * y);
    return sqrtf(xx + yy);
}

// random_move.comp:830-839  update_direction(index, pos).
__device__ __forceinline__ float2 heading(float2 tgt, float2 p, float speed) {
    const float dx = tgt.x - p.x;
    const float dy = tgt.y - p.y;
    const float len = vlen(dx, dy);
    if (len == 0.0f) {
        return make_float2(0.0f, 0.0f);
    }
    const float nx = dx / len;
    const float ny = dy / len;
    return make_float2(nx * speed, ny * speed);
}

struct Endpoint {
    float x, y;
    uint32_t idx, cnt;
};

__device__ __forceinline__ void load_road(const msim_road* roads, uint32_t r, Endpoint& a, Endpoint& b) {
    const uint4* p = reinterpret_cast<const uint4*>(roads + r);
    const uint4 u = p[0];
    const uint4 v = p[1];
    a = Endpoint{__uint_as_float(u.x), __uint_as_float(u.y), u.z, u.w};
    b = Endpoint{__uint_as_float(v.x), __uint_as_float(v.y), v.z, v.w};
}

// random_move.comp:778-828  new_target(index).  Every read is in bounds for a
// map that passed the host-side validation in msim_create.
__device__ __forceinline__ void choose_route(float2& tgt, uint32_t& road, uint4& rng,
                                             const msim_road* __restrict__ roads,
                                             const uint32_t* __restrict__ conns) {
    Endpoint s, e;
    load_road(roads, road, s, e);
    Endpoint c;
    if (tgt.x == s.x && tgt.y == s.y) {
        c = s;
        if (c.cnt <= 1u) {
            tgt = make_float2(e.x, e.y);
            return;
        }
    } else {
        c = e;
        if (c.cnt <= 1u) {
            tgt = make_float2(s.x, s.y);
            return;
        }
    }
    const uint32_t off = (c.cnt == 2u) ? 1u : xs128_range(rng, 1u, c.cnt);
    const uint32_t nr = conns[c.idx + off];
    Endpoint ns, ne;
    load_road(roads, nr, ns, ne);
    if (ns.x == tgt.x && ns.y == tgt.y) {
        tgt = make_float2(ne.x, ne.y);
    } else {
        tgt = make_float2(ns.x, ns.y);
    }
    road = nr;
}

// The endpoint choose_route calls c for an agent on road r heading to tgt:
// the start when tgt equals it, else the end.  Returned as (connection index,
// connection count).  Target and road change only in the route choice, so
// this value, taken when they are set, is the one the next route choice sees.
__device__ __forceinline__ uint2 route_endpoint(const Endpoint& s, const Endpoint& e, float2 tgt) {
    return (tgt.x == s.x && tgt.y == s.y) ? make_uint2(s.idx, s.cnt) : make_uint2(e.idx, e.cnt);
}

// choose_route for an agent standing on its target, with that endpoint's
// connection index and count (c, from route_endpoint) known: the same
// decisions and PRNG draws; the current road's record is read only at a dead
// end.  c receives the endpoint of the new road and target.
__device__ __forceinline__ void choose_route_at(float2& tgt, uint32_t& road, uint2& c, uint4& rng,
                                                const msim_road* __restrict__ roads,
                                                const uint32_t* __restrict__ conns) {
    if (c.y <= 1u) {  // dead end: back along the same road
        Endpoint s, e;
        load_road(roads, road, s, e);
        tgt = (tgt.x == s.x && tgt.y == s.y) ? make_float2(e.x, e.y) : make_float2(s.x, s.y);
        c = route_endpoint(s, e, tgt);
        return;
    }
    const uint32_t off = (c.y == 2u) ? 1u : xs128_range(rng, 1u, c.y);
    const uint32_t nr = conns[c.x + off];
    Endpoint ns, ne;
    load_road(roads, nr, ns, ne);
    if (ns.x == tgt.x && ns.y == tgt.y) {
        tgt = make_float2(ne.x, ne.y);
    } else {
        tgt = make_float2(ns.x, ns.y);
    }
    road = nr;
    c = route_endpoint(ns, ne, tgt);
}

// Uniform-grid cell coordinate: floor(v / cell) with cell a power of two, so
// v * inv_cell is exact; clamped to +-2^30 (NaN maps to the low clamp).
__device__ __forceinline__ int32_t cell_coord(float v, float inv_cell) {
    float c = floorf(v * inv_cell);
    c = (c >= -1073741824.0f) ? c : -1073741824.0f;
    c = (c <= 1073741823.0f) ? c : 1073741823.0f;
    return static_cast<int32_t>(c);
}

// Toroidal bucket of a cell: the world's cell grid folded onto a
// (wmask+1) x (hmask+1) table, row-major.
__device__ __forceinline__ uint32_t bucket_of(int32_t cx, int32_t cy, uint32_t lw, uint32_t wmask,
                                              uint32_t hmask) {
    return ((static_cast<uint32_t>(cy) & hmask) << lw) | (static_cast<uint32_t>(cx) & wmask);
}

// random_move.comp:551-562 quad_tree_in_range(pj, pi, R), exactly:
//   |dx| > R -> false; |dy| > R -> false; sqrt(dx*dx + dy*dy) < R.
// sqrtf is correctly rounded and non-decreasing, so sqrt(d2) < R <=>
// d2 < d2_limit (the smallest binary32 x with sqrtf(x) >= R).  The two box
// tests are implied: |dx| > R gives fl(dx*dx) >= fl(R*R), so d2 >= fl(R*R),
// and the host checks fl(R*R) >= d2_limit when it picks d2_limit.
__device__ __forceinline__ bool in_range(float2 pj, float2 pi, float d2_limit) {
    // decides neighbour counts: one rounding per operation in every build
    const float ddx = pj.x - pi.x;
    const float ddy = pj.y - pi.y;
    const float d2 = opaque(ddx * ddx) + opaque(ddy * ddy);
    return d2 < d2_limit;
}

// Workgroup index remapped so that each XCD (workgroups are dealt round-robin
// over the 8 XCDs) works on one contiguous slice of the index space, keeping
// neighbouring tiles in the same L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nblocks) {
    constexpr uint32_t kXcd = 8;
    const uint32_t xcd = bid % kXcd;
    const uint32_t local = bid / kXcd;
    const uint32_t per = nblocks / kXcd;
    const uint32_t rem = nblocks % kXcd;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

}  // namespace msim
