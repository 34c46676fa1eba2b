/*
 * msim/types.h — plain-old-data layouts shared by the C-ABI, the host driver
 * and every caller.  Each struct is byte-identical to the buffer format the
 * movement-sim compute shader binds, so a caller that already fills the
 * reference's buffers can hand them over unchanged.
 *
 *   msim_entity      <-> sim::Entity            src/sim/Entity.hpp:33-46
 *                        EntityDescriptor       src/sim/shader/random_move.comp:5-13
 *   msim_coordinate  <-> sim::Coordinate        src/sim/Map.hpp:12-20
 *                        CoordinateDescriptor   src/sim/shader/random_move.comp:15-19
 *   msim_road        <-> sim::Road              src/sim/Map.hpp:22-25
 *                        RoadDescriptor         src/sim/shader/random_move.comp:21-24
 *   msim_push_consts <-> sim::PushConsts        src/sim/PushConsts.hpp:9-20
 *                        PushConstants          src/sim/shader/random_move.comp:26-37
 *
 * Field names follow the reference so that code written against sim::Entity
 * and friends reads the same.
 */
#ifndef MSIM_TYPES_H
#define MSIM_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct msim_vec2 {
    float x;
    float y;
} msim_vec2;

typedef struct msim_vec4u {
    uint32_t x;
    uint32_t y;
    uint32_t z;
    uint32_t w;
} msim_vec4u;

typedef struct msim_rgba {
    float r;
    float g;
    float b;
    float a;
} msim_rgba;

/* 64-byte agent record (std430 layout of EntityDescriptor). */
typedef struct msim_entity {
    msim_rgba color;        /* @0  */
    msim_vec4u randomState; /* @16 xorshift128 state, shader element order */
    msim_vec2 pos;          /* @32 */
    msim_vec2 target;       /* @40 */
    msim_vec2 direction;    /* @48 */
    uint32_t roadIndex;     /* @56 */
    uint32_t initialized;   /* @60 */
} msim_entity;

/* Road endpoint: position plus its group in the connection list. */
typedef struct msim_coordinate {
    msim_vec2 pos;
    uint32_t connectedIndex;
    uint32_t connectedCount;
} msim_coordinate;

typedef struct msim_road {
    msim_coordinate start;
    msim_coordinate end;
} msim_road;

/* 28-byte per-dispatch constants.  Only worldSizeX/Y, collisionRadius and
 * tick are read by this implementation; nodeCount, maxDepth and
 * entityNodeCap describe the reference's quadtree and are kept for layout. */
typedef struct msim_push_consts {
    float worldSizeX;
    float worldSizeY;
    uint32_t nodeCount;
    uint32_t maxDepth;
    uint32_t entityNodeCap;
    float collisionRadius;
    uint32_t tick;
} msim_push_consts;

#ifdef __cplusplus
#define MSIM_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define MSIM_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

MSIM_STATIC_ASSERT(sizeof(msim_vec2) == 8, "vec2 is 8 B");
MSIM_STATIC_ASSERT(sizeof(msim_entity) == 64, "Entity is 64 B");
MSIM_STATIC_ASSERT(offsetof(msim_entity, color) == 0, "color @0");
MSIM_STATIC_ASSERT(offsetof(msim_entity, randomState) == 16, "randState @16");
MSIM_STATIC_ASSERT(offsetof(msim_entity, pos) == 32, "pos @32");
MSIM_STATIC_ASSERT(offsetof(msim_entity, target) == 40, "target @40");
MSIM_STATIC_ASSERT(offsetof(msim_entity, direction) == 48, "direction @48");
MSIM_STATIC_ASSERT(offsetof(msim_entity, roadIndex) == 56, "roadIndex @56");
MSIM_STATIC_ASSERT(offsetof(msim_entity, initialized) == 60, "initialized @60");
MSIM_STATIC_ASSERT(sizeof(msim_coordinate) == 16, "Coordinate is 16 B");
MSIM_STATIC_ASSERT(sizeof(msim_road) == 32, "Road is 32 B");
MSIM_STATIC_ASSERT(offsetof(msim_road, end) == 16, "Road.end @16");
MSIM_STATIC_ASSERT(sizeof(msim_push_consts) == 28, "PushConsts is 28 B");
MSIM_STATIC_ASSERT(offsetof(msim_push_consts, collisionRadius) == 20, "collisionRadius @20");
MSIM_STATIC_ASSERT(offsetof(msim_push_consts, tick) == 24, "tick @24");

#ifdef __cplusplus
}
#endif

#endif /* MSIM_TYPES_H */
