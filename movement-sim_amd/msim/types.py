"""numpy dtypes with the byte layouts of include/msim/types.h.

Each dtype matches the reference's buffer format:
  ENTITY_DTYPE  <-> sim::Entity (src/sim/Entity.hpp:33-46), 64 B
  ROAD_DTYPE    <-> sim::Road   (src/sim/Map.hpp:12-25), 32 B
  PUSH_DTYPE    <-> sim::PushConsts (src/sim/PushConsts.hpp:9-20), 28 B
"""
import numpy as np

ENTITY_DTYPE = np.dtype(
    [
        ("color", "<f4", (4,)),
        ("randomState", "<u4", (4,)),
        ("pos", "<f4", (2,)),
        ("target", "<f4", (2,)),
        ("direction", "<f4", (2,)),
        ("roadIndex", "<u4"),
        ("initialized", "<u4"),
    ]
)

ROAD_DTYPE = np.dtype(
    [
        ("start_pos", "<f4", (2,)),
        ("start_index", "<u4"),
        ("start_count", "<u4"),
        ("end_pos", "<f4", (2,)),
        ("end_index", "<u4"),
        ("end_count", "<u4"),
    ]
)

PUSH_DTYPE = np.dtype(
    [
        ("worldSizeX", "<f4"),
        ("worldSizeY", "<f4"),
        ("nodeCount", "<u4"),
        ("maxDepth", "<u4"),
        ("entityNodeCap", "<u4"),
        ("collisionRadius", "<f4"),
        ("tick", "<u4"),
    ]
)

assert ENTITY_DTYPE.itemsize == 64
assert ENTITY_DTYPE.fields["pos"][1] == 32
assert ENTITY_DTYPE.fields["roadIndex"][1] == 56
assert ROAD_DTYPE.itemsize == 32
assert PUSH_DTYPE.itemsize == 28

# The reference's quadtree constants (src/sim/Simulator.hpp:37-38, Simulator.cpp:94-101);
# kept in PushConsts for layout, not read by the grid implementation.
QUADTREE_NODE_COUNT = 21845
QUADTREE_MAX_DEPTH = 8
QUADTREE_ENTITY_NODE_CAP = 10

GREEN = np.array([0.0, 1.0, 0.0, 1.0], np.float32)
BLUE = np.array([0.0, 0.0, 1.0, 1.0], np.float32)


def push_consts(world_x, world_y, radius, tick=1):
    pc = np.zeros((), PUSH_DTYPE)
    pc["worldSizeX"] = world_x
    pc["worldSizeY"] = world_y
    pc["nodeCount"] = QUADTREE_NODE_COUNT
    pc["maxDepth"] = QUADTREE_MAX_DEPTH
    pc["entityNodeCap"] = QUADTREE_ENTITY_NODE_CAP
    pc["collisionRadius"] = radius
    pc["tick"] = tick
    return pc
