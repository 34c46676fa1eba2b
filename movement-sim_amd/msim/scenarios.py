"""Seeded synthetic maps and agents for the benchmark configurations.

Every generator emits the reference's own buffer formats: Road records with
per-endpoint connection groups (src/sim/Map.hpp:12-25, Map.cpp:130) and the
64-byte Entity (src/sim/Entity.hpp:33-46), so the same arrays drive the HIP
path and the CPU oracle.

Connection-list layouts (the shader reads connections[connectedIndex + k],
random_move.comp:803,809):
  "own_first" (default): one group per (road, endpoint), own road first, then
      the other roads at that node in ascending index order — the layout of
      the reference's test_map.json.
  "node_csr": one group per node shared by every road ending there, in
      ascending road index order — the per-node layout map/generate_map.py
      builds (one group per coordinate, generate_map.py:234-258).
Both append one guard entry (road 0): the draw at random_move.comp:808 can
return offset == connectedCount and read one entry past the group.

Road placement never yields zero-length roads (the reference loader would drop
them and shift road indices, Map.cpp:124-128).
"""
from dataclasses import dataclass, field

import numpy as np

from .types import ENTITY_DTYPE, ROAD_DTYPE, push_consts

MASTER_SEED = 0x6D53494D  # per-configuration seed = MASTER_SEED + config index

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """splitmix64 finaliser over a uint64 array (Steele, Lea, Flood 2014)."""
    with np.errstate(over="ignore"):
        z = (np.asarray(x, np.uint64) + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def agent_rng_states(seed, n, first_id=0):
    """Per-agent xorshift128 state (4 x uint32, never all zero) from a master seed."""
    ids = np.arange(first_id, first_id + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * np.uint64(0x100000001B3)
        w0 = splitmix64(base ^ (ids * np.uint64(2)))
        w1 = splitmix64(base ^ (ids * np.uint64(2) + np.uint64(1)))
    st = np.empty((n, 4), np.uint32)
    st[:, 0] = (w0 & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    st[:, 1] = (w0 >> np.uint64(32)).astype(np.uint32)
    st[:, 2] = (w1 & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    st[:, 3] = (w1 >> np.uint64(32)).astype(np.uint32)
    zero = ~st.any(axis=1)
    st[zero, 0] = 1
    return st


# ---- graphs --------------------------------------------------------------

def _spanning_tree(n, edges, weights):
    """Minimum spanning forest edge mask (scipy.sparse.csgraph)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import minimum_spanning_tree

    w = np.asarray(weights, np.float64) + 1e-9  # zero weights would vanish from the sparse matrix
    m = coo_matrix((w, (edges[:, 0], edges[:, 1])), shape=(n, n)).tocsr()
    t = minimum_spanning_tree(m).tocoo()
    r, c = t.row.astype(np.int64), t.col.astype(np.int64)
    key = np.concatenate([r * n + c, c * n + r])
    return np.isin(edges[:, 0].astype(np.int64) * n + edges[:, 1].astype(np.int64), key)


def lattice_graph(side=32, size=1000.0, jitter=0.3, drop=0.15, seed=0):
    """side x side jittered lattice on a size x size square, 4-neighbour edges,
    round(drop * E) edges removed at random among those outside a random
    spanning tree (so the graph stays connected)."""
    rng = np.random.default_rng(seed)
    sp = size / (side - 1)
    ii, jj = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    xy = np.stack([ii.ravel() * sp, jj.ravel() * sp], axis=1)
    xy += rng.uniform(-jitter * sp, jitter * sp, size=xy.shape)
    xy = np.clip(xy, 0.0, size).astype(np.float32)
    idx = ii * side + jj
    e = [np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], 1),
         np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], 1)]
    edges = np.concatenate(e).astype(np.int64)
    tree = _spanning_tree(side * side, edges, rng.random(edges.shape[0]))
    n_drop = int(round(drop * edges.shape[0]))
    cand = np.flatnonzero(~tree)
    n_drop = min(n_drop, cand.size)
    dropped = rng.choice(cand, n_drop, replace=False)
    keep = np.ones(edges.shape[0], bool)
    keep[dropped] = False
    return xy, edges[keep]


def city_graph(n_nodes=150_000, size=12_000.0, centres=8, sigma=1_500.0, mean_degree=2.6,
               min_len=5.0, max_len=400.0, max_degree=6, seed=0):
    """Random geometric planar road graph: nodes from a Gaussian mixture
    (clipped to the square), Delaunay triangulation, edges outside
    [min_len, max_len] removed, largest connected component kept, Euclidean
    spanning tree plus random extra Delaunay edges up to mean_degree with
    node degree capped at max_degree."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    c = rng.uniform(0.15 * size, 0.85 * size, size=(centres, 2))
    pts = np.empty((0, 2))
    while pts.shape[0] < n_nodes:
        k = n_nodes - pts.shape[0]
        cid = rng.integers(0, centres, k)
        p = c[cid] + rng.normal(0.0, sigma, size=(k, 2))
        p = p[(p >= 0).all(1) & (p <= size).all(1)]
        pts = np.concatenate([pts, p])
    pts = np.unique(pts.astype(np.float32), axis=0)
    pts = pts[rng.permutation(pts.shape[0])]
    tri = Delaunay(pts.astype(np.float64))
    s = tri.simplices
    e = np.concatenate([s[:, [0, 1]], s[:, [1, 2]], s[:, [2, 0]]])
    e = np.sort(e, axis=1)
    e = np.unique(e, axis=0)
    d = pts[e[:, 0]].astype(np.float64) - pts[e[:, 1]]
    length = np.hypot(d[:, 0], d[:, 1])
    ok = (length >= min_len) & (length <= max_len)
    e, length = e[ok], length[ok]
    n = pts.shape[0]
    m = coo_matrix((np.ones(e.shape[0]), (e[:, 0], e[:, 1])), shape=(n, n))
    ncomp, lab = connected_components(m, directed=False)
    big = np.argmax(np.bincount(lab))
    keep_node = lab == big
    remap = -np.ones(n, np.int64)
    remap[keep_node] = np.arange(keep_node.sum())
    sel = keep_node[e[:, 0]] & keep_node[e[:, 1]]
    e, length = remap[e[sel]], length[sel]
    pts = pts[keep_node]
    n = pts.shape[0]
    tree = _tree_by_length(n, e, length)
    deg = np.bincount(e[tree].ravel(), minlength=n)
    target = int(round(mean_degree * n / 2))
    chosen = tree.copy()
    need = target - int(tree.sum())
    if need > 0:
        cand = np.flatnonzero(~tree)
        cand = cand[rng.permutation(cand.size)]
        a_all, b_all = e[cand, 0], e[cand, 1]
        for k in range(cand.size):
            a, b = a_all[k], b_all[k]
            if deg[a] < max_degree and deg[b] < max_degree:
                deg[a] += 1
                deg[b] += 1
                chosen[cand[k]] = True
                need -= 1
                if need == 0:
                    break
    return pts.astype(np.float32), e[chosen]


def _tree_by_length(n, e, length):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import minimum_spanning_tree

    m = coo_matrix((length + 1e-9, (e[:, 0], e[:, 1])), shape=(n, n)).tocsr()
    t = minimum_spanning_tree(m).tocoo()
    a = np.minimum(t.row, t.col).astype(np.int64)
    b = np.maximum(t.row, t.col).astype(np.int64)
    tk = a * n + b
    ek = np.minimum(e[:, 0], e[:, 1]) * n + np.maximum(e[:, 0], e[:, 1])
    return np.isin(ek, tk)


def build_roads(nodes, edges, layout="own_first"):
    """Road records and the connection list for an undirected edge list.
    Road k runs from node edges[k, 0] (start) to edges[k, 1] (end)."""
    nodes = np.asarray(nodes, np.float32)
    edges = np.asarray(edges, np.int64)
    n_nodes = nodes.shape[0]
    n_roads = edges.shape[0]
    if n_roads and (nodes[edges[:, 0]] == nodes[edges[:, 1]]).all(1).any():
        raise ValueError("zero-length road")
    ends = edges.ravel()  # endpoint 2k = start of road k, 2k+1 = end
    road_of_end = np.repeat(np.arange(n_roads), 2)
    order = np.lexsort((road_of_end, ends))  # by node, then road index
    deg = np.bincount(ends, minlength=n_nodes)
    node_ptr = np.concatenate([[0], np.cumsum(deg)])
    node_roads = road_of_end[order]  # roads at each node, ascending
    roads = np.zeros(n_roads, ROAD_DTYPE)
    roads["start_pos"] = nodes[edges[:, 0]]
    roads["end_pos"] = nodes[edges[:, 1]]
    roads["start_count"] = deg[edges[:, 0]]
    roads["end_count"] = deg[edges[:, 1]]
    if layout == "node_csr":
        roads["start_index"] = node_ptr[edges[:, 0]]
        roads["end_index"] = node_ptr[edges[:, 1]]
        conns = np.concatenate([node_roads, [0]]).astype(np.uint32)
        return roads, conns
    if layout != "own_first":
        raise ValueError(layout)
    # One group per endpoint (endpoint 2k = start of road k, 2k+1 = its end):
    # the road itself, then the node's other roads in ascending order.
    gsize = deg[ends]
    gstart = np.concatenate([[0], np.cumsum(gsize)])[:-1]
    total = int(gsize.sum())
    endpoint = np.repeat(np.arange(2 * n_roads), gsize)  # endpoint owning each slot
    k = np.arange(total) - gstart[endpoint]  # slot within its group
    own = road_of_end[endpoint]
    node = ends[endpoint]
    r_own = _rank_in_node(order, ends, node_ptr, n_roads)[endpoint]
    j = k - 1
    j = np.maximum(j + (j >= r_own), 0)  # skip the road's own place in the node's list
    other = node_roads[node_ptr[node] + j] if total else np.zeros(0, np.int64)
    conns = np.empty(total + 1, np.uint32)
    conns[:total] = np.where(k == 0, own, other).astype(np.uint32)
    conns[total] = 0
    roads["start_index"] = gstart[0::2]
    roads["end_index"] = gstart[1::2]
    return roads, conns


def _rank_in_node(order, ends, node_ptr, n_roads):
    """Position of each endpoint's road within its node's ascending road list."""
    rank = np.empty(2 * n_roads, np.int64)
    sorted_nodes = ends[order]
    rank[order] = np.arange(2 * n_roads) - node_ptr[sorted_nodes]
    return rank


def star_map():
    """Four roads from the corners of a 100 m square to its centre: the
    structure of the reference's test_map.json (own-first groups; corner
    endpoints count 1, centre endpoints count 4), re-derived here; the centre
    groups list the other roads in ascending order (the fixture's own order
    differs for roads 2 and 3), plus a guard entry."""
    nodes = np.array([[50, 50], [5, 5], [95, 5], [5, 95], [95, 95]], np.float32)
    edges = np.array([[1, 0], [2, 0], [3, 0], [4, 0]])
    roads, conns = build_roads(nodes, edges, "own_first")
    return roads, conns, 100.0, 100.0


# ---- agents --------------------------------------------------------------

def place_agents(roads, n, seed, speed_range=(0.8, 1.6), occupancy="uniform", zipf_s=1.1, zipf_q=0.0,
                 uniform_speed=False):
    """n agents on roads: road by occupancy ("uniform" over road indices like
    Map::get_random_road_index, Map.cpp:152-157, or "zipf": weight
    (rank + q)^-s over a random permutation of the roads), position at a
    uniform fraction along the road, target a random endpoint, speed uniform in
    speed_range (float32), xorshift128 state from agent_rng_states."""
    rng = np.random.default_rng(seed)
    nr = roads.shape[0]
    if occupancy == "uniform":
        r = rng.integers(0, nr, n)
    elif occupancy == "zipf":
        perm = rng.permutation(nr)
        w = (np.arange(1, nr + 1, dtype=np.float64) + zipf_q) ** (-zipf_s)
        cdf = np.cumsum(w)
        cdf /= cdf[-1]
        rank = np.minimum(np.searchsorted(cdf, rng.random(n), side="right"), nr - 1)
        r = perm[rank]
    else:
        raise ValueError(occupancy)
    u = rng.random(n).astype(np.float32)
    s = roads["start_pos"][r]
    e = roads["end_pos"][r]
    pos = (s + (e - s) * u[:, None]).astype(np.float32)
    to_end = rng.random(n) < 0.5
    tgt = np.where(to_end[:, None], e, s).astype(np.float32)
    ents = np.zeros(n, ENTITY_DTYPE)
    ents["color"][:, :3] = rng.random((n, 3)).astype(np.float32)
    ents["color"][:, 3] = 1.0
    ents["randomState"] = agent_rng_states(seed, n)
    ents["pos"] = pos
    ents["target"] = tgt
    ents["roadIndex"] = r.astype(np.uint32)
    ents["initialized"] = 0
    if uniform_speed:
        speed = None
    else:
        lo, hi = speed_range
        speed = (lo + (hi - lo) * rng.random(n)).astype(np.float32)
    return ents, speed


def reference_agents(roads, n, seed):
    """Simulator::add_entities (Simulator.cpp:114-129) with a seeded generator:
    pos = road start, target = road end, direction 0, uniform road, random
    colour and state, initialized = 0; speed is the shader constant."""
    rng = np.random.default_rng(seed)
    r = rng.integers(0, roads.shape[0], n)
    ents = np.zeros(n, ENTITY_DTYPE)
    ents["color"][:, :3] = rng.random((n, 3)).astype(np.float32)
    ents["color"][:, 3] = 1.0
    ents["randomState"] = agent_rng_states(seed, n)
    ents["pos"] = roads["start_pos"][r]
    ents["target"] = roads["end_pos"][r]
    ents["roadIndex"] = r.astype(np.uint32)
    return ents


def max_cell_load(pos, cell=1.0):
    c = np.floor(np.asarray(pos, np.float64) / cell).astype(np.int64)
    key = (c[:, 1] << 32) + c[:, 0]
    _, cnt = np.unique(key, return_counts=True)
    return int(cnt.max()) if cnt.size else 0


# ---- benchmark configurations -----------------------------------------

@dataclass
class Scenario:
    name: str
    roads: np.ndarray
    conns: np.ndarray
    entities: np.ndarray
    speed: object
    pc: np.ndarray
    ticks: int
    world: tuple
    meta: dict = field(default_factory=dict)


def config(k, scale=1.0, seed=None, layout="own_first"):
    """Benchmark configuration k (1..3 plain, 5 adjoint base graph).
    scale < 1 shrinks node and agent counts proportionally (parity tests)."""
    seed = MASTER_SEED + k if seed is None else seed
    if k == 1:
        side = max(4, int(round(32 * np.sqrt(scale)))) if scale < 1 else 32
        nodes, edges = lattice_graph(side=side, size=1000.0, seed=seed)
        roads, conns = build_roads(nodes, edges, layout)
        n = max(1, int(round(10_000 * scale)))
        ents, speed = place_agents(roads, n, seed + 1)
        return Scenario("plain-small-lattice32", roads, conns, ents, speed,
                        push_consts(1000.0, 1000.0, 1.0), 1000, (1000.0, 1000.0),
                        {"nodes": nodes.shape[0], "config": 1})
    if k == 2:
        nn = max(1000, int(round(150_000 * scale)))
        size = 12_000.0 * np.sqrt(nn / 150_000)
        nodes, edges = city_graph(n_nodes=nn, size=size, sigma=1_500.0 * size / 12_000.0, seed=seed)
        roads, conns = build_roads(nodes, edges, layout)
        n = max(1, int(round(1_000_000 * scale)))
        ents, speed = place_agents(roads, n, seed + 1)
        return Scenario("city-150k-nodes-1M-agents", roads, conns, ents, speed,
                        push_consts(size, size, 1.0), 10_000, (size, size),
                        {"nodes": nodes.shape[0], "config": 2})
    if k == 3:
        nn = max(1000, int(round(1_000_000 * scale)))
        size = 31_000.0 * np.sqrt(nn / 1_000_000)
        nodes, edges = city_graph(n_nodes=nn, size=size, sigma=1_500.0 * size / 12_000.0, seed=seed)
        roads, conns = build_roads(nodes, edges, layout)
        n = max(1, int(round(16_000_000 * scale)))
        ents, speed = place_agents(roads, n, seed + 1, occupancy="zipf", zipf_s=1.1, zipf_q=ZIPF_Q)
        return Scenario("dense-1M-nodes-16M-agents-zipf", roads, conns, ents, speed,
                        push_consts(size, size, 1.0), 2_000, (size, size),
                        {"nodes": nodes.shape[0], "config": 3,
                         "max_cell_load": max_cell_load(ents["pos"])})
    raise ValueError(f"no plain configuration {k} (the batched workload is batch_config())")



def batch_config(n_scen=512, scale=1.0, seed=None, layout="own_first"):
    """The batched workload: n_scen independent lattice scenarios (the plain
    small construction at 32.3 m node spacing), node counts log-uniform in
    [1e3, 1e5] and agent counts log-uniform in [1e3, 5e5] (about 80k agents
    per scenario on average), each with its own seed.  scale < 1 shrinks both
    counts (parity tests)."""
    seed = MASTER_SEED + 4 if seed is None else seed
    rng = np.random.default_rng(seed)
    nodes = np.exp(rng.uniform(np.log(1e3), np.log(1e5), n_scen))
    agents = np.exp(rng.uniform(np.log(1e3), np.log(5e5), n_scen))
    out = []
    for s in range(n_scen):
        ss = int(splitmix64(np.uint64(seed) * np.uint64(1000003) + np.uint64(s)) & np.uint64(0x7FFFFFFF))
        side = max(4, int(round(np.sqrt(nodes[s] * scale))))
        size = 1000.0 * (side - 1) / 31.0
        nd, edges = lattice_graph(side=side, size=size, seed=ss)
        roads, conns = build_roads(nd, edges, layout)
        n = max(1, int(round(agents[s] * scale)))
        ents, speed = place_agents(roads, n, ss + 1)
        out.append(Scenario(f"batch-{s}", roads, conns, ents, speed, push_consts(size, size, 1.0), 5_000,
                            (size, size), {"nodes": nd.shape[0], "config": 4, "index": s}))
    return out


def pack_scenarios(scs):
    """Scenarios packed the way the batched handle packs them: roads,
    connections and agents concatenated, road indices (in connections and
    Entity.roadIndex) and connection indices rebased to the packed arrays.
    Returns (roads, conns, entities, speed or None, road_off, conn_off,
    agent_off)."""
    road_off = np.concatenate([[0], np.cumsum([sc.roads.shape[0] for sc in scs])]).astype(np.uint32)
    conn_off = np.concatenate([[0], np.cumsum([sc.conns.shape[0] for sc in scs])]).astype(np.uint32)
    agent_off = np.concatenate([[0], np.cumsum([sc.entities.shape[0] for sc in scs])]).astype(np.uint32)
    roads, conns, ents = [], [], []
    for k, sc in enumerate(scs):
        r = sc.roads.copy()
        r["start_index"] += conn_off[k]
        r["end_index"] += conn_off[k]
        roads.append(r)
        conns.append(np.asarray(sc.conns, np.uint32) + road_off[k])
        e = sc.entities.copy()
        e["roadIndex"] += road_off[k]
        ents.append(e)
    speed = None if scs[0].speed is None else np.concatenate([sc.speed for sc in scs]).astype(np.float32)
    return (np.concatenate(roads), np.concatenate(conns).astype(np.uint32), np.concatenate(ents), speed,
            road_off, conn_off, agent_off)


def adjoint_config(n_scen=64, agents=100_000, nodes=20_000, scale=1.0, seed=None, layout="own_first"):
    """The calibration workload: n_scen scenarios on one jittered-lattice map
    of about `nodes` nodes (the plain small construction at 32.3 m spacing),
    `agents` agents each with their own seed.  scale < 1 shrinks nodes and
    agents (tests)."""
    seed = MASTER_SEED + 5 if seed is None else seed
    side = max(4, int(round(np.sqrt(nodes * scale))))
    size = 1000.0 * (side - 1) / 31.0
    nd, edges = lattice_graph(side=side, size=size, seed=seed)
    roads, conns = build_roads(nd, edges, layout)
    n = max(1, int(round(agents * scale)))
    out = []
    for s in range(n_scen):
        ents, speed = place_agents(roads, n, seed + 1 + s)
        out.append(Scenario(f"adjoint-{s}", roads, conns, ents, speed, push_consts(size, size, 1.0), 1000,
                            (size, size), {"nodes": nd.shape[0], "config": 5, "index": s}))
    return out

# Zipf-Mandelbrot offset for the dense configuration: with q = 0 the heaviest
# of 1.3M roads would carry ~2M agents (~5e4 per 1 m cell, ~1e11 pair tests
# per tick); q = 1000 keeps the top road at a few thousand agents so cell
# loads of several hundred remain (asserted >= 200 by the generator's caller).
ZIPF_Q = 1000.0
