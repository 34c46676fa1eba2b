"""The rounding edges of the route draw and the snap test, forced on the GPU.

random_move.comp:738-746 turns a xorshift output u into an offset
ceil(1 + f * cnt) - 1 with f = float(u) * 2^-32, two roundings: u = 2^32 - 129
(f = 0.99999994), u >= 2^32 - 128 (f = 1.0, the top offset cnt, Q1), u = 0
(offset 0), and the u where 1 + f * cnt crosses an integer are where a fused
multiply-add or a wrong rounding would change the offset.  random_move.comp
:841-852 snaps an agent when |target - pos| <= speed, so dist == speed and one
ulp either side decide between a step and an arrival.  The random draws of
the other tests reach these values about once in 10^7 draws; here every
agent is placed on one of them, and one move dispatch through the parity
build and the contraction-on build (both rebinning layouts) must give the
oracle's road, PRNG state and target bit for bit (positions bit for bit in
the parity build, within 2 ulp with contraction on).
"""
import numpy as np
import pytest

import msim
import orc
from msim import scenarios as S

pytestmark = pytest.mark.gpu

COUNTS = (3, 4, 5, 6, 7, 8, 100)


def state_emitting(u):
    """xorshift128 state (shader order) whose next output is u (w = 0: the
    output is x ^ (x >> 19), inverted by x = u ^ (u >> 19))."""
    return np.array([u ^ (u >> 19), 7, 11, 0], np.uint32)


def offset(u, cnt):
    return orc.rng_range(state_emitting(int(u)), 1, cnt)[0]


def boundary_draws(cnt):
    """u values at every place where the drawn offset changes (the last u of
    each offset and the first of the next), plus the fixed edges."""
    us = {0, 1, 2**31, 4294967167, 4294967168, 4294967295}
    for k in range(1, cnt + 1):
        c = (k * 2**32) // cnt
        lo, hi = max(0, c - 4096), min(2**32 - 1, c + 4096)
        if offset(lo, cnt) == offset(hi, cnt):
            continue
        while hi - lo > 1:  # the offset is non-decreasing in u
            mid = (lo + hi) // 2
            if offset(mid, cnt) == offset(lo, cnt):
                lo = mid
            else:
                hi = mid
        us.update({lo, hi})
    return sorted(us)


def hub_map():
    """One hub per count: cnt roads from leaves (count 1) into the hub, one
    connection group per hub holding all of them, a guard entry at the end.
    Plus one long road for the snap-threshold agents."""
    roads, conns = [], []
    hubs = []
    for h, cnt in enumerate(COUNTS):
        cx, cy = 200.0 + 300.0 * h, 200.0
        g_hub = len(conns)
        first = len(roads)
        conns.extend(range(first, first + cnt))
        for j in range(cnt):
            a = 2 * np.pi * j / cnt
            lx, ly = cx + 20.0 * np.cos(a), cy + 20.0 * np.sin(a)
            roads.append(((lx, ly), (cx, cy), g_hub, cnt))
        hubs.append((cnt, first, (np.float32(cx), np.float32(cy))))
    long_road = len(roads)
    roads.append(((100.0, 1000.0), (2100.0, 1000.0), -1, 1))
    rec = np.zeros(len(roads), msim.ROAD_DTYPE)
    for r, (s, e, g, cnt) in enumerate(roads):
        leaf_group = len(conns)
        conns.append(r)
        rec[r]["start_pos"] = s
        rec[r]["start_index"] = leaf_group
        rec[r]["start_count"] = 1
        rec[r]["end_pos"] = e
        rec[r]["end_index"] = g if g >= 0 else leaf_group
        rec[r]["end_count"] = cnt
    conns.append(0)  # guard: the top offset of the last group reads one past it
    return rec, np.array(conns, np.uint32), hubs, long_road


def boundary_agents():
    roads, conns, hubs, long_road = hub_map()
    ents, spd = [], []
    for cnt, first, (hx, hy) in hubs:
        for i, u in enumerate(boundary_draws(cnt)):
            e = np.zeros((), msim.ENTITY_DTYPE)
            e["pos"] = (hx, hy)
            e["target"] = (hx, hy)  # standing on the target: arrives, draws
            e["roadIndex"] = first + i % cnt
            e["randomState"] = state_emitting(u)
            e["initialized"] = 1
            ents.append(e)
            spd.append(np.float32(1.4))
    # snap threshold: speed == |target - pos| as the shader computes it, and
    # one ulp either side
    rng = np.random.default_rng(11)
    tgt = np.array([2100.0, 1000.0], np.float32)
    for k in range(300):
        p = np.array([2100.0 - rng.uniform(0.5, 1.6), 1000.0 + rng.uniform(-0.7, 0.7)], np.float32)
        d = tgt - p
        dist = np.sqrt(np.float32(d[0] * d[0]) + np.float32(d[1] * d[1]), dtype=np.float32)
        for s in (dist, np.nextafter(dist, np.float32(0)), np.nextafter(dist, np.float32(9))):
            e = np.zeros((), msim.ENTITY_DTYPE)
            e["pos"] = p
            e["target"] = tgt
            e["roadIndex"] = long_road
            e["randomState"] = state_emitting(int(rng.integers(0, 2**32)))
            e["initialized"] = 1
            ents.append(e)
            spd.append(np.float32(s))
    return roads, conns, np.array(ents, msim.ENTITY_DTYPE), np.array(spd, np.float32)


def ulps(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


def test_boundary_draws_cover_every_flip():
    for cnt in COUNTS:
        us = boundary_draws(cnt)
        offs = {offset(u, cnt) for u in us}
        assert offs == set(range(0, cnt + 1)), cnt  # offset 0 (u = 0) up to the top offset cnt
    assert offset(4294967168, 5) == 5 and offset(4294967167, 5) == 5 and offset(0, 5) == 0


@pytest.mark.parametrize("build", ["parity", "contract"])
@pytest.mark.parametrize("binning", [0, 3])
def test_one_move_at_the_rounding_edges(build, binning):
    roads, conns, ents, spd = boundary_agents()
    lib = msim.load_library(msim.LIB_CONTRACT_PATH if build == "contract" else msim.LIB_PATH)
    pc = msim.push_consts(3000.0, 2000.0, 1.0, tick=3)  # next dispatch (tick 4) moves
    g = msim.Simulation(roads, conns, ents, pc, speed=spd, library=lib, binning=binning)
    o = orc.OracleSim(roads, conns, ents, pc, speed=spd)
    g.dispatch(4)
    o.dispatch(4)
    st = g.state()
    np.testing.assert_array_equal(st["road"], o.st.road)
    np.testing.assert_array_equal(st["rng"], o.st.rng)
    np.testing.assert_array_equal(st["tgt"].view(np.uint32), o.st.tgt.view(np.uint32))
    d = ulps(st["pos"], o.st.pos)
    if build == "parity":
        assert d.max() == 0
    else:
        assert d.max() <= 2
    # every draw edge was exercised: arrivals at the hubs took every offset
    n_hub = sum(len(boundary_draws(c)) for c in COUNTS)
    assert (o.st.road[:n_hub] != ents["roadIndex"][:n_hub]).any()
    # speed == dist and speed one ulp above dist arrive (the dead end turns
    # them round: new target = the road's start); speed one ulp below steps
    turned = (st["tgt"][n_hub:] == roads["start_pos"][ents["roadIndex"][n_hub:]]).all(axis=1)
    assert turned[0::3].all() and not turned[1::3].any() and turned[2::3].all()
    g.sync()
