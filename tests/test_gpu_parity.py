"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Integer state (road index, xorshift state, neighbour counts, pair totals,
cells) must be bit-exact; binary32 positions, targets and headings are
compared bit for bit too (the library is built with -ffp-contract=off).
"""
import numpy as np
import pytest

import msim
import orc
from msim import scenarios as S

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_same(gpu, ref_state, what="", check_dir=True):
    st = ref_state
    for name, g, o in (("pos", gpu["pos"], st.pos), ("tgt", gpu["tgt"], st.tgt)):
        gb, ob = bits(g), bits(o)
        bad = np.flatnonzero((gb != ob).any(axis=1))
        assert bad.size == 0, f"{what}: {name} differs for {bad.size} agents, first {bad[:5]} " \
                              f"gpu={g[bad[:3]]} oracle={o[bad[:3]]}"
    if check_dir:
        assert (bits(gpu["dir"]) == bits(st.dir)).all(), f"{what}: dir differs"
    np.testing.assert_array_equal(gpu["road"], st.road, err_msg=f"{what}: road")
    np.testing.assert_array_equal(gpu["rng"], st.rng, err_msg=f"{what}: rng")
    np.testing.assert_array_equal(gpu["ncount"], st.ncount, err_msg=f"{what}: neighbour counts")


def pair(sc, threads=1, **opts):
    g = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, **opts)
    o = orc.OracleSim(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, threads=threads)
    return g, o


def star_scenario(n=64, seed=3):
    roads, conns, w, h = S.star_map()
    ents = S.reference_agents(roads, n, seed)
    return S.Scenario("star", roads, conns, ents, None, msim.push_consts(w, h, 10.0), 0, (w, h))


def test_star_map_reference_protocol_every_dispatch():
    """Simulator::sim_tick protocol on the star fixture with the shader's
    constants (speed 1.4, R = 10): state and pair count after every dispatch."""
    sc = star_scenario()
    g, o = pair(sc)
    for t in range(300):
        g.dispatch()
        o.dispatch()
        if t < 40 or t % 25 == 0:
            assert_same(g.state(), o.st, f"dispatch {t}")
    assert_same(g.state(), o.st, "end")
    total, hist = g.pairs()
    assert total == o.pairs_total
    assert hist.tolist() == o.pair_hist[-len(hist):]
    dbg = g.debug_data()
    assert dbg[1] == o.pairs_total % 2**32


def test_config1_full_parity_1000_ticks():
    sc = S.config(1)
    g, o = pair(sc)
    for chunk in range(10):
        g.run_ticks(100)
        o.run_ticks(100)
        assert_same(g.state(), o.st, f"after {100 * (chunk + 1)} ticks")
    total, hist = g.pairs(4096)
    assert total == o.pairs_total
    assert hist.tolist() == o.pair_hist
    assert o.arrivals > 0


def test_dispatch_loop_equals_fused_graph_path():
    sc = S.config(1, scale=0.3)
    a = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, graph_ticks=8)
    b = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed)
    a.run_ticks(77)
    for _ in range(77):
        b.sim_tick()
    sa, sb = a.state(), b.state()
    for k in sa:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
    (ta, ha), (tb, hb) = a.pairs(), b.pairs()
    assert ta == tb
    np.testing.assert_array_equal(ha, hb)


@pytest.mark.parametrize("reorder", [1, 4])
def test_run_pieces_match_dispatch_loop_after_each_piece(reorder):
    """Renumberings inside captured run segments skip the heading and
    neighbour-count arrays (the run's last tick rewrites both).  After every
    run_ticks call of uneven length (some replay captured segments, some end
    in direct ticks) the whole state equals a dispatch loop's."""
    sc = S.config(1, scale=0.3)
    a = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, graph_ticks=8,
                        reorder_interval=reorder)
    b = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, reorder_interval=reorder)
    done = 0
    for k in (1, 9, 17, 2, 33, 16, 8):
        a.run_ticks(k)
        for _ in range(k):
            b.sim_tick()
        done += k
        sa, sb = a.state(), b.state()
        for key in sa:
            np.testing.assert_array_equal(sa[key], sb[key], err_msg=f"{key} after {done} ticks")


@pytest.mark.parametrize("opts", [dict(reorder_interval=1), dict(reorder_interval=13, graph_ticks=5),
                                  dict(graph_ticks=1), dict(table_log2=12), dict(table_log2=20),
                                  dict(binning=1), dict(binning=1, reorder_interval=5),
                                  dict(segment_log2=9, table_width_log2=3),
                                  dict(table_log2=16, segment_log2=13, table_width_log2=8),
                                  dict(route_queue=1), dict(collide_mode=1),
                                  dict(route_queue=1, binning=1, collide_mode=1),
                                  dict(binning=2), dict(binning=2, reorder_interval=5, segment_log2=9),
                                  dict(binning=2, table_log2=16, segment_log2=13, table_width_log2=8),
                                  dict(binning=2, route_queue=1, graph_ticks=1),
                                  dict(collide_mode=2), dict(collide_mode=2, segment_log2=9, table_width_log2=3),
                                  dict(bin_unroll=4), dict(bin_unroll=8, binning=2, table_log2=16, segment_log2=14),
                                  dict(table_log2=16, segment_log2=14)])
def test_options_do_not_change_results(opts):
    sc = S.config(1, scale=0.5)
    g, o = pair(sc, **opts)
    g.run_ticks(60)
    o.run_ticks(60)
    assert_same(g.state(), o.st, str(opts))
    assert g.pairs()[0] == o.pairs_total


def test_route_choice_off_node_targets_and_shared_groups():
    """Route choice away from the usual geometry (random_move.comp:778-819):
    agents whose target is off their road's endpoints (the endpoint test
    falls to the road's end, the new target to the next road's start), and
    a corner endpoint sharing the centre's connection list (one entry drawn
    from two positions).  Dispatch loop, fused ticks and a lattice map all
    match the oracle."""
    roads, conns, w, h = S.star_map()
    roads = roads.copy()
    roads["start_index"][0] = roads["end_index"][0]  # corner (5, 5) draws the centre's list
    roads["start_count"][0] = roads["end_count"][0]
    ents = S.reference_agents(roads, 96, 11)
    rng = np.random.default_rng(4)
    k = ents.shape[0] // 3
    ents["target"][:k] += rng.uniform(-3.0, 3.0, (k, 2)).astype(np.float32)  # off the node
    sc = S.Scenario("star-shared", roads, conns, ents, None, msim.push_consts(w, h, 10.0), 0, (w, h))
    g, o = pair(sc)
    for t in range(200):
        g.dispatch()
        o.dispatch()
        if t % 20 == 0:
            assert_same(g.state(), o.st, f"dispatch {t}")
    assert_same(g.state(), o.st, "dispatch loop")
    g2, o2 = pair(sc, graph_ticks=16)
    g2.run_ticks(150)
    o2.run_ticks(150)
    assert_same(g2.state(), o2.st, "fused ticks")
    big = S.config(1, scale=0.2)
    e2 = big.entities.copy()
    k = e2.shape[0] // 3
    e2["target"][:k] += np.random.default_rng(5).uniform(-2.0, 2.0, (k, 2)).astype(np.float32)
    sc3 = S.Scenario("lattice-off-node", big.roads, big.conns, e2, big.speed, big.pc, 0, big.world)
    g3, o3 = pair(sc3)
    g3.run_ticks(120)
    o3.run_ticks(120)
    assert_same(g3.state(), o3.st, "lattice, off-node targets")


def test_mixed_initialized_flags_and_collision_first():
    """Agents with Entity.initialized == 0 only enter the index on their first
    dispatch and do not search in a collision dispatch
    (random_move.comp:863-867)."""
    sc = S.config(1, scale=0.2)
    rng = np.random.default_rng(5)
    ents = sc.entities.copy()
    ents["initialized"] = (rng.random(ents.shape[0]) < 0.5).astype(np.uint32)
    # bunch agents so that pairs exist
    ents["pos"][: ents.shape[0] // 2] = ents["pos"][ents.shape[0] // 2: 2 * (ents.shape[0] // 2)]
    pc = sc.pc.copy()
    pc["tick"] = 2  # next dispatch (tick 3) is a collision dispatch
    sc2 = S.Scenario("mixed", sc.roads, sc.conns, ents, sc.speed, pc, 0, sc.world)
    g, o = pair(sc2)
    for t in range(6):
        g.dispatch()
        o.dispatch()
        assert_same(g.state(), o.st, f"dispatch {t}", check_dir=(t > 0))
    assert g.pairs()[0] == o.pairs_total
    assert g.info()["n_uninitialized"] == 0


def test_config2_scaled_parity():
    sc = S.config(2, scale=0.1)
    g, o = pair(sc, threads=8)
    g.run_ticks(150)
    o.run_ticks(150)
    assert_same(g.state(), o.st, "config 2 at 10%")
    assert g.pairs()[0] == o.pairs_total


def test_config2_full_size_parity():
    """1,000,000 agents on the 150k-node city graph, 30 ticks: full state and
    every per-tick pair count."""
    sc = S.config(2)
    g, o = pair(sc, threads=16)
    g.run_ticks(30)
    o.run_ticks(30)
    assert_same(g.state(), o.st, "config 2 full")
    total, hist = g.pairs()
    assert total == o.pairs_total and hist.tolist() == o.pair_hist


def test_one_pass_binning_equals_two_pass_at_full_size():
    """1,000,000 agents (512 rebinning segments chained by look-back,
    across every XCD), 200 ticks: the one-pass rebinning gives the state and
    pair history of the two-pass one."""
    sc = S.config(2)
    a = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed)
    b = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, binning=2)
    assert b.info()["binning"] == 2
    a.run_ticks(200)
    b.run_ticks(200)
    sa, sb = a.state(), b.state()
    for k in sa:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
    (ta, ha), (tb, hb) = a.pairs(), b.pairs()
    assert ta == tb
    np.testing.assert_array_equal(ha, hb)
    assert "k_bin_fused" in b.kernel_names()


def test_entity_readback_layout_and_colours():
    sc = S.config(1, scale=0.1)
    g = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed)
    e0 = g.entities()
    # before any dispatch the records read back exactly as uploaded (device pack of the
    # device-unpacked state)
    np.testing.assert_array_equal(e0.view(np.uint8), sc.entities.view(np.uint8))
    assert (e0["initialized"] == 0).all()
    g.run_ticks(5)
    e = g.entities()
    st = g.state()
    assert (bits(e["pos"]) == bits(st["pos"])).all()
    assert (bits(e["direction"]) == bits(st["dir"])).all()
    assert (e["roadIndex"] == st["road"]).all() and (e["randomState"] == st["rng"]).all()
    assert (e["initialized"] == 1).all()
    blue = st["ncount"] > 0
    np.testing.assert_array_equal(e["color"][blue], np.tile([0, 0, 1, 1], (blue.sum(), 1)))
    np.testing.assert_array_equal(e["color"][~blue], np.tile([0, 1, 0, 1], ((~blue).sum(), 1)))
    part = g.entities(10, 7)
    np.testing.assert_array_equal(part.view(np.uint8), e[10:17].view(np.uint8))
    # after spatial renumbering every few ticks the records still come back in caller order
    h = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, reorder_interval=3)
    h.run_ticks(5)
    np.testing.assert_array_equal(h.entities().view(np.uint8), e.view(np.uint8))


def test_cells_export_matches_oracle():
    sc = S.config(1, scale=0.2)
    g = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed)
    g.run_ticks(3)
    pos = g.state()["pos"]
    np.testing.assert_array_equal(g.cells(), orc.cells(pos, orc.cell_size(1.0)))


def test_heavy_cell_and_duplicates():
    """3,000 agents stacked in one 1 m cell (every pair in range) plus sparse
    agents: counts n-1 in the stack, pairs n(n-1)/2."""
    sc = S.config(1, scale=0.5)
    ents = sc.entities.copy()
    k = 3000
    ents["pos"][:k] = ents["pos"][0]
    ents["initialized"] = 1
    sc2 = S.Scenario("stack", sc.roads, sc.conns, ents, sc.speed, sc.pc, 0, sc.world)
    g, o = pair(sc2)
    g.dispatch(3)
    o.dispatch(3)
    st = g.state()
    assert (st["ncount"][:k] >= k - 1).all()
    assert_same(st, o.st, "stack", check_dir=False)
    assert g.pairs()[0] == o.pairs_total >= k * (k - 1) // 2


def test_tiny_and_empty():
    roads, conns, w, h = S.star_map()
    for n in (1, 2, 3):
        ents = S.reference_agents(roads, n, 7)
        g = msim.Simulation(roads, conns, ents, msim.push_consts(w, h, 10.0))
        o = orc.OracleSim(roads, conns, ents, msim.push_consts(w, h, 10.0))
        g.run_ticks(40)
        o.run_ticks(40)
        assert_same(g.state(), o.st, f"n={n}")
    g = msim.Simulation(roads, conns, np.zeros(0, msim.ENTITY_DTYPE), msim.push_consts(w, h, 10.0))
    g.run_ticks(10)
    assert g.pairs()[0] == 0 and g.entities().shape == (0,)


def test_radius_change_through_push_constants():
    sc = S.config(1, scale=0.2)
    g, o = pair(sc)
    g.run_ticks(3)
    o.run_ticks(3)
    pc = g.pc.copy()
    pc["collisionRadius"] = 0.75  # within the 1 m cell chosen at creation
    pc["tick"] = int(pc["tick"]) + 1
    g.pc = pc
    g.dispatch(int(pc["tick"]))
    o.radius = 0.75
    o.dispatch(int(pc["tick"]))
    g.dispatch()
    o.dispatch()
    assert_same(g.state(), o.st, "radius 0.75")
    pc["collisionRadius"] = 3.0
    with pytest.raises(msim.MsimError):
        g.pc = pc
        g.dispatch(int(pc["tick"]) + 2)


def test_two_handles_are_deterministic():
    sc = S.config(1, scale=0.3)
    a = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed)
    b = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, reorder_interval=3)
    a.run_ticks(40)
    b.run_ticks(40)
    sa, sb = a.state(), b.state()
    for k in sa:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)


def test_config3_full_size_parity():
    """16,000,000 agents with Zipf road occupancy on the 1M-node graph (max
    cell load >= 200, ~5e8 pairs per tick): full state and pair counts after
    3 ticks; every agent is binned exactly once (checked on the device)."""
    sc = S.config(3)
    assert sc.meta["max_cell_load"] >= 200
    g, o = pair(sc, threads=16)
    g.run_ticks(3)
    o.run_ticks(3)
    g.sync()
    assert_same(g.state(), o.st, "config 3 full")
    total, hist = g.pairs()
    assert total == o.pairs_total and hist.tolist() == o.pair_hist


@pytest.mark.parametrize("radius", [1.0, 0.75, 10.0, 2.5])
def test_boundary_pairs_match_brute_force(radius):
    """Crafted pairs at exactly R, one ulp inside, stacked duplicates and
    cell-edge positions, through one collision dispatch, against the oracle's
    O(N^2) pass."""
    roads, conns, w, h = S.star_map()
    rng = np.random.default_rng(int(radius * 10))
    n = 6000
    r32 = np.float32(radius)
    p = rng.uniform(2, 98, (n, 2)).astype(np.float32)
    p[:500, 0] = np.floor(p[:500, 0])  # on cell edges
    p[500:1000] = p[1000:1500]  # duplicates
    p[1500:2000, 0] = p[2000:2500, 0] + r32  # exactly R apart
    p[1500:2000, 1] = p[2000:2500, 1]
    p[2500:3000, 0] = np.nextafter(p[3000:3500, 0] + r32, np.float32(-1))  # just inside
    p[2500:3000, 1] = p[3000:3500, 1]
    ents = S.reference_agents(roads, n, 3)
    ents["pos"] = p
    ents["initialized"] = 1
    pc = msim.push_consts(w, h, radius, tick=2)
    g = msim.Simulation(roads, conns, ents, pc)
    g.dispatch(3)
    st = g.state()
    c = np.zeros(n, np.uint32)
    pairs = orc.L.orc_collide_brute(n, orc._p(np.ascontiguousarray(p)), radius, None, orc._p(c))
    np.testing.assert_array_equal(st["ncount"], c)
    assert g.pairs()[0] == pairs


def test_reference_map_file_drives_the_hip_path():
    """The reference's fixture map, loaded from its JSON (guard entry added),
    with the shader's constants: every dispatch against the oracle."""
    import os

    from msim import mapio

    m = mapio.load_map(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "test_map.json"))
    conns = mapio.with_guard(m.conns, m.roads)
    ents = S.reference_agents(m.roads, 96, 17)
    pc = msim.push_consts(m.width, m.height, 10.0)
    g = msim.Simulation(m.roads, conns, ents, pc)
    o = orc.OracleSim(m.roads, conns, ents, pc)
    for t in range(400):
        g.dispatch()
        o.dispatch()
    assert_same(g.state(), o.st, "reference map")
    assert g.pairs()[0] == o.pairs_total
