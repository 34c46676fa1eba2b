"""The sorted-state move dispatch (binning 3, k_move_sorted) against the CPU
oracle, through the C-ABI.

Agent state lives in bucket order and is advanced and rebinned by one kernel
per move dispatch; road and PRNG state stay in caller order.  Every field is
compared bit for bit with the oracle (positions included: -ffp-contract=off)
for the reference's dispatch protocol, fused graph runs of uneven length,
every tuning option, readbacks between dispatches, two move dispatches in a
row, heavy cells, tiny and empty agent sets, and the city and dense workloads
at full size.
"""
import numpy as np
import pytest

import msim
import orc
from msim import scenarios as S
from test_gpu_parity import assert_same, bits, pair, star_scenario

pytestmark = pytest.mark.gpu

SORTED = dict(binning=3)


def test_star_map_reference_protocol_every_dispatch_sorted():
    """Simulator::sim_tick protocol on the star fixture (speed 1.4, R = 10):
    state and pair count after every dispatch."""
    sc = star_scenario()
    g, o = pair(sc, **SORTED)
    assert g.info()["binning"] == 3
    for t in range(300):
        g.dispatch()
        o.dispatch()
        if t < 40 or t % 25 == 0:
            assert_same(g.state(), o.st, f"dispatch {t}")
    assert_same(g.state(), o.st, "end")
    total, hist = g.pairs()
    assert total == o.pairs_total
    assert hist.tolist() == o.pair_hist[-len(hist):]
    assert g.debug_data()[1] == o.pairs_total % 2**32
    assert g.kernel_names() == ["k_seg_scan", "k_move_sorted", "k_route_sorted", "k_collide_staged", "k_collide_finish"]


def test_config1_full_parity_1000_ticks_sorted():
    sc = S.config(1)
    g, o = pair(sc, **SORTED)
    for chunk in range(10):
        g.run_ticks(100)
        o.run_ticks(100)
        assert_same(g.state(), o.st, f"after {100 * (chunk + 1)} ticks")
    total, hist = g.pairs(4096)
    assert total == o.pairs_total
    assert hist.tolist() == o.pair_hist
    assert o.arrivals > 0
    g.sync()


@pytest.mark.parametrize("opts", [dict(), dict(graph_ticks=1), dict(graph_ticks=5), dict(bin_unroll=2),
                                  dict(bin_unroll=8), dict(segment_log2=9, table_width_log2=3),
                                  dict(table_log2=16, segment_log2=14), dict(table_log2=12),
                                  dict(table_log2=20), dict(collide_mode=1), dict(collide_mode=2),
                                  dict(table_log2=16, segment_log2=13, table_width_log2=8, bin_unroll=2),
                                  dict(reorder_interval=5)])
def test_sorted_options_do_not_change_results(opts):
    sc = S.config(1, scale=0.5)
    g, o = pair(sc, **SORTED, **opts)
    g.run_ticks(60)
    o.run_ticks(60)
    assert_same(g.state(), o.st, str(opts))
    assert g.pairs()[0] == o.pairs_total
    g.sync()


def test_sorted_run_pieces_and_readbacks_match_dispatch_loop():
    """Runs of uneven length (captured segments and direct ticks) with full
    readbacks in between equal a dispatch loop of the classic layout."""
    sc = S.config(1, scale=0.3)
    a = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, graph_ticks=8, **SORTED)
    b = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed)
    done = 0
    for k in (1, 9, 17, 2, 33, 16, 8):
        a.run_ticks(k)
        for _ in range(k):
            b.sim_tick()
        done += k
        sa, sb = a.state(), b.state()
        for key in sa:
            np.testing.assert_array_equal(sa[key], sb[key], err_msg=f"{key} after {done} ticks")
        np.testing.assert_array_equal(a.entities().view(np.uint8), b.entities().view(np.uint8))
        np.testing.assert_array_equal(a.cells(), b.cells())
    (ta, ha), (tb, hb) = a.pairs(), b.pairs()
    assert ta == tb
    np.testing.assert_array_equal(ha, hb)
    np.testing.assert_array_equal(a.debug_data(), b.debug_data())


def test_sorted_two_move_dispatches_in_a_row():
    """Two move dispatches without a collision dispatch between them (the
    neighbour counts of the last collision dispatch then follow a slot order
    the second move overwrites), with readbacks after each dispatch."""
    sc = S.config(1, scale=0.2)
    g, o = pair(sc, **SORTED)
    g.run_ticks(4)
    o.run_ticks(4)
    t = int(g.pc["tick"])
    e = t + 1 if (t + 1) % 2 == 0 else t + 2  # the next move dispatch (even tick)
    for tick in (e, e + 2, e + 4, e + 5, e + 6, e + 8, e + 9):  # moves, moves, collision, ...
        g.dispatch(tick)
        o.dispatch(tick)
        assert_same(g.state(), o.st, f"tick {tick}")
    for tick in (e + 10, e + 12):  # two moves, readback only at the end
        g.dispatch(tick)
        o.dispatch(tick)
    assert_same(g.state(), o.st, "two moves")
    g.run_ticks(10)
    o.run_ticks(10)
    assert_same(g.state(), o.st, "run after")


def test_sorted_collision_dispatches_in_a_row_then_captured_runs():
    """Collision dispatches without a move between them, moves after them, a
    run that starts after a move dispatch (each tick a collision then a move
    dispatch), and captured runs right after repeated collision dispatches."""
    sc = S.config(1, scale=0.2)
    g, o = pair(sc, graph_ticks=4, **SORTED)
    g.run_ticks(3)
    o.run_ticks(3)
    t = int(g.pc["tick"])  # the last dispatch was a collision dispatch (odd tick)
    assert t % 2 == 1
    # collision, collision, collision, move, collision, collision, move, move,
    # collision, move
    for tick in (t + 2, t + 4, t + 6, t + 7, t + 8, t + 10, t + 11, t + 13, t + 14, t + 15):
        g.dispatch(tick)
        o.dispatch(tick)
        assert_same(g.state(), o.st, f"tick {tick}")
        assert g.pairs()[0] == o.pairs_total
    g.run_ticks(13)  # after a move dispatch: collision and move dispatches
    o.run_ticks(13)
    assert_same(g.state(), o.st, "run after a move dispatch")
    t = int(g.pc["tick"])
    assert t % 2 == 0
    g.dispatch(t + 1)  # two collision dispatches
    o.dispatch(t + 1)
    g.dispatch(t + 3)
    o.dispatch(t + 3)
    g.run_ticks(9)  # after a collision dispatch: captured segments at once
    o.run_ticks(9)
    assert_same(g.state(), o.st, "run after collision dispatches")
    assert g.pairs()[0] == o.pairs_total
    g.sync()


def test_sorted_entities_and_cells():
    sc = S.config(1, scale=0.1)
    g = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, **SORTED)
    e0 = g.entities()
    np.testing.assert_array_equal(e0.view(np.uint8), sc.entities.view(np.uint8))
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
    np.testing.assert_array_equal(g.entities(10, 7).view(np.uint8), e[10:17].view(np.uint8))
    np.testing.assert_array_equal(g.cells(), orc.cells(st["pos"], orc.cell_size(1.0)))


def test_sorted_heavy_cell_and_mixed_initialized():
    """3,000 agents stacked in one cell, half of all agents uninitialised,
    the first dispatch a collision dispatch: the classic initialisation
    dispatches, then the sorted state from the first move of initialised
    agents on."""
    sc = S.config(1, scale=0.5)
    ents = sc.entities.copy()
    k = 3000
    ents["pos"][:k] = ents["pos"][0]
    ents["initialized"] = (np.random.default_rng(5).random(ents.shape[0]) < 0.5).astype(np.uint32)
    pc = sc.pc.copy()
    pc["tick"] = 2
    sc2 = S.Scenario("stack-mixed", sc.roads, sc.conns, ents, sc.speed, pc, 0, sc.world)
    g, o = pair(sc2, **SORTED)
    for t in range(8):
        g.dispatch()
        o.dispatch()
        assert_same(g.state(), o.st, f"dispatch {t}", check_dir=(t > 0))
    g.run_ticks(20)
    o.run_ticks(20)
    assert_same(g.state(), o.st, "after 20 ticks")
    assert g.pairs()[0] == o.pairs_total


def test_sorted_tiny_and_empty():
    roads, conns, w, h = S.star_map()
    for n in (1, 2, 3, 5):
        ents = S.reference_agents(roads, n, 7)
        g = msim.Simulation(roads, conns, ents, msim.push_consts(w, h, 10.0), **SORTED)
        o = orc.OracleSim(roads, conns, ents, msim.push_consts(w, h, 10.0))
        g.run_ticks(40)
        o.run_ticks(40)
        assert_same(g.state(), o.st, f"n={n}")
    g = msim.Simulation(roads, conns, np.zeros(0, msim.ENTITY_DTYPE), msim.push_consts(w, h, 10.0), **SORTED)
    g.run_ticks(10)
    assert g.pairs()[0] == 0 and g.entities().shape == (0,)


def test_sorted_radius_change_through_push_constants():
    sc = S.config(1, scale=0.2)
    g, o = pair(sc, **SORTED)
    g.run_ticks(3)
    o.run_ticks(3)
    pc = g.pc.copy()
    pc["collisionRadius"] = 0.75
    pc["tick"] = int(pc["tick"]) + 1
    g.pc = pc
    g.dispatch(int(pc["tick"]))
    o.radius = 0.75
    o.dispatch(int(pc["tick"]))
    g.dispatch()
    o.dispatch()
    assert_same(g.state(), o.st, "radius 0.75")
    g.run_ticks(7)
    o.run_ticks(7)
    assert_same(g.state(), o.st, "radius 0.75, fused")


def test_sorted_config2_full_size_parity():
    """1,000,000 agents on the 150k-node city graph, 30 ticks."""
    sc = S.config(2)
    g, o = pair(sc, threads=16, **SORTED)
    g.run_ticks(30)
    o.run_ticks(30)
    assert_same(g.state(), o.st, "config 2 full")
    total, hist = g.pairs()
    assert total == o.pairs_total and hist.tolist() == o.pair_hist
    g.sync()


def test_sorted_config3_full_size_parity():
    """16,000,000 agents with Zipf road occupancy on the 1M-node graph, 3
    ticks (1,024 segments chained by look-back); the device checks that every
    agent is binned exactly once and stays in the bucket it was counted in."""
    sc = S.config(3)
    g, o = pair(sc, threads=16, **SORTED)
    g.run_ticks(3)
    o.run_ticks(3)
    g.sync()
    assert_same(g.state(), o.st, "config 3 full")
    total, hist = g.pairs()
    assert total == o.pairs_total and hist.tolist() == o.pair_hist
