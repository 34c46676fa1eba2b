"""The batched form (msim_batch_*) through the C-ABI.

Every scenario of a batched run must end bit-identical to the same scenario
run alone through the plain handle (itself parity-tested against the oracle),
and a few scenarios are checked against the oracle directly.  Integer state,
positions, targets and headings are compared bit for bit.
"""
import numpy as np
import pytest

import msim
import orc
from msim import scenarios as S

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_state_equal(a, b, what):
    for k in ("pos", "tgt", "dir"):
        bad = np.flatnonzero((bits(a[k]) != bits(b[k])).any(axis=1))
        assert bad.size == 0, f"{what}: {k} differs for {bad.size} agents, first {bad[:5]}"
    for k in ("road", "rng", "ncount"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")


def oracle_state(o):
    st = o.st
    return {"pos": st.pos, "tgt": st.tgt, "dir": st.dir, "road": st.road, "rng": st.rng, "ncount": st.ncount}


def plain_run(sc, ticks, pc=None, **opts):
    g = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc if pc is None else pc, speed=sc.speed, **opts)
    g.run_ticks(ticks)
    return g


def check_against_plain(scs, ticks, pc=None, **opts):
    print(f"batch of {len(scs)} scenarios, {sum(s.entities.shape[0] for s in scs)} agents, {ticks} ticks, "
          f"options {opts}", flush=True)
    b = msim.Batch(scs, pc=pc, **opts)
    b.run_ticks(ticks)
    tot, last = b.pairs()
    for s, sc in enumerate(scs):
        g = plain_run(sc, ticks, pc=pc)
        assert_state_equal(b.state(s), g.state(), f"scenario {s} ({sc.entities.shape[0]} agents)")
        total, hist = g.pairs()
        assert int(tot[s]) == total, f"scenario {s}: pair total"
        assert int(last[s]) == (int(hist[-1]) if hist.size else 0), f"scenario {s}: last pairs"
        g.close()
    return b


def test_batch_equals_plain_per_scenario():
    """Uneven lattice scenarios (the batched workload's construction, scaled
    down), two 100-tick graph segments plus direct ticks."""
    scs = S.batch_config(12, scale=0.02)
    check_against_plain(scs, 230)


def test_batch_matches_oracle():
    scs = S.batch_config(4, scale=0.01, seed=77)
    b = msim.Batch(scs)
    b.run_ticks(120)
    tot, _ = b.pairs()
    for s, sc in enumerate(scs):
        o = orc.OracleSim(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed)
        o.run_ticks(120)
        assert_state_equal(b.state(s), oracle_state(o), f"scenario {s} vs oracle")
        assert int(tot[s]) == o.pairs_total


def _star(n, seed):
    roads, conns, w, h = S.star_map()
    ents = S.reference_agents(roads, n, seed)
    return S.Scenario(f"star-{n}", roads, conns, ents, None, msim.push_consts(w, h, 10.0), 0, (w, h))


def test_batch_ragged_empty_and_heavy_cells():
    """Constant shader speed, R = 10 m on the star fixture: an empty scenario,
    a single agent, and scenarios whose agents start stacked on the few road
    endpoints (heavy cells), checked against plain runs and the oracle."""
    scs = [_star(64, 3), _star(0, 4), _star(1, 5), _star(700, 6), _star(33, 7)]
    b = check_against_plain(scs, 300)
    for s, sc in enumerate(scs):
        o = orc.OracleSim(sc.roads, sc.conns, sc.entities, sc.pc, speed=None)
        o.run_ticks(300)
        assert_state_equal(b.state(s), oracle_state(o), f"star scenario {s} vs oracle")


@pytest.mark.parametrize("mode", [0, 2])
def test_batch_many_tiny_scenarios_share_waves(mode):
    """60 scenarios of 0 to 9 agents: most waves of the neighbour pass span
    several scenarios (and skip empty ones)."""
    scs = [_star(k % 10, 100 + k) for k in range(60)]
    check_against_plain(scs, 120, collide_mode=mode)


@pytest.mark.parametrize("opts", [dict(binning=2), dict(collide_mode=1), dict(collide_mode=2)])
def test_batch_ragged_empty_and_heavy_cells_other_modes(opts):
    scs = [_star(64, 3), _star(0, 4), _star(1, 5), _star(700, 6), _star(33, 7)]
    check_against_plain(scs, 300, **opts)


@pytest.mark.parametrize("opts", [dict(reorder_interval=7, graph_ticks=10), dict(reorder_interval=0),
                                  dict(reorder_interval=1, graph_ticks=3), dict(graph_ticks=1),
                                  dict(binning=2), dict(binning=2, reorder_interval=1, graph_ticks=3),
                                  dict(binning=2, segment_log2=9, graph_ticks=1),
                                  dict(collide_mode=1), dict(collide_mode=2, binning=2),
                                  dict(bin_unroll=4, binning=2), dict(bin_unroll=8, segment_log2=9),
                                  dict(table_width_log2=7), dict(binning=2, table_width_log2=4)])
def test_batch_options_do_not_change_results(opts):
    scs = S.batch_config(5, scale=0.01, seed=5)
    check_against_plain(scs, 64, **opts)


def test_batch_run_pieces_match_plain_dispatch_loops():
    """Batched renumberings inside captured segments skip the heading and
    neighbour-count arrays; after every run_ticks call of uneven length each
    scenario's state equals a plain dispatch loop's."""
    scs = S.batch_config(3, scale=0.01, seed=31)
    b = msim.Batch(scs, graph_ticks=8, reorder_interval=3)
    plain = [msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed) for sc in scs]
    done = 0
    for k in (1, 10, 2, 25, 9):
        b.run_ticks(k)
        for g in plain:
            for _ in range(k):
                g.sim_tick()
        done += k
        for s, g in enumerate(plain):
            assert_state_equal(b.state(s), g.state(), f"scenario {s} after {done} ticks")
    for g in plain:
        g.close()


def test_batch_initialised_agents_move_on_the_first_tick():
    scs = S.batch_config(3, scale=0.01, seed=9)
    for sc in scs:
        sc.entities["initialized"] = 1
    check_against_plain(scs, 41)


@pytest.mark.parametrize("binning", [0, 2])
def test_batch_repeated_runs_and_time_ticks(binning):
    """run_ticks in pieces equals one run; time_ticks advances the state."""
    scs = S.batch_config(4, scale=0.01, seed=21)
    a = msim.Batch(scs, binning=binning)
    b = msim.Batch(scs, binning=binning)
    for k in (1, 5, 100, 2, 133):
        a.run_ticks(k)
    ms, _ = b.time_ticks(141)
    assert ms > 0.0
    ms, kms = b.time_ticks(50, kernels=True)
    assert ms > 0.0 and set(kms) == set(b.kernel_names()) and all(v > 0.0 for v in kms.values())
    assert ("k_bbin_fused" in kms) == (binning == 2)
    for s in range(len(scs)):
        assert_state_equal(a.state(s), b.state(s), f"scenario {s}")
    np.testing.assert_array_equal(a.pairs()[0], b.pairs()[0])
