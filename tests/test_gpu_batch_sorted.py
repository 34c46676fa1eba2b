"""The batched form on bucket-sorted state (binning 3, k_bmove_sorted).

Every scenario must end bit-identical to its own plain run (and, for a few,
to the oracle), for uneven scenario sizes, empty and one-agent scenarios,
heavy cells, many tiny scenarios sharing waves, every option, runs of uneven
length, and the hand-over to the differentiable tick.
"""
import numpy as np
import pytest

import msim
import orc
from msim import scenarios as S
from test_gpu_batch import _star, assert_state_equal, check_against_plain, oracle_state

pytestmark = pytest.mark.gpu

SORTED = dict(binning=3)


def test_batch_sorted_equals_plain_per_scenario():
    scs = S.batch_config(12, scale=0.02)
    b = check_against_plain(scs, 230, **SORTED)
    assert b.kernel_names() == ["k_seg_scan", "k_bmove_sorted", "k_broute_sorted", "k_bcollide_staged", "k_bfinish"]
    b.sync()


def test_batch_sorted_matches_oracle():
    scs = S.batch_config(4, scale=0.01, seed=77)
    b = msim.Batch(scs, **SORTED)
    b.run_ticks(120)
    tot, _ = b.pairs()
    for s, sc in enumerate(scs):
        o = orc.OracleSim(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed)
        o.run_ticks(120)
        assert_state_equal(b.state(s), oracle_state(o), f"scenario {s} vs oracle")
        assert int(tot[s]) == o.pairs_total


def test_batch_sorted_ragged_empty_and_heavy_cells():
    scs = [_star(64, 3), _star(0, 4), _star(1, 5), _star(700, 6), _star(33, 7)]
    b = check_against_plain(scs, 300, **SORTED)
    for s, sc in enumerate(scs):
        o = orc.OracleSim(sc.roads, sc.conns, sc.entities, sc.pc, speed=None)
        o.run_ticks(300)
        assert_state_equal(b.state(s), oracle_state(o), f"star scenario {s} vs oracle")


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_batch_sorted_many_tiny_scenarios_share_waves(mode):
    scs = [_star(k % 10, 100 + k) for k in range(60)]
    check_against_plain(scs, 120, collide_mode=mode, **SORTED)


@pytest.mark.parametrize("opts", [dict(graph_ticks=10), dict(graph_ticks=1), dict(graph_ticks=3, segment_log2=9),
                                  dict(bin_unroll=2), dict(bin_unroll=8), dict(segment_log2=12),
                                  dict(collide_mode=1), dict(collide_mode=2), dict(reorder_interval=4),
                                  dict(table_width_log2=3), dict(table_width_log2=7, segment_log2=11),
                                  dict(table_width_log2=30)])
def test_batch_sorted_options_do_not_change_results(opts):
    scs = S.batch_config(5, scale=0.01, seed=5)
    check_against_plain(scs, 64, **SORTED, **opts)


def test_batch_sorted_run_pieces_match_plain_dispatch_loops():
    scs = S.batch_config(3, scale=0.01, seed=31)
    b = msim.Batch(scs, graph_ticks=8, **SORTED)
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


def test_batch_sorted_initialised_agents_and_time_ticks():
    scs = S.batch_config(4, scale=0.01, seed=21)
    for sc in scs:
        sc.entities["initialized"] = 1
    a = msim.Batch(scs, **SORTED)
    b = msim.Batch(scs, **SORTED)
    for k in (1, 5, 100, 2, 133):
        a.run_ticks(k)
    ms, _ = b.time_ticks(141)
    assert ms > 0.0
    ms, kms = b.time_ticks(50, kernels=True)  # two passes of 50 ticks: 241 in all, like a
    assert ms > 0.0 and set(kms) == set(b.kernel_names()) and all(v > 0.0 for v in kms.values())
    for s in range(len(scs)):
        assert_state_equal(a.state(s), b.state(s), f"scenario {s}")
    np.testing.assert_array_equal(a.pairs()[0], b.pairs()[0])
    check_against_plain(scs, 41, **SORTED)


def test_batch_sorted_hands_over_to_the_differentiable_tick():
    """A sorted-state handle set up for the differentiable tick after some
    plain ticks: its neutral forward run (theta = 1, alpha = 0) continues
    exactly like a classic-layout handle's."""
    import torch

    scs = S.batch_config(3, scale=0.01, seed=41)
    a = msim.Batch(scs, **SORTED)
    b = msim.Batch(scs)
    a.run_ticks(17)
    b.run_ticks(17)
    for s in range(len(scs)):
        assert_state_equal(a.state(s), b.state(s), f"scenario {s} before the hand-over")
    outs = []
    for h in (a, b):
        h.diff_setup(30, 10)
        theta = torch.ones(h.n_roads, dtype=torch.float32, device="cuda")
        out = torch.empty(h.n, 2, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        h.diff_forward(theta.data_ptr(), 0.0, out.data_ptr())
        outs.append(out.cpu().numpy())
    np.testing.assert_array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    for s in range(len(scs)):
        assert_state_equal(a.state(s), b.state(s), f"scenario {s} after the differentiable run")
