"""The multi-threaded oracle (orc_move_phase_mt, orc_collide_grid_mt), which
checks the full-size GPU runs, against the single-threaded oracle that the
golden vectors and brute force pin: identical state and pair history."""
import numpy as np
import pytest

import orc
from msim import scenarios as S


def _run(sc, threads, ticks):
    o = orc.OracleSim(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, threads=threads)
    o.run_ticks(ticks)
    return o


@pytest.mark.parametrize("cfg,scale,ticks", [(1, 1.0, 150), (2, 0.02, 60), (3, 0.002, 40)])
def test_threaded_oracle_equals_single_threaded(cfg, scale, ticks):
    sc = S.config(cfg, scale=scale)
    a = _run(sc, 1, ticks)
    b = _run(sc, 7, ticks)
    for k in ("pos", "tgt", "dir"):
        np.testing.assert_array_equal(getattr(a.st, k).view(np.uint32), getattr(b.st, k).view(np.uint32), err_msg=k)
    for k in ("road", "rng", "ncount"):
        np.testing.assert_array_equal(getattr(a.st, k), getattr(b.st, k), err_msg=k)
    assert a.pair_hist == b.pair_hist and a.pairs_total == b.pairs_total
    assert a.arrivals > 0  # the threaded move does not count arrivals
