"""Grid occupancy export (msim_read_grid_occupancy), the uniform-grid
counterpart of the quadtree nodes the UI overlay draws
(Simulator.cpp:149-153, QuadTreeGridGlObject.cpp:30-50): every occupied cell
with its agent count, against a histogram of the oracle's per-agent cells."""
import numpy as np
import pytest

import msim
import orc
from msim import scenarios as S

pytestmark = pytest.mark.gpu


def histogram(pos):
    c = orc.cells(pos, orc.cell_size(1.0))
    cells, counts = np.unique(c, axis=0, return_counts=True)
    return cells, counts.astype(np.uint32)


def sorted_entries(cells, counts):
    """Rows in np.unique's order: by cx, then cy."""
    order = np.lexsort((cells[:, 1], cells[:, 0]))
    return cells[order], counts[order]


@pytest.mark.parametrize("opts", [dict(), dict(binning=2, reorder_interval=4), dict(binning=3),
                                  dict(binning=1), dict(table_log2=12)])
def test_occupancy_matches_histogram_of_cells(opts):
    sc = S.config(1, scale=0.5)
    g, o = (msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, **opts),
            orc.OracleSim(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed))
    for ticks in (0, 1, 13):
        g.run_ticks(ticks)
        o.run_ticks(ticks)
        cells, counts, total = g.grid_occupancy()
        want_c, want_n = histogram(o.st.pos)
        assert total == want_c.shape[0] == cells.shape[0]
        got_c, got_n = sorted_entries(cells, counts)
        np.testing.assert_array_equal(got_c, want_c)
        np.testing.assert_array_equal(got_n, want_n)
        assert counts.sum() == sc.entities.shape[0]


def test_occupancy_folded_table_stacks_and_capacity():
    """A small table folds many world cells onto each bucket; 2,000 agents
    stacked in one cell; a capacity below the number of occupied cells
    writes that many entries and reports the full count."""
    sc = S.config(2, scale=0.1)
    ents = sc.entities.copy()
    ents["pos"][:2000] = ents["pos"][7]
    sc2 = S.Scenario("stack", sc.roads, sc.conns, ents, sc.speed, sc.pc, 0, sc.world)
    g = msim.Simulation(sc2.roads, sc2.conns, sc2.entities, sc2.pc, speed=sc2.speed, table_log2=12, binning=3)
    # before the first tick the 2,000 stacked agents share one cell
    cells, counts, total = g.grid_occupancy()
    assert counts.max() >= 2000
    got_c, got_n = sorted_entries(cells, counts)
    want_c, want_n = histogram(ents["pos"])
    np.testing.assert_array_equal(got_c, want_c)
    np.testing.assert_array_equal(got_n, want_n)
    g.run_ticks(3)
    pos = g.state()["pos"]
    cells, counts, total = g.grid_occupancy()
    want_c, want_n = histogram(pos)
    got_c, got_n = sorted_entries(cells, counts)
    np.testing.assert_array_equal(got_c, want_c)
    np.testing.assert_array_equal(got_n, want_n)
    part, pcnt, total2 = g.grid_occupancy(capacity=100)
    assert total2 == total and part.shape == (100, 2)
    # every returned entry is one of the occupied cells with its count
    lookup = {tuple(c): n for c, n in zip(want_c.tolist(), want_n.tolist())}
    assert all(lookup[tuple(c)] == n for c, n in zip(part.tolist(), pcnt.tolist()))
