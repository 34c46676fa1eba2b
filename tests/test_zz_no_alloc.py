"""msim_read_cells_noalloc (the allocation-free readback) against
msim_read_cells, through the C-ABI: two calls in a row, the second on a
handle with fewer agents, in the classic and the sorted-state layouts; a
handle beyond the staging array's size is refused."""
import numpy as np
import pytest

import msim
import orc
from msim import scenarios as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("binning", [0, 3])
def test_cells_noalloc_equals_cells(binning):
    big = S.config(2, scale=0.5)      # 500,000 agents
    small = S.config(1)               # 10,000 agents
    prev = None
    for sc in (big, small):
        g = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, binning=binning)
        g.run_ticks(7)
        a = g.cells()
        b = g.cells_noalloc()
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(b, orc.cells(g.state()["pos"], orc.cell_size(1.0)))
        if prev is not None:
            assert b.shape[0] < prev
        prev = b.shape[0]
        g.close()


def test_cells_noalloc_refuses_more_agents_than_its_array():
    roads, conns, w, h = S.star_map()
    n = (1 << 21) + 1
    ents = S.reference_agents(roads, n, 5)
    g = msim.Simulation(roads, conns, ents, msim.push_consts(w, h, 10.0))
    with pytest.raises(msim.MsimError, match="at most"):
        g.cells_noalloc()
    assert g.cells().shape == (n, 2)
