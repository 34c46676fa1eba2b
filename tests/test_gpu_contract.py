"""The contraction-on build (libmsim_contract.so, -ffp-contract=fast) against
the parity build and against a binary64 replay of the step (error within
0.5 ulp of the new position plus 3 ulp of the speed, like the parity build).

From the same state, one tick of each build makes the same decisions (the
integer-deciding arithmetic -- route offset, distance-vs-speed, neighbour
predicate -- keeps one rounding per operation in both builds), and the
contracted positions stay within 2 ulp of the parity build's.  Both builds'
errors against the binary64 step are stated in the test output.
"""
import numpy as np
import pytest

import msim
import orc
from msim import scenarios as S

pytestmark = pytest.mark.gpu


def ulps(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


def f64_ulps(x32, x64):
    """|x32 - x64| in ulps of x32."""
    x32 = np.asarray(x32, np.float32)
    sp = np.spacing(np.abs(x32)).astype(np.float64)
    return np.abs(x32.astype(np.float64) - x64) / sp


@pytest.mark.parametrize("cfg,scale,warm", [(1, 1.0, 57), (2, 0.2, 23)])
def test_contracted_tick_within_two_ulp(cfg, scale, warm):
    sc = S.config(cfg, scale=scale)
    fast_lib = msim.load_library(msim.LIB_CONTRACT_PATH)
    assert b"fp-contract=fast" in fast_lib.msim_version()
    base = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed)
    base.run_ticks(warm)
    ents = base.entities()  # a mid-run state, every agent initialised
    assert (ents["initialized"] == 1).all()
    pc = msim.push_consts(*sc.world, float(sc.pc["collisionRadius"]), tick=int(base.pc["tick"]))
    a = msim.Simulation(sc.roads, sc.conns, ents, pc, speed=sc.speed)
    b = msim.Simulation(sc.roads, sc.conns, ents, pc, speed=sc.speed, library=fast_lib)
    a.run_ticks(1)
    b.run_ticks(1)
    sa, sb = a.state(), b.state()
    for k in ("road", "rng", "ncount", "tgt"):
        np.testing.assert_array_equal(np.asarray(sa[k]).view(np.uint32), np.asarray(sb[k]).view(np.uint32),
                                      err_msg=k)
    d = ulps(sa["pos"], sb["pos"])
    p64, arr = orc.advance_f64(ents["pos"], ents["target"], sc.speed)
    moving = arr == 0
    e_off = f64_ulps(sa["pos"][moving], p64[moving])
    e_fast = f64_ulps(sb["pos"][moving], p64[moving])
    print(f"config {cfg}: {moving.sum()} moving agents; contracted vs parity build max {d.max()} ulp "
          f"({(d > 0).mean():.1%} of coordinates differ); error vs binary64 step: parity build max "
          f"{e_off.max():.2f} ulp (mean {e_off.mean():.3f}), contracted build max {e_fast.max():.2f} ulp "
          f"(mean {e_fast.mean():.3f})")
    assert d.max() <= 2
    # the step's error bound (as for the oracle): 0.5 ulp of the new position
    # plus 3 ulp of the heading's magnitude (the speed)
    spd = np.full(ents.shape[0], 1.4, np.float32) if sc.speed is None else sc.speed
    for pos in (sa["pos"], sb["pos"]):
        err = np.abs(pos[moving].astype(np.float64) - p64[moving])
        bound = 0.5 * np.spacing(np.abs(pos[moving])) + 3.0 * np.spacing(spd[moving])[:, None]
        assert (err <= bound).all()
