"""The C++ Simulator mirror (msim_host_sim_*) and the headless binary on the
GPU: the worker's ticks equal a plain msim_sim run of the same agents bit for
bit, the entity hand-over refreshes after a take, the per-tick CSV log has the
reference's format (Simulator.cpp:388-398), and continue/pause gate the
worker."""
import os
import re
import subprocess

import numpy as np
import pytest

import msim
from msim import host

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "test_map.json")
HEADLESS = os.path.join(os.path.dirname(msim.LIB_PATH), "msim_headless")
LINE = re.compile(r"^(\d{2}):(\d{1,2}):(\d{2})\.(\d{3});(\d+);([0-9.e+-]+);([0-9.e+-]+);([0-9.e+-]+)$")


def plain_run(n, seed, ticks, radius=10.0):
    m = host.HostMap(GOLDEN)
    m.add_guard()
    roads, conns = m.arrays()
    ents = m.make_entities(n, seed)
    i = m.info()
    sim = msim.Simulation(roads, conns, ents, msim.push_consts(i["width"], i["height"], radius))
    sim.run_ticks(ticks)
    return ents, sim.entities()


def test_worker_ticks_equal_a_plain_run_and_hand_over_refreshes(tmp_path):
    n, seed, T = 3000, 5, 30
    csv = tmp_path / "ticks.csv"
    s = host.HostSimulator(GOLDEN, n, 10.0, seed, csv_path=csv, tick_limit=T)
    try:
        first = s.take_entities()  # the creation copy
        ents0, _ = plain_run(n, seed, 0)
        assert first is not None
        np.testing.assert_array_equal(first.view(np.uint8), ents0.view(np.uint8))
        assert s.take_entities() is None  # nothing new until a tick has run
        s.start_worker()
        assert not s.is_simulating()
        s.continue_simulation()
        s.wait_ticks(T)
        st = s.stats()
        assert st["ticks"] == T and not s.is_simulating()
        assert st["tick_ns"] > 0 and st["update_ns"] > 0 and st["collision_ns"] > 0
        # the copy read after the first tick (the creation copy was taken)
        _, after1 = plain_run(n, seed, 1)
        got = s.take_entities()
        np.testing.assert_array_equal(got.view(np.uint8), after1.view(np.uint8))
        # past the limit each continue runs one more tick
        s.continue_simulation()
        s.wait_ticks(T + 1)
        assert s.stats()["ticks"] == T + 1
        _, ref = plain_run(n, seed, T + 1)
        got = s.take_entities()
        np.testing.assert_array_equal(got.view(np.uint8), ref.view(np.uint8))
        s.stop_worker()
    finally:
        s.close()
    lines = csv.read_text().splitlines()
    assert len(lines) == T + 1
    for k, ln in enumerate(lines, start=1):
        m = LINE.match(ln)
        assert m, ln
        assert int(m.group(5)) == k  # PushConsts.tick / 2 after the collision dispatch
        upd, col, tot = (float(m.group(g)) for g in (6, 7, 8))
        assert upd >= 0 and col >= 0 and tot >= 0


def _records_of(o):
    """Entity records the oracle's state reads back as
    (random_move.comp:545-548,876 colours: blue in a pair, else green)."""
    st = o.st
    rec = np.zeros(st.pos.shape[0], msim.ENTITY_DTYPE)
    rec["color"] = np.where((st.ncount > 0)[:, None], [0, 0, 1, 1], [0, 1, 0, 1])
    rec["randomState"] = st.rng
    rec["pos"] = st.pos
    rec["target"] = st.tgt
    rec["direction"] = st.dir
    rec["roadIndex"] = st.road
    rec["initialized"] = 1
    return rec


def test_worker_hand_over_matches_the_oracle(tmp_path):
    """The worker thread's handed-over entity buffer (Simulator.cpp:159-295:
    read after the tick that follows a take) and its pair counts against the
    CPU oracle driven through the same protocol, after ticks 1 and T + 1."""
    import ctypes

    import orc

    n, seed, T = 3000, 9, 40
    m = host.HostMap(GOLDEN)
    m.add_guard()
    roads, conns = m.arrays()
    ents = m.make_entities(n, seed)
    i = m.info()
    o = orc.OracleSim(roads, conns, ents, msim.push_consts(i["width"], i["height"], 10.0))
    s = host.HostSimulator(GOLDEN, n, 10.0, seed, csv_path=tmp_path / "o.csv", tick_limit=T)
    try:
        first = s.take_entities()
        np.testing.assert_array_equal(first.view(np.uint8), ents.view(np.uint8))
        s.start_worker()
        s.continue_simulation()
        s.wait_ticks(T)
        o.run_ticks(1)
        got = s.take_entities()  # read after tick 1, the first tick after the take
        np.testing.assert_array_equal(got.view(np.uint8), _records_of(o).view(np.uint8))
        s.continue_simulation()
        s.wait_ticks(T + 1)
        o.run_ticks(T)
        got = s.take_entities()  # read after tick T + 1
        np.testing.assert_array_equal(got.view(np.uint8), _records_of(o).view(np.uint8))
        lib = msim.load_library()
        total = np.zeros(1, np.uint64)
        per = np.zeros(T + 1, np.uint32)
        k = ctypes.c_uint32()
        assert lib.msim_read_pairs(ctypes.c_void_p(s.handle()), total.ctypes.data, per.ctypes.data, T + 1,
                                   ctypes.byref(k)) == 0
        assert int(total[0]) == o.pairs_total
        assert per[: k.value].tolist() == o.pair_hist[-k.value:]
        dbg = np.zeros(10, np.uint32)
        assert lib.msim_read_debug(ctypes.c_void_p(s.handle()), dbg.ctypes.data) == 0
        assert int(dbg[1]) == o.pairs_total % 2**32
        s.stop_worker()
    finally:
        s.close()


def test_pause_stops_the_worker(tmp_path):
    s = host.HostSimulator(GOLDEN, 500, 10.0, 2, csv_path=tmp_path / "p.csv")
    try:
        s.start_worker()
        s.continue_simulation()
        s.wait_ticks(20)
        s.pause_simulation()
        s.wait_ticks(1 << 62, timeout_ms=10_000)  # returns once paused
        n1 = s.stats()["ticks"]
        import time
        time.sleep(0.2)
        assert s.stats()["ticks"] <= n1 + 1 and not s.is_simulating()
        s.stop_worker()
    finally:
        s.close()


def test_headless_run(tmp_path):
    csv = tmp_path / "h.csv"
    p = subprocess.run([HEADLESS, "--map", GOLDEN, "--agents", "2000", "--ticks", "50", "--csv", str(csv), "--quiet"],
                       capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0
    assert p.stdout.startswith("ticks 50 ")
    lines = csv.read_text().splitlines()
    assert len(lines) == 50 and all(LINE.match(x) for x in lines)
