"""The C++ host driver (include/msim/msim_host.h) on the CPU: the map loader
against msim.mapio and the reference fixture, seeded agent creation, the
tick-rate and tick-duration statistics (src/utils/TickRate.cpp,
TickDurationHistory.cpp) against closed forms, the tick log's time stamp
(Simulator.cpp:348-385) and the headless binary's argument handling."""
import json
import os
import subprocess

import numpy as np
import pytest

import msim
from msim import host, mapio
from msim import scenarios as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "test_map.json")
HEADLESS = os.path.join(os.path.dirname(msim.LIB_PATH), "msim_headless")


def same_map(cpp, py):
    roads, conns = cpp.arrays()
    i = cpp.info()
    assert (i["width"], i["height"], i["skipped"]) == (py.width, py.height, py.skipped)
    np.testing.assert_array_equal(roads.view(np.uint32), py.roads.view(np.uint32))
    np.testing.assert_array_equal(conns, py.conns)


def test_cpp_loader_reads_the_reference_fixture_like_mapio():
    same_map(host.HostMap(GOLDEN), mapio.load_map(GOLDEN))


@pytest.mark.parametrize("layout", ["own_first", "node_csr"])
def test_cpp_loader_on_a_generated_city(tmp_path, layout):
    sc = S.config(2, scale=0.02, layout=layout)
    p = tmp_path / "city.json"
    mapio.save_map(str(p), *sc.world, sc.roads, sc.conns)
    m = host.HostMap(p)
    same_map(m, mapio.load_map(str(p)))
    roads, conns = m.arrays()
    np.testing.assert_array_equal(roads.view(np.uint32), sc.roads.view(np.uint32))
    np.testing.assert_array_equal(conns, sc.conns)


def test_cpp_loader_skips_zero_length_roads(tmp_path):
    j = json.load(open(GOLDEN))
    zero = dict(j["roads"][0])
    zero["end"] = dict(zero["start"])
    j["roads"].insert(1, zero)
    p = tmp_path / "m.json"
    p.write_text(json.dumps(j))
    m = host.HostMap(p)
    assert m.info()["skipped"] == 1 and m.info()["n_roads"] == 4
    same_map(m, mapio.load_map(str(p)))


@pytest.mark.parametrize("where,drop", [("top", "maxDistLong"), ("top", "roads"),
                                        ("top", "connectionRoadIndexList"), ("road", "connCountStart"),
                                        ("road", "end"), ("coord", "distLong")])
def test_cpp_loader_missing_fields(tmp_path, where, drop):
    j = json.load(open(GOLDEN))
    if where == "top":
        del j[drop]
    elif where == "road":
        del j["roads"][1][drop]
    else:
        del j["roads"][3]["start"][drop]
    p = tmp_path / "m.json"
    p.write_text(json.dumps(j))
    with pytest.raises(msim.MsimError, match=f"Failed to parse map. '{drop}' field missing."):
        host.HostMap(p)
    with pytest.raises(ValueError, match=f"'{drop}' field missing"):
        mapio.load_map(str(p))


def test_cpp_loader_missing_file(tmp_path):
    with pytest.raises(msim.MsimError, match="cannot open"):
        host.HostMap(tmp_path / "absent.json")


def test_cpp_loader_rejects_malformed_json(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"maxDistLat": 100, "roads": [')
    with pytest.raises(msim.MsimError):
        host.HostMap(p)


def test_guard_entry_like_mapio():
    m = host.HostMap(GOLDEN)
    roads, conns = m.arrays()
    m.add_guard()
    _, guarded = m.arrays()
    np.testing.assert_array_equal(guarded, mapio.with_guard(conns, roads))
    assert guarded.shape[0] == conns.shape[0] + 1 and guarded[-1] == 0
    m.add_guard()  # already guarded: unchanged
    np.testing.assert_array_equal(m.arrays()[1], guarded)


def test_entities_follow_add_entities():
    m = host.HostMap(GOLDEN)
    m.add_guard()
    roads, _ = m.arrays()
    e = m.make_entities(20000, seed=7)
    r = e["roadIndex"]
    assert r.max() < roads.shape[0]
    np.testing.assert_array_equal(e["pos"], roads["start_pos"][r])
    np.testing.assert_array_equal(e["target"], roads["end_pos"][r])
    assert (e["direction"] == 0).all() and (e["initialized"] == 0).all()
    assert (e["color"][:, 3] == 1.0).all() and (e["color"][:, :3] >= 0).all() and (e["color"][:, :3] < 1).all()
    assert (e["randomState"] != 0).any(axis=1).all()
    # uniform road choice: every road within 5 sigma of n/roads
    cnt = np.bincount(r, minlength=roads.shape[0])
    exp = e.shape[0] / roads.shape[0]
    assert (np.abs(cnt - exp) < 5 * np.sqrt(exp)).all(), cnt
    # deterministic in the seed
    np.testing.assert_array_equal(m.make_entities(20000, seed=7).view(np.uint8), e.view(np.uint8))
    assert not np.array_equal(m.make_entities(20000, seed=8)["randomState"], e["randomState"])


@pytest.mark.parametrize("period_ms", [7, 10, 300, 999])
def test_tick_rate_of_a_constant_period_is_its_inverse(period_ms):
    # the rate is the tick count over the first full second scaled to exactly
    # one second, with the excess carried over: for a constant period p it is
    # 1/p at every update
    r = host.TickRate()
    p = period_ms * 1_000_000
    seen = []
    for k in range(1, 4000):
        r.tick_at(k * p)
        if k * p >= 1_000_000_000:
            seen.append(r.get_ticks())
    assert seen and np.allclose(seen, 1e9 / p, rtol=1e-9)


def test_tick_rate_waits_for_a_full_second():
    r = host.TickRate()
    for k in range(1, 99):
        r.tick_at(k * 10_000_000)
    assert r.get_ticks() == 0.0
    r.tick_at(99 * 10_000_000)
    r.tick_at(100 * 10_000_000)
    assert r.get_ticks() == pytest.approx(100.0)


def test_tick_rate_follows_a_rate_change():
    r = host.TickRate()
    t = 0
    for _ in range(200):  # 2 s at 100 per second
        t += 10_000_000
        r.tick_at(t)
    for _ in range(3000):  # then 3 s at 1000 per second
        t += 1_000_000
        r.tick_at(t)
    assert r.get_ticks() == pytest.approx(1000.0, rel=1e-6)


def test_tick_history_mean_and_ring():
    h = host.TickHistory()
    assert h.get_avg_time() == 0 and h.get_avg_time_str() == "0.00ns"
    for v in range(1, 101):
        h.add_time(v)
    assert h.get_avg_time() == 50  # 5050 // 100
    for v in range(101, 151):
        h.add_time(v)
    assert h.get_avg_time() == sum(range(51, 151)) // 100  # the last 100 only


@pytest.mark.parametrize("ns,text", [(999, "999.00ns"), (1000, "1.00us"), (1500, "1.50us"), (999_999, "1000.00us"),
                                     (2_500_000, "2.50ms"), (3_000_000_000, "3.00s"),
                                     (12_345_678_901, "12.35s")])
def test_tick_history_string_units(ns, text):
    h = host.TickHistory()
    h.add_time(ns)
    assert h.get_avg_time_str() == text


@pytest.mark.parametrize("ms,text", [(0, "00:0:00.000"), (3_723_004, "01:2:03.004"),
                                     (10 * 3_600_000 + 59 * 60_000 + 5_007, "10:59:05.007"),
                                     (23 * 3_600_000 + 7 * 60_000 + 59_999, "23:7:59.999")])
def test_time_stamp_pads_all_but_minutes(ms, text):
    # Simulator::get_time_stamp pads hours and seconds to 2 and milliseconds
    # to 3 digits; the minute loop tests the millisecond string, so minutes
    # are never padded
    assert host.format_time_of_day(ms) == text


def test_headless_rejects_bad_arguments():
    assert os.path.exists(HEADLESS), "build the library first"
    for args in ([], ["--agents", "10"], ["--map"], ["--map", GOLDEN, "--bogus"]):
        p = subprocess.run([HEADLESS] + args, capture_output=True, text=True, timeout=60)
        assert p.returncode == 2 and "usage" in p.stderr
