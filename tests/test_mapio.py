"""Maps in the reference's JSON format (msim.mapio), checked against the
reference's own fixture test_map.json (committed as tests/golden/test_map.json)
and the loader rules of src/sim/Map.cpp:28-150."""
import json
import os

import numpy as np
import pytest

from msim import mapio
from msim import scenarios as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "test_map.json")


def test_reference_fixture_is_the_star_map():
    m = mapio.load_map(GOLDEN)
    roads, conns, w, h = S.star_map()
    assert (m.width, m.height) == (w, h) and m.skipped == 0
    np.testing.assert_array_equal(m.roads.view(np.uint32), roads.view(np.uint32))
    # same groups: own road first, then the node's other roads (the fixture
    # lists those in its own order)
    for r in m.roads:
        for idx, cnt in ((r["start_index"], r["start_count"]), (r["end_index"], r["end_count"])):
            a, b = m.conns[idx:idx + cnt], conns[idx:idx + cnt]
            assert a[0] == b[0] and sorted(a) == sorted(b)
    # no guard entry in the fixture: its last centre group ends at the last entry
    assert m.conns.shape[0] == conns.shape[0] - 1
    g = mapio.with_guard(m.conns, m.roads)
    assert g.shape[0] == conns.shape[0] and g[-1] == 0


def test_zero_length_roads_are_skipped_without_renumbering(tmp_path):
    j = json.load(open(GOLDEN))
    zero = dict(j["roads"][0])
    zero["end"] = dict(zero["start"])
    j["roads"].insert(1, zero)
    p = tmp_path / "m.json"
    p.write_text(json.dumps(j))
    m = mapio.load_map(str(p))
    assert m.skipped == 1 and m.roads.shape[0] == 4
    assert m.conns.tolist() == j["connectionRoadIndexList"]


@pytest.mark.parametrize("drop", ["maxDistLat", "roads", "connectionRoadIndexList"])
def test_missing_top_level_field(tmp_path, drop):
    j = json.load(open(GOLDEN))
    del j[drop]
    p = tmp_path / "m.json"
    p.write_text(json.dumps(j))
    with pytest.raises(ValueError, match=f"'{drop}' field missing"):
        mapio.load_map(str(p))


@pytest.mark.parametrize("drop", ["connIndexStart", "connCountEnd", "start"])
def test_missing_road_field(tmp_path, drop):
    j = json.load(open(GOLDEN))
    del j["roads"][2][drop]
    p = tmp_path / "m.json"
    p.write_text(json.dumps(j))
    with pytest.raises(ValueError, match=f"'{drop}' field missing"):
        mapio.load_map(str(p))


def test_missing_file_gives_none(tmp_path):
    assert mapio.load_map(str(tmp_path / "absent.json")) is None


@pytest.mark.parametrize("layout", ["own_first", "node_csr"])
def test_round_trip_of_a_generated_map(tmp_path, layout):
    sc = S.config(2, scale=0.02, layout=layout)
    p = tmp_path / "city.json"
    mapio.save_map(str(p), *sc.world, sc.roads, sc.conns)
    m = mapio.load_map(str(p))
    np.testing.assert_array_equal(m.roads.view(np.uint32), sc.roads.view(np.uint32))
    np.testing.assert_array_equal(m.conns, sc.conns)
    assert (m.width, m.height) == tuple(float(np.float32(v)) for v in sc.world)
