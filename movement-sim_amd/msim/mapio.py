"""Maps in the reference's JSON format (src/sim/Map.cpp:28-150,
test_map.json): load into the Road / connection arrays msim_create takes,
and write generated maps back out.

Format: {"maxDistLat": W, "maxDistLong": H, "roads": [{"start": {"distLat", "distLong"},
"end": {...}, "connIndexStart", "connCountStart", "connIndexEnd", "connCountEnd"}, ...],
"connectionRoadIndexList": [...]}.  The loader keeps the reference's rules:
every field is required (the same messages as its runtime_errors), coordinates
are binary32, zero-length roads are skipped (later road indices shift down and
the connection list is not renumbered, as in the reference), a missing file
gives None.
"""
import json
import logging
import os
from dataclasses import dataclass

import numpy as np

from .types import ROAD_DTYPE

log = logging.getLogger("msim.mapio")


@dataclass
class MapData:
    width: float
    height: float
    roads: np.ndarray        # ROAD_DTYPE
    conns: np.ndarray        # uint32
    skipped: int = 0         # zero-length roads dropped


def _field(obj, name):
    if not isinstance(obj, dict) or name not in obj:
        raise ValueError(f"Failed to parse map. '{name}' field missing.")
    return obj[name]


def _unsigned(v, name):
    if isinstance(v, bool) or not isinstance(v, int) or v < 0 or v >= 2 ** 32:
        raise ValueError(f"Failed to parse map. '{name}' is not an unsigned 32-bit integer.")
    return v


def load_map(path):
    """Map::load_from_file: MapData, or None when the file does not exist."""
    if not os.path.exists(path):
        log.error("Failed to open map from '%s'. File does not exist.", path)
        return None
    with open(path) as f:
        j = json.load(f)
    width = np.float32(_field(j, "maxDistLat"))
    height = np.float32(_field(j, "maxDistLong"))
    rows = []
    skipped = 0
    for jr in _field(j, "roads"):
        i_s = _unsigned(_field(jr, "connIndexStart"), "connIndexStart")
        c_s = _unsigned(_field(jr, "connCountStart"), "connCountStart")
        js = _field(jr, "start")
        sx = np.float32(_field(js, "distLat"))
        sy = np.float32(_field(js, "distLong"))
        i_e = _unsigned(_field(jr, "connIndexEnd"), "connIndexEnd")
        c_e = _unsigned(_field(jr, "connCountEnd"), "connCountEnd")
        je = _field(jr, "end")
        ex = np.float32(_field(je, "distLat"))
        ey = np.float32(_field(je, "distLong"))
        if sx == ex and sy == ey:  # Map.cpp:124-128
            log.warning("Zero length road detected. Skipping...")
            skipped += 1
            continue
        rows.append(((sx, sy), i_s, c_s, (ex, ey), i_e, c_e))
    roads = np.array(rows, ROAD_DTYPE) if rows else np.zeros(0, ROAD_DTYPE)
    conns = np.array([_unsigned(c, "connectionRoadIndexList") for c in _field(j, "connectionRoadIndexList")],
                     np.uint32)
    log.info("Map loaded from '%s'. Found %d roads with %d connections.", path, roads.shape[0], conns.shape[0])
    return MapData(float(width), float(height), roads, conns, skipped)


def _num(v):
    """Shortest decimal that reads back as the same binary32."""
    return float(np.float32(v))


def save_map(path, width, height, roads, conns):
    """Writes roads / connections in the reference's format (lat/long 0)."""
    out = {
        "minDistLat": 0,
        "maxDistLat": _num(width),
        "minDistLong": 0,
        "maxDistLong": _num(height),
        "roads": [
            {
                "start": {"lat": 0, "long": 0, "distLat": _num(r["start_pos"][0]), "distLong": _num(r["start_pos"][1])},
                "end": {"lat": 0, "long": 0, "distLat": _num(r["end_pos"][0]), "distLong": _num(r["end_pos"][1])},
                "connIndexStart": int(r["start_index"]),
                "connCountStart": int(r["start_count"]),
                "connIndexEnd": int(r["end_index"]),
                "connCountEnd": int(r["end_count"]),
            }
            for r in np.asarray(roads, ROAD_DTYPE)
        ],
        "connectionRoadIndexList": [int(c) for c in np.asarray(conns, np.uint32)],
    }
    with open(path, "w") as f:
        json.dump(out, f)


def with_guard(conns, roads):
    """The connection list with one guard entry (road 0) appended when some
    group's top draw would read one past its end (random_move.comp:808 can
    return offset == count), which msim_create otherwise rejects."""
    conns = np.asarray(conns, np.uint32)
    r = np.asarray(roads, ROAD_DTYPE)
    top = 0
    for idx, cnt in ((r["start_index"], r["start_count"]), (r["end_index"], r["end_count"])):
        idx = idx.astype(np.int64)
        cnt = cnt.astype(np.int64)
        reach = np.where(cnt >= 3, idx + cnt, np.where(cnt == 2, idx + 1, -1))
        top = max(top, int(reach.max()) if reach.size else -1)
    if top >= conns.shape[0]:
        pad = np.zeros(top + 1 - conns.shape[0], np.uint32)
        conns = np.concatenate([conns, pad])
    return conns
