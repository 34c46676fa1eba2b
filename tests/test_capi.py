"""C-ABI checks that need no GPU: the library loads, exports every symbol the
public headers declare, and rejects malformed input before touching a device."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import msim
from msim import _lib
from msim import scenarios as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "msim", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names.update(re.findall(r"\b(msim_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    lib = msim.load_library()
    declared = _declared_symbols()
    assert "msim_create" in declared and "msim_run_ticks" in declared
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # the binding declares a signature for each of them
    assert declared <= set(_lib.SIGNATURES), declared - set(_lib.SIGNATURES)


def test_version_string():
    v = msim.load_library().msim_version().decode()
    assert "gfx950" in v and "fp-contract=off" in v


def _create(roads, conns, ents, pc, speed=None):
    lib = msim.load_library()
    h = ctypes.c_void_p()
    rc = lib.msim_create(_lib._ptr(roads), roads.shape[0], _lib._ptr(conns), conns.shape[0], _lib._ptr(ents),
                         ents.shape[0], _lib._ptr(speed), _lib._ptr(pc), None, ctypes.byref(h))
    return rc, lib.msim_get_last_error().decode(), h


def test_rejects_connection_list_without_guard():
    # The reference's test_map.json layout: the last centre group ends at the
    # last entry, so the draw's top offset would read one past the buffer.
    roads, conns, w, h = S.star_map()
    ents = S.reference_agents(roads, 8, 1)
    pc = msim.push_consts(w, h, 10.0)
    rc, msg, _ = _create(roads, conns[:-1].copy(), ents, pc)
    assert rc == 1 and "guard" in msg


def test_rejects_bad_road_index_and_nan():
    roads, conns, w, h = S.star_map()
    pc = msim.push_consts(w, h, 10.0)
    ents = S.reference_agents(roads, 8, 1)
    ents["roadIndex"][3] = 4
    rc, msg, _ = _create(roads, conns, ents, pc)
    assert rc == 1 and "agent 3" in msg
    ents = S.reference_agents(roads, 8, 1)
    ents["pos"][5, 0] = np.nan
    rc, msg, _ = _create(roads, conns, ents, pc)
    assert rc == 1 and "agent 5" in msg
    bad = conns.copy()
    bad[2] = 9
    rc, msg, _ = _create(roads, bad, S.reference_agents(roads, 8, 1), pc)
    assert rc == 1 and "connections[2]" in msg


def test_rejects_bad_radius():
    roads, conns, w, h = S.star_map()
    ents = S.reference_agents(roads, 8, 1)
    for r in (0.0, -1.0, float("inf"), float("nan")):
        rc, msg, _ = _create(roads, conns, ents, msim.push_consts(w, h, r))
        assert rc == 1 and "collisionRadius" in msg


def test_simulation_raises_on_invalid():
    roads, conns, w, h = S.star_map()
    with pytest.raises(msim.MsimError):
        msim.Simulation(roads, conns[:-1], S.reference_agents(roads, 4, 2), msim.push_consts(w, h, 10.0))


def _batch_create(scs, pc, opts=None):
    lib = msim.load_library()
    offs = [np.concatenate([[0], np.cumsum([getattr(s, f).shape[0] for s in scs])]).astype(np.uint32)
            for f in ("roads", "conns", "entities")]
    r = np.concatenate([s.roads for s in scs])
    c = np.concatenate([s.conns for s in scs]).astype(np.uint32)
    e = np.concatenate([s.entities for s in scs])
    h = ctypes.c_void_p()
    rc = lib.msim_batch_create(len(scs), _lib._ptr(offs[0]), _lib._ptr(r), _lib._ptr(offs[1]), _lib._ptr(c),
                               _lib._ptr(offs[2]), _lib._ptr(e), None, _lib._ptr(pc),
                               None if opts is None else ctypes.byref(opts), ctypes.byref(h))
    return rc, lib.msim_batch_last_error(None).decode()


def _star_scenarios(k):
    roads, conns, w, h = S.star_map()
    return [S.Scenario(f"s{i}", roads, conns.copy(), S.reference_agents(roads, 8 + i, i), None,
                       msim.push_consts(w, h, 10.0), 0, (w, h)) for i in range(k)], msim.push_consts(w, h, 10.0)


def test_batch_rejects_malformed_scenarios():
    scs, pc = _star_scenarios(4)
    scs[2].conns = scs[2].conns[:-1].copy()  # no guard entry in scenario 2 only
    rc, msg = _batch_create(scs, pc)
    assert rc == 1 and "scenario 2" in msg and "guard" in msg
    scs, pc = _star_scenarios(3)
    scs[1].entities["roadIndex"][0] = 4  # local index past scenario 1's roads
    rc, msg = _batch_create(scs, pc)
    assert rc == 1 and "scenario 1" in msg


def test_batch_rejects_mixed_init_and_even_tick():
    scs, pc = _star_scenarios(3)
    scs[2].entities["initialized"][3] = 1
    rc, msg = _batch_create(scs, pc)
    assert rc == 1 and "initialized" in msg
    scs, pc = _star_scenarios(2)
    pc["tick"] = 2
    rc, msg = _batch_create(scs, pc)
    assert rc == 1 and "tick" in msg


def test_batch_rejects_atomic_binning():
    scs, pc = _star_scenarios(2)
    rc, msg = _batch_create(scs, pc, _lib.make_options(binning=1))
    assert rc == 1 and "binning" in msg


def test_batch_binding_rejects_mixed_speed_modes():
    sa = S.batch_config(2, scale=0.01)
    sa[1].speed = None
    with pytest.raises(ValueError):
        msim.Batch(sa)
