"""Python handles on the C++ host driver (include/msim/msim_host.h).

The reference's host side is C++ (sim::Map, sim::Simulator, utils::TickRate,
utils::TickDurationHistory); its mirror lives in movement-sim_amd/host/ and is
compiled into libmsim.so.  These classes are thin ctypes wrappers for tests
and Python callers:

* HostMap: Map::load_from_file (src/sim/Map.cpp:28-150) in C++;
* TickRate / TickHistory: src/utils/TickRate.cpp, TickDurationHistory.cpp;
* HostSimulator: sim::Simulator (src/sim/Simulator.cpp) with its worker
  thread, continue/pause, the per-tick CSV log and the entity hand-over.
"""
import ctypes

import numpy as np

from ._lib import MsimError, load_library
from .types import ENTITY_DTYPE, ROAD_DTYPE


def _check(lib, rc):
    if rc != 0:
        raise MsimError(rc, lib.msim_host_last_error().decode())


class HostMap:
    """A map loaded by the C++ loader."""

    def __init__(self, path):
        self.lib = load_library()
        self.h = ctypes.c_void_p()
        _check(self.lib, self.lib.msim_host_map_load(str(path).encode(), ctypes.byref(self.h)))

    def info(self):
        w, h = ctypes.c_float(), ctypes.c_float()
        nr, nc, sk = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(self.lib, self.lib.msim_host_map_info(self.h, ctypes.byref(w), ctypes.byref(h), ctypes.byref(nr),
                                                     ctypes.byref(nc), ctypes.byref(sk)))
        return {"width": w.value, "height": h.value, "n_roads": nr.value, "n_connections": nc.value,
                "skipped": sk.value}

    def arrays(self):
        i = self.info()
        roads = np.zeros(i["n_roads"], ROAD_DTYPE)
        conns = np.zeros(i["n_connections"], np.uint32)
        _check(self.lib, self.lib.msim_host_map_read(self.h, roads.ctypes.data_as(ctypes.c_void_p),
                                                     conns.ctypes.data_as(ctypes.c_void_p)))
        return roads, conns

    def add_guard(self):
        _check(self.lib, self.lib.msim_host_map_add_guard(self.h))

    def make_entities(self, n, seed):
        out = np.zeros(n, ENTITY_DTYPE)
        _check(self.lib, self.lib.msim_host_make_entities(self.h, n, seed, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def close(self):
        if self.h:
            self.lib.msim_host_map_free(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TickRate:
    """utils::TickRate fed explicit tick times (ns from an origin of 0)."""

    def __init__(self):
        self.lib = load_library()
        self.h = ctypes.c_void_p(self.lib.msim_tick_rate_new())

    def tick_at(self, ns):
        self.lib.msim_tick_rate_tick_at(self.h, int(ns))

    def get_ticks(self):
        return self.lib.msim_tick_rate_get(self.h)

    def __del__(self):
        try:
            self.lib.msim_tick_rate_free(self.h)
        except Exception:
            pass


class TickHistory:
    """utils::TickDurationHistory."""

    def __init__(self):
        self.lib = load_library()
        self.h = ctypes.c_void_p(self.lib.msim_tick_history_new())

    def add_time(self, ns):
        self.lib.msim_tick_history_add(self.h, int(ns))

    def get_avg_time(self):
        return self.lib.msim_tick_history_avg_ns(self.h)

    def get_avg_time_str(self):
        buf = ctypes.create_string_buffer(64)
        _check(self.lib, self.lib.msim_tick_history_avg_str(self.h, buf, 64))
        return buf.value.decode()

    def __del__(self):
        try:
            self.lib.msim_tick_history_free(self.h)
        except Exception:
            pass


def format_time_of_day(ms_since_midnight):
    lib = load_library()
    buf = ctypes.create_string_buffer(64)
    _check(lib, lib.msim_host_format_time_of_day(int(ms_since_midnight), buf, 64))
    return buf.value.decode()


class HostSimulator:
    """sim::Simulator over the HIP pipeline (worker thread, CSV tick log)."""

    def __init__(self, map_path, agents, collision_radius=10.0, seed=1, csv_path=None, log_stderr=False,
                 tick_limit=0):
        self.lib = load_library()
        self.n = agents
        self.h = ctypes.c_void_p()
        _check(self.lib, self.lib.msim_host_sim_create(str(map_path).encode(), agents, collision_radius, seed,
                                                       None if csv_path is None else str(csv_path).encode(),
                                                       int(log_stderr), tick_limit, ctypes.byref(self.h)))

    def start_worker(self):
        _check(self.lib, self.lib.msim_host_sim_start_worker(self.h))

    def stop_worker(self):
        _check(self.lib, self.lib.msim_host_sim_stop_worker(self.h))

    def continue_simulation(self):
        _check(self.lib, self.lib.msim_host_sim_continue(self.h))

    def pause_simulation(self):
        _check(self.lib, self.lib.msim_host_sim_pause(self.h))

    def is_simulating(self):
        return bool(self.lib.msim_host_sim_is_simulating(self.h))

    def wait_ticks(self, n, timeout_ms=120_000):
        _check(self.lib, self.lib.msim_host_sim_wait_ticks(self.h, n, timeout_ms))

    def stats(self):
        t, tps = ctypes.c_uint64(), ctypes.c_double()
        a, u, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(self.lib, self.lib.msim_host_sim_stats(self.h, ctypes.byref(t), ctypes.byref(tps), ctypes.byref(a),
                                                      ctypes.byref(u), ctypes.byref(c)))
        return {"ticks": t.value, "tps": tps.value, "tick_ns": a.value, "update_ns": u.value,
                "collision_ns": c.value}

    def take_entities(self):
        """The worker's last entity copy, or None when it was taken already."""
        out = np.zeros(self.n, ENTITY_DTYPE)
        taken = ctypes.c_int()
        _check(self.lib, self.lib.msim_host_sim_take_entities(self.h, out.ctypes.data_as(ctypes.c_void_p), self.n,
                                                              ctypes.byref(taken)))
        return out if taken.value else None

    def handle(self):
        return self.lib.msim_host_sim_handle(self.h)

    def close(self):
        if self.h:
            self.lib.msim_host_sim_destroy(self.h)
            self.h = ctypes.c_void_p()
