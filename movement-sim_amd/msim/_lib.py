"""ctypes binding of libmsim.so, the C-ABI declared in include/msim/msim.h.

The library is the only compute path: if it has not been built, loading
raises instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

from .types import ENTITY_DTYPE, PUSH_DTYPE, ROAD_DTYPE

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libmsim.so")
# the same library built with -ffp-contract=fast (integer-deciding arithmetic
# stays single-rounding; positions may differ from LIB_PATH's by fused roundings)
LIB_CONTRACT_PATH = os.path.join(PKG_ROOT, "lib", "libmsim_contract.so")

MSIM_OK = 0
STATUS_NAMES = {1: "MSIM_ERR_INVALID", 2: "MSIM_ERR_HIP", 3: "MSIM_ERR_OOM", 4: "MSIM_ERR_STATE",
                5: "MSIM_ERR_TIMEOUT"}


class MsimError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("cell_size", ctypes.c_float),
        ("table_log2", ctypes.c_uint32),
        ("block_size", ctypes.c_uint32),
        ("reorder_interval", ctypes.c_uint32),
        ("graph_ticks", ctypes.c_uint32),
        ("history_capacity", ctypes.c_uint32),
        ("device", ctypes.c_int32),
        ("watchdog_ms", ctypes.c_uint32),
        ("table_width_log2", ctypes.c_uint32),
        ("segment_log2", ctypes.c_uint32),
        ("binning", ctypes.c_uint32),
        ("route_queue", ctypes.c_uint32),
        ("collide_mode", ctypes.c_uint32),
        ("bin_unroll", ctypes.c_uint32),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("n_agents", ctypes.c_uint32),
        ("n_roads", ctypes.c_uint32),
        ("n_connections", ctypes.c_uint32),
        ("tick", ctypes.c_uint32),
        ("ticks_simulated", ctypes.c_uint64),
        ("cell_size", ctypes.c_float),
        ("collision_radius", ctypes.c_float),
        ("table_width", ctypes.c_uint32),
        ("table_height", ctypes.c_uint32),
        ("uniform_speed", ctypes.c_uint32),
        ("n_uninitialized", ctypes.c_uint32),
        ("device_bytes", ctypes.c_uint64),
        ("kernels_per_tick", ctypes.c_uint32),
        ("binning", ctypes.c_uint32),
    ]


# Every symbol include/msim/msim.h and include/msim/msim_host.h declare, with
# its ctypes signature.
_VP = ctypes.c_void_p
_U32 = ctypes.c_uint32
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
SIGNATURES = {
    "msim_create": (ctypes.c_int, [_VP, _U32, _VP, _U32, _VP, _U32, _VP, _VP, _VP, ctypes.POINTER(_VP)]),
    "msim_dispatch": (ctypes.c_int, [_VP, _VP]),
    "msim_run_ticks": (ctypes.c_int, [_VP, _U32]),
    "msim_sync": (ctypes.c_int, [_VP]),
    "msim_read_entities": (ctypes.c_int, [_VP, _VP, _U32, _U32]),
    "msim_read_debug": (ctypes.c_int, [_VP, _VP]),
    "msim_read_state": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "msim_read_pairs": (ctypes.c_int, [_VP, _VP, _VP, _U32, _VP]),
    "msim_read_cells": (ctypes.c_int, [_VP, _VP]),
    "msim_read_cells_noalloc": (ctypes.c_int, [_VP, _VP]),
    "msim_read_grid_occupancy": (ctypes.c_int, [_VP, _VP, _VP, _U32, _VP]),
    "msim_get_info": (ctypes.c_int, [_VP, _VP]),
    "msim_time_ticks": (ctypes.c_int, [_VP, _U32, _VP, _VP, _U32]),
    "msim_kernel_name": (ctypes.c_char_p, [_VP, _U32]),
    "msim_get_stream": (ctypes.c_int, [_VP, ctypes.POINTER(_VP)]),
    "msim_destroy": (ctypes.c_int, [_VP]),
    "msim_last_error": (ctypes.c_char_p, [_VP]),
    "msim_get_last_error": (ctypes.c_char_p, []),
    "msim_version": (ctypes.c_char_p, []),
    "msim_batch_create": (ctypes.c_int, [_U32, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP,
                                         ctypes.POINTER(_VP)]),
    "msim_batch_run_ticks": (ctypes.c_int, [_VP, _U32]),
    "msim_batch_sync": (ctypes.c_int, [_VP]),
    "msim_batch_read_state": (ctypes.c_int, [_VP, _U32, _VP, _VP, _VP, _VP, _VP, _VP]),
    "msim_batch_read_pairs": (ctypes.c_int, [_VP, _VP, _VP]),
    "msim_batch_time_ticks": (ctypes.c_int, [_VP, _U32, _VP, _VP, _U32]),
    "msim_batch_kernel_name": (ctypes.c_char_p, [_VP, _U32]),
    "msim_batch_destroy": (ctypes.c_int, [_VP]),
    "msim_batch_last_error": (ctypes.c_char_p, [_VP]),
    "msim_batch_diff_setup": (ctypes.c_int, [_VP, _U32, _U32, ctypes.c_int]),
    "msim_batch_diff_forward": (ctypes.c_int, [_VP, _VP, ctypes.c_float, ctypes.c_int, _VP]),
    "msim_batch_diff_backward": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
    "msim_batch_diff_read_schedule": (ctypes.c_int, [_VP, _U32, _VP, _VP, _VP]),
    "msim_batch_diff_set_arrival_gradient": (ctypes.c_int, [_VP, ctypes.c_int]),
    "msim_batch_diff_generation": (ctypes.c_int, [_VP, _VP]),
    # host driver (msim_host.h)
    "msim_host_map_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_VP)]),
    "msim_host_map_info": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "msim_host_map_read": (ctypes.c_int, [_VP, _VP, _VP]),
    "msim_host_map_add_guard": (ctypes.c_int, [_VP]),
    "msim_host_map_free": (None, [_VP]),
    "msim_host_make_entities": (ctypes.c_int, [_VP, _U32, _U64, _VP]),
    "msim_tick_rate_new": (_VP, []),
    "msim_tick_rate_tick_at": (None, [_VP, _I64]),
    "msim_tick_rate_get": (ctypes.c_double, [_VP]),
    "msim_tick_rate_free": (None, [_VP]),
    "msim_tick_history_new": (_VP, []),
    "msim_tick_history_add": (None, [_VP, _I64]),
    "msim_tick_history_avg_ns": (_I64, [_VP]),
    "msim_tick_history_avg_str": (ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.c_size_t]),
    "msim_tick_history_free": (None, [_VP]),
    "msim_host_format_time_of_day": (ctypes.c_int, [_U64, ctypes.c_char_p, ctypes.c_size_t]),
    "msim_host_sim_create": (ctypes.c_int, [ctypes.c_char_p, _U32, ctypes.c_float, _U64, ctypes.c_char_p,
                                            ctypes.c_int, _U64, ctypes.POINTER(_VP)]),
    "msim_host_sim_start_worker": (ctypes.c_int, [_VP]),
    "msim_host_sim_stop_worker": (ctypes.c_int, [_VP]),
    "msim_host_sim_continue": (ctypes.c_int, [_VP]),
    "msim_host_sim_pause": (ctypes.c_int, [_VP]),
    "msim_host_sim_is_simulating": (ctypes.c_int, [_VP]),
    "msim_host_sim_wait_ticks": (ctypes.c_int, [_VP, _U64, _U32]),
    "msim_host_sim_stats": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "msim_host_sim_take_entities": (ctypes.c_int, [_VP, _VP, _U32, _VP]),
    "msim_host_sim_handle": (_VP, [_VP]),
    "msim_host_sim_destroy": (ctypes.c_int, [_VP]),
    "msim_host_last_error": (ctypes.c_char_p, []),
}

_lib = None


def load_library(path=None):
    """Load libmsim.so once; raise if it is missing (there is no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is not built: run __graft_entry__.build() (or make -C movement-sim_amd/csrc) first")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def make_options(**kw):
    o = Options()
    o.struct_size = ctypes.sizeof(Options)
    o.device = -1
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Simulation:
    """One handle of the plain pipeline (the HIP drop-in for the shader tick).

    roads: ROAD_DTYPE array, connections: uint32 array, entities: ENTITY_DTYPE
    array, speed: float32 array or None (constant 1.4 m/tick), pc: PUSH_DTYPE.
    """

    def __init__(self, roads, connections, entities, pc, speed=None, library=None, **options):
        self.lib = library if library is not None else load_library()
        roads = np.ascontiguousarray(roads, ROAD_DTYPE)
        conns = np.ascontiguousarray(connections, np.uint32)
        ents = np.ascontiguousarray(entities, ENTITY_DTYPE)
        pcv = np.array(pc, PUSH_DTYPE).reshape(())  # 0-d record
        spd = None if speed is None else np.ascontiguousarray(speed, np.float32)
        if spd is not None and spd.shape[0] != ents.shape[0]:
            raise ValueError("speed must have one entry per agent")
        opts = make_options(**options) if options else None
        h = ctypes.c_void_p()
        rc = self.lib.msim_create(_ptr(roads), roads.shape[0], _ptr(conns), conns.shape[0], _ptr(ents),
                                  ents.shape[0], _ptr(spd), _ptr(pcv),
                                  None if opts is None else ctypes.byref(opts), ctypes.byref(h))
        if rc != MSIM_OK:
            raise MsimError(rc, self.lib.msim_get_last_error().decode())
        self.h = h
        self.n = ents.shape[0]
        self.pc = pcv.copy()

    def _check(self, rc):
        if rc != MSIM_OK:
            raise MsimError(rc, self.lib.msim_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.msim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- driving -----------------------------------------------------------
    def dispatch(self, tick=None):
        """One shader dispatch (even tick = move, odd = collision)."""
        if tick is None:
            tick = int(self.pc["tick"]) + 1
        self.pc["tick"] = tick
        self._check(self.lib.msim_dispatch(self.h, _ptr(self.pc)))

    def sim_tick(self):
        """Simulator::sim_tick: tick+=1 -> dispatch, tick+=1 -> dispatch."""
        self.dispatch()
        self.dispatch()

    def run_ticks(self, n):
        self._check(self.lib.msim_run_ticks(self.h, n))
        self.pc["tick"] = self.info()["tick"]

    def sync(self):
        self._check(self.lib.msim_sync(self.h))

    # -- readback ----------------------------------------------------------
    def entities(self, first=0, count=None):
        count = self.n - first if count is None else count
        out = np.zeros(count, ENTITY_DTYPE)
        self._check(self.lib.msim_read_entities(self.h, _ptr(out), first, count))
        return out

    def state(self):
        n = self.n
        st = {
            "pos": np.zeros((n, 2), np.float32),
            "tgt": np.zeros((n, 2), np.float32),
            "dir": np.zeros((n, 2), np.float32),
            "road": np.zeros(n, np.uint32),
            "rng": np.zeros((n, 4), np.uint32),
            "ncount": np.zeros(n, np.uint32),
        }
        self._check(self.lib.msim_read_state(self.h, _ptr(st["pos"]), _ptr(st["tgt"]), _ptr(st["dir"]),
                                             _ptr(st["road"]), _ptr(st["rng"]), _ptr(st["ncount"])))
        return st

    def debug_data(self):
        out = np.zeros(10, np.uint32)
        self._check(self.lib.msim_read_debug(self.h, _ptr(out)))
        return out

    def pairs(self, n=4096):
        total = np.zeros(1, np.uint64)
        per = np.zeros(n, np.uint32)
        got = ctypes.c_uint32()
        self._check(self.lib.msim_read_pairs(self.h, _ptr(total), _ptr(per), n, ctypes.byref(got)))
        return int(total[0]), per[: got.value].copy()

    def cells(self):
        out = np.zeros((self.n, 2), np.int32)
        self._check(self.lib.msim_read_cells(self.h, _ptr(out)))
        return out

    def grid_occupancy(self, capacity=None):
        """(cells [k, 2] int32, counts [k] uint32, number of occupied cells):
        the occupied cells of the last binning, in no particular order."""
        cap = self.n if capacity is None else capacity
        cells = np.zeros((max(cap, 1), 2), np.int32)
        counts = np.zeros(max(cap, 1), np.uint32)
        total = ctypes.c_uint32()
        self._check(self.lib.msim_read_grid_occupancy(self.h, _ptr(cells), _ptr(counts), cap, ctypes.byref(total)))
        k = min(total.value, cap)
        return cells[:k].copy(), counts[:k].copy(), total.value

    def cells_noalloc(self):
        """cells() through the allocation-free entry point."""
        out = np.zeros((self.n, 2), np.int32)
        self._check(self.lib.msim_read_cells_noalloc(self.h, _ptr(out)))
        return out

    def info(self):
        i = Info()
        self._check(self.lib.msim_get_info(self.h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in Info._fields_}

    def kernel_names(self):
        out = []
        while True:
            nm = self.lib.msim_kernel_name(self.h, len(out))
            if not nm:
                return out
            out.append(nm.decode())

    def time_ticks(self, n, kernels=True):
        """(ms for n ticks through the graph path, {kernel: summed ms over a
        second pass of n ticks with events around every launch})."""
        total = np.zeros(1, np.float32)
        names = self.kernel_names()
        k = np.zeros(len(names), np.float32)
        self._check(self.lib.msim_time_ticks(self.h, n, _ptr(total), _ptr(k) if kernels else None, len(names)))
        return float(total[0]), dict(zip(names, (float(v) for v in k)))

    def stream(self):
        s = ctypes.c_void_p()
        self._check(self.lib.msim_get_stream(self.h, ctypes.byref(s)))
        return s.value


class Batch:
    """Many independent scenarios advanced together (msim_batch_*).

    scenarios: sequence of objects with .roads, .conns, .entities, .speed
    (float32 array or None) and one shared .pc (collisionRadius and tick must
    agree).  Either every scenario has per-agent speeds or none has.
    """

    def __init__(self, scenarios, pc=None, **options):
        self.lib = load_library()
        scs = list(scenarios)
        if not scs:
            raise ValueError("no scenarios")
        pcv = np.array(scs[0].pc if pc is None else pc, PUSH_DTYPE).reshape(())
        speeds = [sc.speed for sc in scs]
        if any(v is None for v in speeds) and not all(v is None for v in speeds):
            raise ValueError("either every scenario has per-agent speeds or none has")
        roads = [np.ascontiguousarray(sc.roads, ROAD_DTYPE) for sc in scs]
        conns = [np.ascontiguousarray(sc.conns, np.uint32) for sc in scs]
        ents = [np.ascontiguousarray(sc.entities, ENTITY_DTYPE) for sc in scs]

        def offsets(parts):
            return np.concatenate([[0], np.cumsum([p.shape[0] for p in parts])]).astype(np.uint32)

        self.road_off, self.conn_off, self.agent_off = offsets(roads), offsets(conns), offsets(ents)
        r = np.concatenate(roads)
        c = np.concatenate(conns)
        e = np.concatenate(ents)
        spd = None if speeds[0] is None else np.ascontiguousarray(np.concatenate(speeds), np.float32)
        opts = make_options(**options) if options else None
        h = ctypes.c_void_p()
        rc = self.lib.msim_batch_create(len(scs), _ptr(self.road_off), _ptr(r), _ptr(self.conn_off), _ptr(c),
                                        _ptr(self.agent_off), _ptr(e), _ptr(spd), _ptr(pcv),
                                        None if opts is None else ctypes.byref(opts), ctypes.byref(h))
        if rc != MSIM_OK:
            raise MsimError(rc, self.lib.msim_batch_last_error(None).decode())
        self.h = h
        self.n_scen = len(scs)
        self.n = int(self.agent_off[-1])
        self.n_roads = int(self.road_off[-1])
        self.diff_ticks = 0

    def _check(self, rc):
        if rc != MSIM_OK:
            raise MsimError(rc, self.lib.msim_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.msim_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run_ticks(self, n):
        self._check(self.lib.msim_batch_run_ticks(self.h, n))

    def sync(self):
        self._check(self.lib.msim_batch_sync(self.h))

    def state(self, s):
        """Scenario s's state in its own agent order and road numbering."""
        n = int(self.agent_off[s + 1] - self.agent_off[s])
        st = {
            "pos": np.zeros((n, 2), np.float32),
            "tgt": np.zeros((n, 2), np.float32),
            "dir": np.zeros((n, 2), np.float32),
            "road": np.zeros(n, np.uint32),
            "rng": np.zeros((n, 4), np.uint32),
            "ncount": np.zeros(n, np.uint32),
        }
        self._check(self.lib.msim_batch_read_state(self.h, s, _ptr(st["pos"]), _ptr(st["tgt"]), _ptr(st["dir"]),
                                                   _ptr(st["road"]), _ptr(st["rng"]), _ptr(st["ncount"])))
        return st

    def pairs(self):
        """(cumulative pairs per scenario, pairs of the last tick per scenario)."""
        tot = np.zeros(self.n_scen, np.uint64)
        last = np.zeros(self.n_scen, np.uint32)
        self._check(self.lib.msim_batch_read_pairs(self.h, _ptr(tot), _ptr(last)))
        return tot, last

    def kernel_names(self):
        out = []
        while True:
            nm = self.lib.msim_batch_kernel_name(self.h, len(out))
            if not nm:
                return out
            out.append(nm.decode())

    def time_ticks(self, n, kernels=False):
        """(ms for n ticks of every scenario, {kernel: summed ms over a second
        pass of n ticks with events around every launch} or {})."""
        ms = np.zeros(1, np.float32)
        names = self.kernel_names() if kernels else []
        k = np.zeros(max(1, len(names)), np.float32)
        self._check(self.lib.msim_batch_time_ticks(self.h, n, _ptr(ms), _ptr(k) if kernels else None, len(names)))
        return float(ms[0]), dict(zip(names, (float(v) for v in k)))

    # -- differentiable tick (device pointers: ints, e.g. torch data_ptr()) --
    def diff_setup(self, ticks, checkpoint_every=50, keep_schedule=False):
        self._check(self.lib.msim_batch_diff_setup(self.h, ticks, checkpoint_every, int(keep_schedule)))
        self.diff_ticks = ticks

    def diff_forward(self, theta_ptr, alpha, final_pos_ptr, replay=False):
        self._check(self.lib.msim_batch_diff_forward(self.h, ctypes.c_void_p(theta_ptr), float(alpha),
                                                     int(replay), ctypes.c_void_p(final_pos_ptr)))

    def diff_backward(self, grad_pos_ptr, grad_theta_ptr, grad_alpha_ptr):
        self._check(self.lib.msim_batch_diff_backward(self.h, ctypes.c_void_p(grad_pos_ptr),
                                                      ctypes.c_void_p(grad_theta_ptr),
                                                      ctypes.c_void_p(grad_alpha_ptr)))

    def diff_generation(self):
        """Number of recorded (non-replay) forward runs; diff_backward
        differentiates the latest."""
        g = np.zeros(1, np.uint64)
        self._check(self.lib.msim_batch_diff_generation(self.h, _ptr(g)))
        return int(g[0])

    def diff_set_arrival_gradient(self, mode):
        """0: arrivals cut the gradient (frozen run); 1: carry-over surrogate."""
        self._check(self.lib.msim_batch_diff_set_arrival_gradient(self.h, int(mode)))

    def diff_schedule(self, tick):
        """(targets [n, 2], packed roads [n], flags [n]) of one recorded tick;
        tick == ticks gives the decisions in force after the run."""
        t = np.zeros((self.n, 2), np.float32)
        r = np.zeros(self.n, np.uint32)
        f = np.zeros(self.n, np.uint8)
        self._check(self.lib.msim_batch_diff_read_schedule(self.h, tick, _ptr(t), _ptr(r), _ptr(f)))
        return t, r, f
