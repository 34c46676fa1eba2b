"""ctypes wrapper of the CPU oracle (liboracle.so) and of the reference's own
CPU quadtree build (_ref/libref_quadtree.so, when it was built).

TEST INFRASTRUCTURE ONLY: tests, the smoke check and the benchmark's CPU
baseline use it as the checker; the product path never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_quadtree.so")

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32

ROAD_FIELDS = ("start_pos", "start_index", "start_count", "end_pos", "end_index", "end_count")


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _load():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} not built (make -C oracle)")
    lib = ctypes.CDLL(LIB)
    sig = {
        "orc_rng_next": (_u32, [_vp]),
        "orc_rng_float": (ctypes.c_float, [_vp]),
        "orc_rng_range": (_u32, [_vp, _u32, _u32]),
        "orc_length": (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
        "orc_update_direction": (None, [_vp, _vp, ctypes.c_float, _vp]),
        "orc_move_phase": (ctypes.c_int, [_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _u32, _vp]),
        "orc_move_phase_mt": (ctypes.c_int, [_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _u32,
                                             ctypes.c_int]),
        "orc_in_range": (ctypes.c_int, [_vp, _vp, ctypes.c_float]),
        "orc_cell_size": (ctypes.c_float, [ctypes.c_float]),
        "orc_cells": (None, [_u32, _vp, ctypes.c_float, _vp]),
        "orc_collide_grid": (ctypes.c_int64, [_u32, _vp, ctypes.c_float, _vp, _vp]),
        "orc_collide_brute": (ctypes.c_int64, [_u32, _vp, ctypes.c_float, _vp, _vp]),
        "orc_collide_grid_mt": (ctypes.c_int64, [_u32, _vp, ctypes.c_float, _vp, ctypes.c_int]),
        "orc_advance_f64": (None, [_u32, _vp, _vp, _vp, _vp, _vp]),
        "orc_diff_record": (ctypes.c_int, [_u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _u32, _vp,
                                           ctypes.c_float, ctypes.c_float, _u32, ctypes.c_int, _vp, _vp, _vp,
                                           _vp]),
        "orc_diff_fix_scale": (ctypes.c_double, [ctypes.c_float]),
        "orc_diff_set_threads": (None, [ctypes.c_int]),
        "orc_diff_replay64": (ctypes.c_int, [_u32, _u32, _vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                             _u32, _vp, _vp, _vp, _vp]),
        "orc_diff_adjoint64": (ctypes.c_int, [_u32, _u32, _vp, _vp, _vp, _vp, _u32, ctypes.c_double,
                                              ctypes.c_double, _u32, _vp, _vp, _vp, _vp, _vp, _vp]),
        "orc_diff_replay64_carry": (ctypes.c_int, [_u32, _u32, _vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                                   _u32, _vp, _vp, _vp, _vp, _vp, _vp]),
        "orc_diff_adjoint64_ex": (ctypes.c_int, [_u32, _u32, _vp, _vp, _vp, _vp, _u32, ctypes.c_double,
                                                 ctypes.c_double, _u32, _vp, _vp, _vp, _vp, _vp, ctypes.c_int,
                                                 ctypes.c_int, _vp, _vp, _vp, _vp]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    return lib


L = _load()


def ref_library():
    """The reference's quadtree collision code, or None where it was not built."""
    if not os.path.exists(REF_LIB):
        return None
    lib = ctypes.CDLL(REF_LIB)
    lib.ref_quadtree_pairs.restype = ctypes.c_longlong
    lib.ref_quadtree_pairs.argtypes = [ctypes.c_uint, _vp, _vp]
    lib.ref_known_answer_test.restype = ctypes.c_int
    lib.ref_world.argtypes = [_vp, _vp, _vp, _vp]
    return lib


# ---- scalar helpers ------------------------------------------------------

def rng_next(state):
    s = np.ascontiguousarray(state, np.uint32).copy()
    v = L.orc_rng_next(_p(s))
    return int(v), s


def rng_float(state):
    s = np.ascontiguousarray(state, np.uint32).copy()
    v = L.orc_rng_float(_p(s))
    return np.float32(v), s


def rng_range(state, lo, hi):
    s = np.ascontiguousarray(state, np.uint32).copy()
    v = L.orc_rng_range(_p(s), lo, hi)
    return int(v), s


def update_direction(tgt, pos, speed):
    t = np.asarray(tgt, np.float32).copy()
    p = np.asarray(pos, np.float32).copy()
    d = np.zeros(2, np.float32)
    L.orc_update_direction(_p(t), _p(p), speed, _p(d))
    return d


def in_range(pj, pi, radius):
    a = np.asarray(pj, np.float32).copy()
    b = np.asarray(pi, np.float32).copy()
    return bool(L.orc_in_range(_p(a), _p(b), radius))


def cell_size(radius):
    return float(L.orc_cell_size(radius))


def cells(pos, cell):
    p = np.ascontiguousarray(pos, np.float32)
    out = np.zeros((p.shape[0], 2), np.int32)
    L.orc_cells(p.shape[0], _p(p), cell, _p(out))
    return out


# ---- state ---------------------------------------------------------------

class State:
    """Structure-of-arrays agent state in caller order (the oracle's view)."""

    def __init__(self, entities, speed=None):
        n = entities.shape[0]
        self.n = n
        self.pos = np.ascontiguousarray(entities["pos"], np.float32).copy()
        self.tgt = np.ascontiguousarray(entities["target"], np.float32).copy()
        self.dir = np.ascontiguousarray(entities["direction"], np.float32).copy()
        self.road = np.ascontiguousarray(entities["roadIndex"], np.uint32).copy()
        self.rng = np.ascontiguousarray(entities["randomState"], np.uint32).copy()
        self.init = np.ascontiguousarray(entities["initialized"], np.uint32).copy()
        self.ncount = np.zeros(n, np.uint32)
        self.speed = None if speed is None else np.ascontiguousarray(speed, np.float32).copy()

    def copy(self):
        c = object.__new__(State)
        for k, v in self.__dict__.items():
            c.__dict__[k] = v.copy() if isinstance(v, np.ndarray) else v
        return c


def roads_buffer(roads):
    """The Road records as the oracle's 32-byte structs (same layout)."""
    return np.ascontiguousarray(roads).view(np.uint8)


class OracleSim:
    """Simulator::sim_tick protocol over the oracle (Simulator.cpp:213-279):
    even tick = move dispatch, odd = collision dispatch."""

    def __init__(self, roads, conns, entities, pc, speed=None, threads=1):
        self.roads = np.ascontiguousarray(roads)
        self.n_roads = roads.shape[0]
        self.conns = np.ascontiguousarray(conns, np.uint32)
        self.st = State(entities, speed)
        self.radius = float(np.asarray(pc)["collisionRadius"])
        self.tick = int(np.asarray(pc)["tick"])
        self.pairs_total = 0
        self.pair_hist = []
        self.inserts = 0
        self.arrivals = 0
        self.threads = threads
        self.any_collide = False

    def move(self):
        st = self.st
        arr = np.zeros(1, np.uint64)
        uninit = None
        if (st.init == 0).any():
            uninit = st.init.copy()
        if self.threads > 1 and uninit is None:
            rc = L.orc_move_phase_mt(st.n, _p(st.pos), _p(st.tgt), _p(st.dir), _p(st.road), _p(st.rng),
                                     _p(st.speed), _p(self.roads), self.n_roads, _p(self.conns),
                                     self.conns.shape[0], self.threads)
        else:
            rc = L.orc_move_phase(st.n, _p(st.pos), _p(st.tgt), _p(st.dir), _p(st.road), _p(st.rng),
                                  _p(st.speed), _p(st.init), _p(self.roads), self.n_roads, _p(self.conns),
                                  self.conns.shape[0], _p(arr))
        if rc != 0:
            raise RuntimeError(f"oracle move failed: {rc}")
        self.arrivals += int(arr[0])

    def collide(self, method="grid"):
        st = self.st
        if self.threads > 1 and not (st.init == 0).any() and method == "grid":
            pairs = L.orc_collide_grid_mt(st.n, _p(st.pos), self.radius, _p(st.ncount), self.threads)
        else:
            fn = L.orc_collide_grid if method == "grid" else L.orc_collide_brute
            pairs = fn(st.n, _p(st.pos), self.radius, _p(st.init), _p(st.ncount))
        if pairs < 0:
            raise RuntimeError("oracle collide failed")
        self.pairs_total += int(pairs)
        self.pair_hist.append(int(pairs))
        self.any_collide = True
        return int(pairs)

    def dispatch(self, tick=None):
        self.tick = self.tick + 1 if tick is None else tick
        if self.tick % 2 == 0:
            self.move()
        else:
            self.collide()

    def sim_tick(self):
        self.dispatch()
        self.dispatch()

    def run_ticks(self, n):
        for _ in range(n):
            self.sim_tick()


def advance_f64(pos, tgt, speed=None):
    p = np.ascontiguousarray(pos, np.float32)
    t = np.ascontiguousarray(tgt, np.float32)
    s = None if speed is None else np.ascontiguousarray(speed, np.float32)
    out = np.zeros((p.shape[0], 2), np.float64)
    arr = np.zeros(p.shape[0], np.int32)
    L.orc_advance_f64(p.shape[0], _p(p), _p(t), _p(s), _p(out), _p(arr))
    return out, arr


# ---- differentiable tick ---------------------------------------------------

class DiffSchedule:
    """Decisions of a recorded differentiable run: per tick the target and
    road each move used and a flag (1 arrived, 2 no move), caller order;
    final_tgt / final_road: the decisions in force after the last tick."""

    def __init__(self, tgt, road, flag, final_tgt=None, final_road=None):
        self.tgt = np.ascontiguousarray(tgt, np.float32)    # [T, n, 2]
        self.road = np.ascontiguousarray(road, np.uint32)   # [T, n]
        self.flag = np.ascontiguousarray(flag, np.uint8)    # [T, n]
        self.ticks = self.flag.shape[0]
        self.final_tgt = None if final_tgt is None else np.ascontiguousarray(final_tgt, np.float32)  # [n, 2]
        self.final_road = None if final_road is None else np.ascontiguousarray(final_road, np.uint32)  # [n]


def diff_record(roads, conns, entities, speed, agent_off, theta, alpha, radius, ticks, trajectory=False):
    """binary32 differentiable run from the entities' state (the first tick
    does not move when the agents are uninitialised): (final pos, schedule),
    and with trajectory=True also every tick's pre-move and post-move
    positions, float32 [ticks, 2, n, 2]."""
    st = State(entities, speed)
    n = st.n
    off = np.ascontiguousarray(agent_off, np.uint32)
    th = np.ascontiguousarray(theta, np.float32)
    rb = roads_buffer(roads)
    cn = np.ascontiguousarray(conns, np.uint32)
    tg = np.zeros((ticks, n, 2), np.float32)
    rd = np.zeros((ticks, n), np.uint32)
    fl = np.zeros((ticks, n), np.uint8)
    no_move = int((st.init == 0).all()) if n else 0
    traj = np.zeros((ticks, 2, n, 2), np.float32) if trajectory else None
    rc = L.orc_diff_record(n, off.shape[0] - 1, _p(off), _p(st.pos), _p(st.tgt), _p(st.road), _p(st.rng),
                           _p(st.speed), _p(rb), roads.shape[0], _p(cn), cn.shape[0], _p(th), alpha, radius, ticks,
                           no_move, _p(tg), _p(rd), _p(fl), _p(traj))
    if rc != 0:
        raise RuntimeError(f"orc_diff_record failed: {rc}")
    sched = DiffSchedule(tg, rd, fl, st.tgt.copy(), st.road.copy())
    return (st.pos, sched, traj) if trajectory else (st.pos, sched)


def diff_replay64(pos0, speed, agent_off, theta, alpha, radius, sched):
    p0 = np.ascontiguousarray(pos0, np.float32)
    n = p0.shape[0]
    off = np.ascontiguousarray(agent_off, np.uint32)
    th = np.ascontiguousarray(theta, np.float64)
    sp = None if speed is None else np.ascontiguousarray(speed, np.float32)
    out = np.zeros((n, 2), np.float64)
    rc = L.orc_diff_replay64(n, off.shape[0] - 1, _p(off), _p(p0), _p(sp), _p(th), float(alpha), float(radius),
                             sched.ticks, _p(sched.tgt), _p(sched.road), _p(sched.flag), _p(out))
    if rc != 0:
        raise RuntimeError("orc_diff_replay64 failed")
    return out


def diff_adjoint64(pos0, speed, agent_off, theta, alpha, radius, sched, grad_pos, carry=False, carry_traj=False,
                   traj32=None):
    """(dL/dtheta, dL/dalpha) in binary64 for dL/d(final pos) = grad_pos.
    carry: arrivals differentiated as carrying their leftover time onto the
    next road (the surrogate); carry_traj: on the carried trajectory
    (diff_replay64_carry's) instead of the snapped one; traj32: linearised
    around this binary32 trajectory (diff_record's) instead of the binary64
    replay."""
    p0 = np.ascontiguousarray(pos0, np.float32)
    n = p0.shape[0]
    off = np.ascontiguousarray(agent_off, np.uint32)
    th = np.ascontiguousarray(theta, np.float64)
    sp = None if speed is None else np.ascontiguousarray(speed, np.float32)
    g = np.ascontiguousarray(grad_pos, np.float64)
    gt = np.zeros(th.shape[0], np.float64)
    ga = np.zeros(1, np.float64)
    if (carry or carry_traj) and sched.final_tgt is None:
        raise ValueError("the carry-over needs the schedule's final decisions")
    tj = None if traj32 is None else np.ascontiguousarray(traj32, np.float32)
    if tj is not None and tj.shape != (sched.ticks, 2, n, 2):
        raise ValueError("traj32 must be [ticks, 2, n, 2]")
    rc = L.orc_diff_adjoint64_ex(n, off.shape[0] - 1, _p(off), _p(p0), _p(sp), _p(th), th.shape[0], float(alpha),
                                 float(radius), sched.ticks, _p(sched.tgt), _p(sched.road), _p(sched.flag),
                                 _p(sched.final_tgt), _p(sched.final_road), int(carry_traj), int(carry),
                                 _p(tj), _p(g),
                                 _p(gt), _p(ga))
    if rc != 0:
        raise RuntimeError("orc_diff_adjoint64 failed")
    return gt, float(ga[0])


def diff_replay64_carry(pos0, speed, agent_off, theta, alpha, radius, sched):
    """binary64 replay in which every arrival carries its leftover time onto
    the next road (the model the carry-over surrogate differentiates)."""
    p0 = np.ascontiguousarray(pos0, np.float32)
    n = p0.shape[0]
    off = np.ascontiguousarray(agent_off, np.uint32)
    th = np.ascontiguousarray(theta, np.float64)
    sp = None if speed is None else np.ascontiguousarray(speed, np.float32)
    out = np.zeros((n, 2), np.float64)
    rc = L.orc_diff_replay64_carry(n, off.shape[0] - 1, _p(off), _p(p0), _p(sp), _p(th), float(alpha),
                                   float(radius), sched.ticks, _p(sched.tgt), _p(sched.road), _p(sched.flag),
                                   _p(sched.final_tgt), _p(sched.final_road), _p(out))
    if rc != 0:
        raise RuntimeError("orc_diff_replay64_carry failed")
    return out


def diff_fix_scale(radius):
    """The binary32 record's fixed-point unit for push terms (1 / its value)."""
    return float(L.orc_diff_fix_scale(radius))


def diff_set_threads(n):
    """Threads for the binary32 record, the binary64 replay and the adjoint
    (diff_record, diff_replay64*, diff_adjoint64)."""
    L.orc_diff_set_threads(int(n))
