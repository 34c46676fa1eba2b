"""Autotuner for the plain and batched pipelines' launch and layout parameters.

Coordinate descent over the msim_options knobs that change speed but never
results (every value is covered by the parity tests): bucket-table row width,
rebinning segment size, binning mode and load batching, inline vs queued
route choice, spatial reorder interval and the neighbour walk.  Each candidate is timed on the real
workload with HIP events around graph-replayed ticks; the winner is cached in
a JSON file keyed by the workload's shape.

    python -m msim.autotune --config 2          (from movement-sim_amd/)
    python -m msim.autotune --batch             (the 512-scenario batched workload)

The batched handle's knobs: rebinning segment size (for every scenario),
reorder interval, ticks per captured graph segment, the neighbour walk and
the rebinning mode (two passes, one pass, or the bucket-sorted state).
"""
import argparse
import json
import os
import time

import numpy as np

from ._lib import Batch, Simulation, MsimError

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuned.json")

# Knob -> candidate values (0 = the library's automatic choice where it has one).
# The descent runs once per layout (binning), so that the segment size and
# load batching are tuned for each layout, and keeps the best.
SPACE = {
    "binning": [0, 1, 2, 3],
    "segment_log2": [0, 9, 10, 11, 12, 13, 14],
    "bin_unroll": [0, 2, 4, 8],
    "table_width_log2": [0, 5, 6, 7, 8, 9, 10, 11, 12],
    "route_queue": [0, 1],
    "reorder_interval": [8, 16, 32, 64],
    "collide_mode": [0, 1, 2],
}
DEFAULT = {"table_width_log2": 0, "segment_log2": 0, "binning": 0, "route_queue": 0,
           "reorder_interval": 16, "collide_mode": 0, "bin_unroll": 0}


BATCH_SPACE = {
    "binning": [0, 2, 3],
    "segment_log2": [0, 11, 12, 13],
    "bin_unroll": [0, 2, 4, 8],
    "reorder_interval": [10, 20, 25, 50],
    "graph_ticks": [50, 100, 200],
    "collide_mode": [0, 1, 2],
}
BATCH_DEFAULT = {"segment_log2": 0, "reorder_interval": 20, "graph_ticks": 100, "collide_mode": 0, "binning": 0,
                 "bin_unroll": 0}


def batch_key(scs):
    n = int(sum(sc.entities.shape[0] for sc in scs))
    r = int(sum(sc.roads.shape[0] for sc in scs))
    uni = "u" if scs[0].speed is None else "s"
    return f"batch{len(scs)}-n{n}-r{r}-R{float(scs[0].pc['collisionRadius']):g}-{uni}"


def _prime(h, warmup, ticks):
    """Warm-up, then one untimed run as long as the timed one (rounded up to
    an even number of ticks, which ends on the buffers it started from), so
    that the graph segments the timed run replays are captured outside it."""
    h.run_ticks(warmup)
    h.run_ticks(ticks + (ticks % 2))
    h.sync()


def time_batch_options(scs, opts, warmup, ticks):
    """Microseconds per batched tick (HIP events around graph segments)."""
    try:
        b = Batch(scs, **opts)
    except MsimError:
        return float("inf")
    try:
        _prime(b, warmup, ticks)
        ms, _ = b.time_ticks(ticks)
        return 1e3 * ms / ticks
    except MsimError:
        return float("inf")
    finally:
        b.close()


def tune_batch(scs, warmup=101, ticks=200, log=print):
    return _descend_per_layout(lambda o: time_batch_options(scs, o, warmup, ticks), BATCH_DEFAULT, BATCH_SPACE, log)


def tuned_batch_options(scs, path=CACHE):
    ent = load_cache(path).get(batch_key(scs))
    return dict(ent["options"]) if ent else dict(BATCH_DEFAULT)


def workload_key(sc):
    n = int(sc.entities.shape[0])
    w, h = (int(round(v)) for v in sc.world)
    uni = "u" if sc.speed is None else "s"
    return f"n{n}-r{int(sc.roads.shape[0])}-w{w}x{h}-R{float(sc.pc['collisionRadius']):g}-{uni}"


def time_options(sc, opts, warmup, ticks):
    """Microseconds per tick through the graph path (HIP events)."""
    try:
        sim = Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, **opts)
    except MsimError:
        return float("inf")
    try:
        _prime(sim, warmup, ticks)
        ms, _ = sim.time_ticks(ticks, kernels=False)
        sim.sync()
        return 1e3 * ms / ticks
    except MsimError:
        return float("inf")
    finally:
        sim.close()


def tune(sc, warmup=20, ticks=100, space=None, log=print):
    space = SPACE if space is None else space
    return _descend_per_layout(lambda o: time_options(sc, o, warmup, ticks), DEFAULT, space, log)


def _descend(timer, default, space, log):
    """Coordinate descent: each knob in turn, keeping the best value found."""
    best = dict(default)
    best_us = timer(best)
    log(f"  start {best} -> {best_us:.1f} us/tick")
    trials = [(dict(best), best_us)]
    for knob, values in space.items():
        for v in values:
            if v == best[knob]:
                continue
            cand = dict(best)
            cand[knob] = v
            us = timer(cand)
            trials.append((cand, us))
            log(f"  {knob}={v}: {us:.1f} us/tick" + (" (new best)" if us < best_us else ""))
            if us < best_us:
                best, best_us = cand, us
    return best, best_us, trials


def _descend_per_layout(timer, default, space, log):
    """_descend over the other knobs from each binning value; the best of
    those descents, with every trial."""
    best, best_us, trials = None, float("inf"), []
    rest = {k: v for k, v in space.items() if k != "binning"}
    for b in space.get("binning", [default.get("binning", 0)]):
        start = dict(default)
        start["binning"] = b
        log(f" binning {b}")
        cand, us, tr = _descend(timer, start, rest, log)
        trials += tr
        if us < best_us:
            best, best_us = cand, us
    return best, best_us, trials


def load_cache(path=CACHE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def tuned_options(sc, path=CACHE):
    """Cached best options for this workload shape, or the defaults."""
    ent = load_cache(path).get(workload_key(sc))
    return dict(ent["options"]) if ent else dict(DEFAULT)


def main():
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    from msim import scenarios as S

    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, nargs="*", default=None, help="plain configurations (default 2, "
                                                                            "none with --batch)")
    ap.add_argument("--batch", action="store_true", help="tune the batched workload")
    ap.add_argument("--ticks", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--out", default=CACHE)
    args = ap.parse_args()
    cache = load_cache(args.out)
    if args.config is None:
        args.config = [] if args.batch else [2]
    if args.batch:
        scs = S.batch_config()
        key = batch_key(scs)
        print(f"batched workload ({key})", flush=True)
        t0 = time.time()
        best, us, trials = tune_batch(scs, log=lambda m: print(m, flush=True))
        n = sum(sc.entities.shape[0] for sc in scs)
        cache[key] = {"config": 4, "options": best, "us_per_tick": us, "agent_ticks_per_s": n / us * 1e6,
                      "default_us_per_tick": trials[0][1], "trials": len(trials),
                      "tuned_s": round(time.time() - t0, 1)}
        print(json.dumps(cache[key]), flush=True)
        with open(args.out, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    for k in args.config:
        sc = S.config(k)
        key = workload_key(sc)
        print(f"config {k} ({key})", flush=True)
        t0 = time.time()
        best, us, trials = tune(sc, args.warmup, args.ticks, log=lambda m: print(m, flush=True))
        n = sc.entities.shape[0]
        cache[key] = {"config": k, "options": best, "us_per_tick": us, "agent_ticks_per_s": n / us * 1e6,
                      "default_us_per_tick": trials[0][1], "trials": len(trials),
                      "tuned_s": round(time.time() - t0, 1)}
        print(json.dumps(cache[key]), flush=True)
        with open(args.out, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
