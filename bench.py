"""Benchmark on one MI355X: agent-ticks/s of the batched form against the
plain per-scenario loop at bit-exact state, and of the plain drop-in.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 4] [--cpu-seconds S]

A step is one tick (Simulator::sim_tick: move dispatch + collision dispatch)
over every agent of the workload.  --config 4 (default) is the batched
workload (512 scenarios); 1-3 are plain single-map workloads.  Inputs are
uploaded to HBM before timing; K ticks run as hipGraph segment replays
between two stream synchronisations.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "movement-sim_amd"))

METRIC = "agent-ticks/s, 1 GPU: batched vs plain at bit-exact state; HBM GB/s % of peak"

WORKLOADS = {
    5: "adjoint calibration: 64 scenarios of 100,000 agents on one 20,000-node jittered lattice (141x141, "
       "32.3 m spacing), 1,000 ticks, per-road speed multipliers (true: log-normal sigma 0.2) and a soft-"
       "avoidance strength (true 0.02), loss = mean squared final-position error, Adam, checkpoints every "
       "50 ticks",
    4: "batched: 512 independent jittered-lattice scenarios (32.3 m node spacing, 15% of edges dropped), "
       "nodes log-uniform in [1e3, 1e5], agents log-uniform in [1e3, 5e5] (~41M agents), speed "
       "U[0.8,1.6) m/tick, R 1 m, 1 m cells",
    1: "plain small: 32x32 jittered lattice (1,024 nodes, 1,686 roads), 10,000 agents, R 1 m, 1 m cells",
    2: "city-scale plain drop-in: 150k-node Gaussian-clustered Delaunay city graph (12 km square, "
       "mean degree 2.6), 1,000,000 agents, speed U[0.8,1.6) m/tick, R 1 m, 1 m cells",
    3: "dense stress: 1M-node city graph (31 km square), 16,000,000 agents with Zipf(1.1) road "
       "occupancy, R 1 m, 1 m cells",
}


def lavapipe_probe():
    """Is the reference's CPU Vulkan path (Mesa lavapipe) available here?"""
    import ctypes.util
    import glob
    import shutil

    icds = sorted(set(glob.glob("/usr/share/vulkan/icd.d/*.json") + glob.glob("/etc/vulkan/icd.d/*.json")))
    found = {"vulkaninfo": shutil.which("vulkaninfo"), "libvulkan": ctypes.util.find_library("vulkan"),
             "icd_files": icds, "glslc": shutil.which("glslc") or shutil.which("glslangValidator")}
    found["lavapipe"] = any("lvp" in os.path.basename(x) for x in icds)
    return found


def cpu_baseline(scs, seconds, threads):
    """The CPU oracle (restatement; the reference itself needs a Vulkan ICD,
    see lavapipe_probe) on a bounded sample of the workload: the first
    scenarios, each from its own initial state (one untimed insert-only
    tick).  A warm-up run sizes each scenario's tick count to its share of
    seconds / 3; then three timed runs of exactly those ticks from the
    initial states, and the median rate is reported."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc

    k = min(len(scs), 8)
    share = seconds / 3 / k

    def run(sc, ticks=None):
        o = orc.OracleSim(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, threads=threads)
        o.sim_tick()  # insert-only move dispatch + first collision dispatch
        done = 0
        t0 = time.perf_counter()
        while (ticks is None and time.perf_counter() - t0 < share) or (ticks is not None and done < ticks):
            o.sim_tick()
            done += 1
        return done, time.perf_counter() - t0

    plan = [max(1, run(sc)[0]) for sc in scs[:k]]  # warm-up, sizes the sample
    rates, secs = [], []
    for _ in range(3):
        at = el = 0.0
        for sc, ticks in zip(scs[:k], plan):
            _, e = run(sc, ticks)
            at += sc.entities.shape[0] * ticks
            el += e
        rates.append(at / el)
        secs.append(el)
    parts = [f"{sc.entities.shape[0]}x{t}" for sc, t in zip(scs[:k], plan)]
    return {"value": sorted(rates)[1], "unit": "agent-ticks/s", "cores": threads, "kind": "port",
            "runs": rates,
            "sample": f"first {k} scenario(s) of the same workload, agents x ticks = {', '.join(parts)}; median "
                      f"of 3 runs ({', '.join(f'{x:.1f}' for x in secs)} s) after one warm-up; "
                      f"oracle/msim_oracle.c with {threads} POSIX threads",
            "reference_cpu_path": lavapipe_probe()}


def prime_graphs(h, steps, opts):
    """Captures the graph segments a run of `steps` ticks will replay, from
    the parity the timed run starts with (a segment of an even number of
    ticks ends on the buffers it started from), so no capture falls inside
    the timed region."""
    h.run_ticks(steps if steps % 2 == 0 else steps + 1)
    h.sync()


def graph_ticks_for(steps, tuned):
    """Segment length for a timed run of `steps` ticks: the tuned length, cut
    to an even number below steps so that the run replays captured segments
    (its last tick always runs direct)."""
    g = min(int(tuned or 100), steps - 1)
    return max(2, g - (g % 2))


def plain_line(args, sc):
    """Plain drop-in on one workload: fused ticks through the graph path."""
    import msim
    from msim import autotune
    from msim import roofline as RL

    n = sc.entities.shape[0]
    opts = dict(autotune.DEFAULT) if args.untuned else autotune.tuned_options(sc)
    for k in ("binning", "collide_mode", "bin_unroll", "reorder_interval", "segment_log2"):
        if getattr(args, k) is not None:
            opts[k] = getattr(args, k)
    opts["graph_ticks"] = graph_ticks_for(args.steps, opts.get("graph_ticks"))
    sim = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, **opts)
    info = sim.info()

    sim.run_ticks(max(1, args.warmup))  # includes the insert-only first dispatch
    sim.sync()
    prime_graphs(sim, args.steps, opts)
    t0 = time.perf_counter()
    sim.run_ticks(args.steps)
    sim.sync()
    el = time.perf_counter() - t0
    print(f"# timed {args.steps} ticks in {el:.3f} s", file=sys.stderr, flush=True)

    # per-kernel durations: HIP events right before and after one launch per
    # tick, on the handle's stream, each kernel probed in turn
    _, kms = sim.time_ticks(args.profile_ticks, kernels=True)
    per_launch_ms = {k: v / args.profile_ticks for k, v in kms.items()}
    roof = roofline(per_launch_ms, n, bool(info["uniform_speed"]), args.config)
    config = {
        "workload": WORKLOADS[args.config],
        "agents": n,
        "roads": int(sc.roads.shape[0]),
        "nodes": int(sc.meta.get("nodes", 0)),
        "collision_radius_m": float(sc.pc["collisionRadius"]),
        "cell_m": info["cell_size"],
        "bucket_table": [info["table_width"], info["table_height"]],
        "options": opts,
    }
    extra = {"fused_tick": fused(el / args.steps, n, per_launch_ms)}
    return n * args.steps / el, 1e3 * el / args.steps, config, roof, extra


def dense_record(args):
    """The dense stress workload (16M agents on one city map) through the
    plain handle with its tuned options: ms per tick, the fused tick's
    fraction of HBM peak and the dominant kernel against its roofline."""
    from msim import scenarios as S

    t = time.perf_counter()
    sc = S.config(3)
    gen_s = time.perf_counter() - t
    a = argparse.Namespace(**vars(args))
    a.config, a.steps, a.warmup = 3, args.dense_steps, 20
    for k in ("binning", "collide_mode", "bin_unroll", "reorder_interval", "segment_log2"):
        setattr(a, k, None)
    print(f"# dense stress workload: {sc.entities.shape[0]} agents (generated in {gen_s:.1f} s)",
          file=sys.stderr, flush=True)
    value, ms, config, roof, extra = plain_line(a, sc)
    kus = roof.pop("kernels_us_per_tick")
    return {"workload": WORKLOADS[3], "agents": config["agents"], "steps": a.steps, "ms_per_tick": ms,
            "agent_ticks_per_s": value, "fused_tick": extra["fused_tick"], "roofline": roof,
            "kernels_us_per_tick": kus, "options": config["options"]}


def roofline(per_launch_ms, n, uniform, cfg):
    """Dominant kernel's algorithmic bytes per launch / its average launch time."""
    from msim import roofline as RL

    dom = max(per_launch_ms, key=per_launch_ms.get)
    dom_bytes = RL.kernel_bytes(dom, uniform) * n
    achieved = dom_bytes / (per_launch_ms[dom] * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_config{cfg}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("kernels", {}).get(dom.split("<")[0], {}).get("bytes_per_launch")
    peak = RL.HBM_PEAK_BYTES_PER_S / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": peak, "unit": "GB/s",
            "frac": achieved / peak, "traffic": traffic, "algorithmic_bytes_per_launch": dom_bytes,
            "avg_launch_us": per_launch_ms[dom] * 1e3,
            "kernels_us_per_tick": {k: v * 1e3 for k, v in per_launch_ms.items()}}


def fused(s_per_tick, n, per_launch_ms):
    from msim import roofline as RL

    gbs = RL.B_FUSED * n / s_per_tick / 1e9
    return {"bytes_per_agent": RL.B_FUSED, "achieved_GBps": gbs, "frac": gbs / (RL.HBM_PEAK_BYTES_PER_S / 1e9),
            "kernel_sum_us": sum(per_launch_ms.values()) * 1e3}


def compare_plain(args, scs, b, n):
    """The same ticks from the same initial state, batched and through the
    plain per-scenario loop; every scenario's state compared bit for bit."""
    import numpy as np

    import msim
    from msim import autotune

    t0 = time.perf_counter()
    b.run_ticks(args.compare_ticks)
    b.sync()
    t_batch_cmp = time.perf_counter() - t0
    t_plain = 0.0
    exact = True
    first_bad = None
    batch_pairs = b.pairs()[0]
    for s, sc in enumerate(scs):
        g = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, **autotune.tuned_options(sc))
        g.sync()
        t0 = time.perf_counter()
        g.run_ticks(args.compare_ticks)
        g.sync()
        t_plain += time.perf_counter() - t0
        bs, ps = b.state(s), g.state()
        same = all(np.array_equal(np.ascontiguousarray(bs[k]).view(np.uint32),
                                  np.ascontiguousarray(ps[k]).view(np.uint32)) for k in bs)
        same = same and int(batch_pairs[s]) == g.pairs()[0]
        if not same and first_bad is None:
            first_bad = s
        exact = exact and same
        g.close()
    print(f"# comparison: {args.compare_ticks} ticks, batched {t_batch_cmp:.3f} s, plain loop {t_plain:.3f} s, "
          f"bit-exact {exact}", file=sys.stderr, flush=True)
    plain_rate = n * args.compare_ticks / t_plain
    batch_cmp_rate = n * args.compare_ticks / t_batch_cmp
    return {
        "ticks": args.compare_ticks,
        "batched_agent_ticks_per_s": batch_cmp_rate,
        "plain_loop_agent_ticks_per_s": plain_rate,
        "speedup": batch_cmp_rate / plain_rate,
        "bit_exact_scenarios": exact,
        "first_mismatch": first_bad,
        "plain_loop": "one plain handle per scenario (tuned options), created and uploaded before its "
                      "timed run_ticks; times summed over the scenarios",
    }


def batched_line(args, scs):
    """Batched workload: every scenario advanced by one set of launches per
    tick (graph segments), against the plain per-scenario loop on the same
    scenarios and ticks, with every scenario's final state compared bit for
    bit."""
    import msim
    from msim import autotune

    n = sum(int(sc.entities.shape[0]) for sc in scs)
    opts = dict(autotune.BATCH_DEFAULT) if args.untuned else autotune.tuned_batch_options(scs)
    for k in ("segment_log2", "reorder_interval", "graph_ticks"):
        v = getattr(args, "batch_" + k)
        if v is not None:
            opts[k] = v
    for k in ("binning", "collide_mode", "bin_unroll"):
        if getattr(args, k) is not None:
            opts[k] = getattr(args, k)
    opts["graph_ticks"] = graph_ticks_for(args.steps, opts.get("graph_ticks"))
    b = msim.Batch(scs, **opts)
    extra = {}
    if args.compare_ticks > 0:
        extra["batched_vs_plain"] = compare_plain(args, scs, b, n)
    b.run_ticks(max(1, args.warmup))
    b.sync()
    prime_graphs(b, args.steps, opts)
    t0 = time.perf_counter()
    b.run_ticks(args.steps)
    b.sync()
    el = time.perf_counter() - t0
    print(f"# timed {args.steps} batched ticks in {el:.3f} s", file=sys.stderr, flush=True)
    _, kms = b.time_ticks(args.profile_ticks, kernels=True)
    per_launch_ms = {k: v / args.profile_ticks for k, v in kms.items()}
    roof = roofline(per_launch_ms, n, scs[0].speed is None, 4)
    agents = [int(sc.entities.shape[0]) for sc in scs]
    config = {
        "workload": WORKLOADS[4],
        "scenarios": len(scs),
        "agents": n,
        "agents_min_max": [min(agents), max(agents)],
        "roads": int(sum(sc.roads.shape[0] for sc in scs)),
        "collision_radius_m": float(scs[0].pc["collisionRadius"]),
        "options": opts,
    }
    extra["fused_tick"] = fused(el / args.steps, n, per_launch_ms)
    return n * args.steps / el, 1e3 * el / args.steps, config, roof, extra


def adjoint_line(args):
    """Calibration through the adjoint: one step = one forward + adjoint
    gradient of the whole run (all scenarios, all ticks) and an Adam update."""
    from msim import calibrate
    from msim import scenarios as S

    scs = S.adjoint_config()
    res = calibrate.calibrate(scs, ticks=1000, steps=args.steps, arrival_gradient=args.arrival_gradient,
                              log=lambda m: print("# " + m, file=sys.stderr, flush=True))
    n = res["agents"]
    sec = res["seconds_per_gradient"]
    h = res["history"]
    out = {
        "metric": METRIC,
        "value": n * res["ticks"] / sec,
        "unit": "agent-ticks/s (differentiable forward + adjoint)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 0,
        "ms_per_step": 1e3 * sec,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (adjoint f64)",
        "data": "synthetic (seeded generator, msim/scenarios.py)",
        "config": {"workload": WORKLOADS[5], "agents": n, "roads_per_map": res["roads_per_map"],
                   "ticks": res["ticks"], "checkpoint_every": 50},
        "adjoint": {
            "arrival_gradient": res["arrival_gradient"],
            "seconds_per_gradient": sec,
            "hidden_run_seconds": res["hidden_run_seconds"],
            "loss_m2_first_last": [h[0]["loss_m2"], h[-1]["loss_m2"]],
            "loss_m2_best": [res["best_step"], h[res["best_step"]]["loss_m2"]],
            "alpha_first_last": [h[0]["alpha"], h[-1]["alpha"]],
            "alpha_true": res["alpha_true"],
            "theta_mean_rel_err_first_last": [h[0]["theta_mean_rel_err"], h[-1]["theta_mean_rel_err"]],
            "loss_m2_every_10": [x["loss_m2"] for x in h[::10]],
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed ticks (default: 1000 batched, 10000 plain)")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--config", type=int, default=4, choices=(1, 2, 3, 4, 5))
    ap.add_argument("--scenarios", type=int, default=512, help="batched workload: number of scenarios")
    ap.add_argument("--compare-ticks", type=int, default=200, help="batched workload: ticks of the plain-loop "
                                                                    "comparison")
    ap.add_argument("--batch-reorder-interval", type=int, default=None, help="override the tuned value")
    ap.add_argument("--batch-graph-ticks", type=int, default=None, help="override the tuned value")
    ap.add_argument("--batch-segment-log2", type=int, default=None, help="override the tuned value")
    ap.add_argument("--binning", type=int, default=None, help="override the tuned rebinning mode (0 two passes, "
                                                               "1 atomic (plain only), 2 one pass)")
    ap.add_argument("--collide-mode", type=int, default=None, help="override the tuned neighbour walk (0 staged, "
                                                                    "2 slots per lane; 1 per lane; 2 staged, 1 slot)")
    ap.add_argument("--reorder-interval", type=int, default=None, help="plain workloads: override the tuned "
                    "renumbering interval (ticks)")
    ap.add_argument("--segment-log2", type=int, default=None, help="plain workloads: override the tuned "
                    "rebinning segment size (log2 buckets)")
    ap.add_argument("--bin-unroll", type=int, default=None, help="override the tuned rebinning load batching (4, 8)")
    ap.add_argument("--arrival-gradient", choices=("frozen", "carry"), default="carry",
                    help="calibration (--config 5): how the adjoint treats arrivals")
    ap.add_argument("--profile-ticks", type=int, default=240)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--untuned", action="store_true", help="library defaults instead of msim/tuned.json")
    ap.add_argument("--dense-steps", type=int, default=500, help="batched workload: timed ticks of the dense "
                    "stress sub-record (0 leaves it out)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs are not part of it "
              "(use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = {1: 10000, 2: 10000, 3: 500, 4: 1000, 5: 50}[args.config]

    from msim import scenarios as S

    t = time.perf_counter()
    if args.config == 5:
        adjoint_line(args)
        return
    if args.config == 4:
        scs = S.batch_config(args.scenarios)
        n = sum(int(sc.entities.shape[0]) for sc in scs)
        print(f"# batched workload: {len(scs)} scenarios, {n} agents (generated in "
              f"{time.perf_counter() - t:.1f} s)", file=sys.stderr, flush=True)
        value, ms_per_step, config, roof, extra = batched_line(args, scs)
        sample_sc = scs
        if args.dense_steps > 0:
            extra["dense16M"] = dense_record(args)
    else:
        sc = S.config(args.config)
        print(f"# workload config {args.config}: {sc.entities.shape[0]} agents, {sc.roads.shape[0]} roads "
              f"(generated in {time.perf_counter() - t:.1f} s)", file=sys.stderr, flush=True)
        value, ms_per_step, config, roof, extra = plain_line(args, sc)
        sample_sc = [sc]

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(sample_sc, args.cpu_seconds, threads)

    kus = roof.pop("kernels_us_per_tick")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "agent-ticks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, msim/scenarios.py)",
        "config": config,
        "roofline": roof,
        "kernels_us_per_tick": kus,
        **extra,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
