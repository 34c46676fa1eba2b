"""Benchmark: fused ticks of the plain pipeline on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2] [--cpu-seconds S]

A step is one tick (Simulator::sim_tick: move dispatch + collision dispatch)
over every agent of the workload.  Inputs are uploaded to HBM before timing;
K ticks run as hipGraph replays between two stream synchronisations.  Prints
one JSON line.
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
    1: "plain small: 32x32 jittered lattice (1,024 nodes, 1,686 roads), 10,000 agents, R 1 m, 1 m cells",
    2: "city-scale plain drop-in: 150k-node Gaussian-clustered Delaunay city graph (12 km square, "
       "mean degree 2.6), 1,000,000 agents, speed U[0.8,1.6) m/tick, R 1 m, 1 m cells",
    3: "dense stress: 1M-node city graph (31 km square), 16,000,000 agents with Zipf(1.1) road "
       "occupancy, R 1 m, 1 m cells",
}


def cpu_baseline(sc, seconds, threads):
    """The CPU oracle (restatement; the reference itself needs a Vulkan ICD)
    on the same initial state: one untimed init tick, then full ticks until
    `seconds` elapse."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc

    o = orc.OracleSim(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, threads=threads)
    o.sim_tick()  # insert-only move dispatch + first collision dispatch
    ticks = 0
    t0 = time.perf_counter()
    while True:
        o.sim_tick()
        ticks += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    n = sc.entities.shape[0]
    return {"value": n * ticks / el, "unit": "agent-ticks/s", "cores": threads, "kind": "port",
            "sample": f"{n} agents x {ticks} ticks of the same workload ({el:.1f} s), oracle/msim_oracle.c "
                      f"with {threads} POSIX threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--config", type=int, default=2, choices=(1, 2, 3))
    ap.add_argument("--profile-ticks", type=int, default=240)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--untuned", action="store_true", help="library defaults instead of msim/tuned.json")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs are not part of it "
              "(use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import msim
    from msim import roofline as RL
    from msim import scenarios as S

    t = time.perf_counter()
    sc = S.config(args.config)
    gen_s = time.perf_counter() - t
    n = sc.entities.shape[0]
    print(f"# workload config {args.config}: {n} agents, {sc.roads.shape[0]} roads, "
          f"{sc.conns.shape[0]} connection entries (generated in {gen_s:.1f} s)", file=sys.stderr, flush=True)
    from msim import autotune

    opts = dict(autotune.DEFAULT) if args.untuned else autotune.tuned_options(sc)
    sim = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, **opts)
    info = sim.info()

    sim.run_ticks(max(1, args.warmup))  # includes the insert-only first dispatch
    sim.sync()
    t0 = time.perf_counter()
    sim.run_ticks(args.steps)
    sim.sync()
    el = time.perf_counter() - t0
    ms_per_step = 1e3 * el / args.steps
    value = n * args.steps / el
    print(f"# timed {args.steps} ticks in {el:.3f} s", file=sys.stderr, flush=True)

    # per-kernel durations: HIP events right before and after one launch per
    # tick, on the handle's stream, each kernel probed in turn
    _, kms = sim.time_ticks(args.profile_ticks, kernels=True)
    per_launch_ms = {k: v / args.profile_ticks for k, v in kms.items()}
    dom = max(per_launch_ms, key=per_launch_ms.get)
    uniform = bool(info["uniform_speed"])
    dom_bytes = RL.kernel_bytes(dom, uniform) * n
    achieved = dom_bytes / (per_launch_ms[dom] * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_config{args.config}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            tj = json.load(f)
        traffic = tj.get("kernels", {}).get(dom, {}).get("bytes_per_launch")
    tick_kernel_ms = sum(per_launch_ms.values())
    fused_gbs = RL.B_FUSED * n / (ms_per_step * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(sc, args.cpu_seconds, threads)

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
        "config": {
            "workload": WORKLOADS[args.config],
            "agents": n,
            "roads": int(sc.roads.shape[0]),
            "nodes": int(sc.meta.get("nodes", 0)),
            "collision_radius_m": float(sc.pc["collisionRadius"]),
            "cell_m": info["cell_size"],
            "bucket_table": [info["table_width"], info["table_height"]],
            "options": opts,
            "graph_ticks": 32,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": RL.HBM_PEAK_BYTES_PER_S / 1e9,
            "unit": "GB/s",
            "frac": achieved / (RL.HBM_PEAK_BYTES_PER_S / 1e9),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": dom_bytes,
            "avg_launch_us": per_launch_ms[dom] * 1e3,
        },
        "kernels_us_per_tick": {k: v * 1e3 for k, v in per_launch_ms.items()},
        "fused_tick": {"bytes_per_agent": RL.B_FUSED, "achieved_GBps": fused_gbs,
                       "frac": fused_gbs / (RL.HBM_PEAK_BYTES_PER_S / 1e9),
                       "kernel_sum_us": tick_kernel_ms * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
