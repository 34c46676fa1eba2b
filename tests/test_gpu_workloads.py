"""The batched and the differentiable workloads at their full sizes.

* Batched: 512 independent jittered-lattice scenarios (38.8M agents), with
  the tuned options and with the sorted state, in 10-tick captured segments
  (the first captured, the next replayed).  Every scenario must end
  bit-identical to its own plain handle (state and pair counts), and the 4
  smallest and the 2 largest scenarios are checked against the oracle.
* Differentiable: 2 scenarios of the calibration workload (100k agents each
  on one 20k-node map), 1,000 ticks, checkpoints every 50 ticks, avoidance
  alpha = 0.02.  The GPU run must equal the oracle's binary32 record bit for
  bit (final positions and every tick's decisions), and dL/dtheta and
  dL/dalpha must agree within 1e-3 (relative) with the binary64 host adjoint
  linearised around that same binary32 run, with frozen arrivals and with
  the carry-over surrogate.  The host adjoint itself is checked at this size
  against central differences of its binary64 replay.

Why the adjoint is compared around the binary32 run: the binary32 and
binary64 trajectories drift apart by rounding (about 3e-4 m after 10 ticks),
the pushes amplify the drift over 1,000 ticks, and the two linearisation
points then give gradients that differ by about 1e-2 (dL/dtheta) and by
more than dL/dalpha itself, whatever the adjoint's accuracy.  The central
differences use eps = 1e-5: the replay is smooth under the recorded
decisions, but the pushes curve it enough over 1,000 ticks that eps = 1e-3
differences miss the derivative by 5e-2 (dL/dtheta) and 0.55 (dL/dalpha) on
this workload and eps = 1e-4 by 4e-3 and 1e-3.  The carry-over surrogate's
replay moves arrivals by the distance L = |T - p| to the target, which has a
kink wherever a perturbed agent passes its target within a tick; its
differences converge only linearly in eps (thousands of arrivals per tick
sit near such a kink), so their check runs on the CPU at a size where the
kinks are few (test_oracle_adjoint.py).
"""
import os
import time

import numpy as np
import pytest

import msim
import orc
from msim import scenarios as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from msim.adjoint import DiffRun  # noqa: E402

THREADS = min(16, os.cpu_count() or 1)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_state_equal(a, b, what):
    for k in ("pos", "tgt", "dir"):
        bad = np.flatnonzero((bits(a[k]) != bits(b[k])).any(axis=1))
        assert bad.size == 0, f"{what}: {k} differs for {bad.size} agents, first {bad[:5]}"
    for k in ("road", "rng", "ncount"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")


def say(msg):
    print(msg, flush=True)


CONFIG4_TICKS = 30
CONFIG4_OPTIONS = {
    "tuned": dict(binning=2, bin_unroll=4, collide_mode=0, reorder_interval=14, segment_log2=12, graph_ticks=10),
    "sorted": dict(binning=3, bin_unroll=4, segment_log2=12, graph_ticks=10),
}


def test_config4_full_size_every_scenario_equals_plain_and_oracle_sample():
    t0 = time.perf_counter()
    scs = S.batch_config(512)
    n = np.array([sc.entities.shape[0] for sc in scs])
    say(f"512 scenarios, {n.sum()} agents, generated in {time.perf_counter() - t0:.1f} s")
    batches = {}
    for name, opts in CONFIG4_OPTIONS.items():
        b = msim.Batch(scs, **opts)
        b.run_ticks(CONFIG4_TICKS)
        batches[name] = (b, b.pairs())
        say(f"{name}: {CONFIG4_TICKS} ticks done")
    for s, sc in enumerate(scs):
        g = msim.Simulation(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed)
        g.run_ticks(CONFIG4_TICKS)
        want = g.state()
        total, hist = g.pairs()
        last = int(hist[-1]) if hist.size else 0
        g.close()
        for name, (b, (tot, lst)) in batches.items():
            assert_state_equal(b.state(s), want, f"{name}, scenario {s} ({n[s]} agents)")
            assert int(tot[s]) == total, f"{name}, scenario {s}: pair total"
            assert int(lst[s]) == last, f"{name}, scenario {s}: pairs of the last tick"
    say(f"every scenario equals its plain handle ({time.perf_counter() - t0:.1f} s)")
    order = np.argsort(n, kind="stable")
    for s in list(order[:4]) + list(order[-2:]):
        sc = scs[s]
        o = orc.OracleSim(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed, threads=THREADS)
        o.run_ticks(CONFIG4_TICKS)
        st = o.st
        want = {"pos": st.pos, "tgt": st.tgt, "dir": st.dir, "road": st.road, "rng": st.rng, "ncount": st.ncount}
        for name, (b, (tot, _)) in batches.items():
            assert_state_equal(b.state(s), want, f"{name}, scenario {s} ({n[s]} agents) vs oracle")
            assert int(tot[s]) == o.pairs_total, f"{name}, scenario {s}: pair total vs oracle"
        say(f"scenario {s} ({n[s]} agents) equals the oracle")
    for b, _ in batches.values():
        b.close()


CONFIG5_TICKS = 1000
CONFIG5_ALPHA = 0.02


@pytest.fixture(scope="module")
def config5():
    """2 full calibration scenarios, their multipliers, and the oracle's
    binary32 record of the 1,000-tick run with its trajectory."""
    orc.diff_set_threads(THREADS)
    scs = S.adjoint_config(2)
    roads, conns, ents, speed, _, _, off = S.pack_scenarios(scs)
    rng = np.random.default_rng(1)
    theta = np.exp(rng.normal(0.0, 0.2, roads.shape[0])).astype(np.float32)
    t0 = time.perf_counter()
    fin, sched, traj = orc.diff_record(roads, conns, ents, speed, off, theta, CONFIG5_ALPHA, 1.0, CONFIG5_TICKS,
                                       trajectory=True)
    say(f"oracle binary32 record: {ents.shape[0]} agents, {CONFIG5_TICKS} ticks, {int((sched.flag & 1).sum())} "
        f"arrivals, {time.perf_counter() - t0:.1f} s on {THREADS} threads")
    target = fin.astype(np.float64) + rng.normal(0.0, 1.0, fin.shape)
    return dict(scs=scs, p0=ents["pos"], speed=speed, off=off, theta=theta, fin=fin, sched=sched, traj=traj,
                target=target)


def gpu_gradient(c, arrival_gradient):
    """GPU run and gradient of L = sum |p_T - target|^2; the run's final
    positions and decisions must equal the oracle's record bit for bit."""
    b = msim.Batch(c["scs"])
    run = DiffRun(b, CONFIG5_TICKS, checkpoint_every=50, keep_schedule=True, arrival_gradient=arrival_gradient)
    th = torch.from_numpy(c["theta"]).to("cuda").requires_grad_(True)
    al = torch.tensor(CONFIG5_ALPHA, device="cuda", requires_grad=True)
    t0 = time.perf_counter()
    pos = run(th, al)
    loss = ((pos.double() - torch.from_numpy(c["target"]).to("cuda")) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    say(f"{arrival_gradient}: GPU forward + backward {time.perf_counter() - t0:.2f} s")
    np.testing.assert_array_equal(pos.detach().cpu().numpy().view(np.uint32), c["fin"].view(np.uint32))
    sched = c["sched"]
    for t in range(CONFIG5_TICKS + 1):
        tg, rd, fl = b.diff_schedule(t)
        if t < CONFIG5_TICKS:
            want = (sched.tgt[t], sched.road[t], sched.flag[t])
        else:
            want = (sched.final_tgt, sched.final_road, np.zeros_like(fl))
        assert np.array_equal(tg.view(np.uint32), want[0].view(np.uint32)), f"tick {t}: targets"
        assert np.array_equal(rd, want[1]) and np.array_equal(fl, want[2]), f"tick {t}: roads or flags"
    gt, ga = th.grad.double().cpu().numpy(), float(al.grad)
    b.close()
    return gt, ga


@pytest.mark.parametrize("arrival_gradient", ["frozen", "carry"])
def test_config5_run_equals_oracle_and_gradient_matches_binary64_adjoint(config5, arrival_gradient):
    c = config5
    gt, ga = gpu_gradient(c, arrival_gradient)
    gpos = 2.0 * (c["fin"].astype(np.float64) - c["target"])
    t0 = time.perf_counter()
    gt64, ga64 = orc.diff_adjoint64(c["p0"], c["speed"], c["off"], c["theta"].astype(np.float64), CONFIG5_ALPHA, 1.0,
                                    c["sched"], gpos, carry=arrival_gradient == "carry", traj32=c["traj"])
    rel = np.linalg.norm(gt - gt64) / np.linalg.norm(gt64)
    rel_a = abs(ga - ga64) / abs(ga64)
    say(f"{arrival_gradient}: binary64 adjoint around the binary32 run {time.perf_counter() - t0:.1f} s; "
        f"dL/dtheta relative {rel:.2e}, dL/dalpha {ga:.8g} vs {ga64:.8g} (relative {rel_a:.2e})")
    assert rel < 1e-3
    assert rel_a < 1e-3


def test_config5_binary64_adjoint_is_the_derivative_of_its_replay(config5):
    """The host adjoint at this workload's size against central differences
    of the binary64 replay under the recorded decisions (eps = 1e-5)."""
    c = config5
    p0, sp, off, sched, tgt = c["p0"], c["speed"], c["off"], c["sched"], c["target"]
    th64 = c["theta"].astype(np.float64)
    t0 = time.perf_counter()
    base = orc.diff_replay64(p0, sp, off, th64, CONFIG5_ALPHA, 1.0, sched)
    gt, ga = orc.diff_adjoint64(p0, sp, off, th64, CONFIG5_ALPHA, 1.0, sched, 2.0 * (base - tgt))

    def loss64(t, a):
        q = orc.diff_replay64(p0, sp, off, t, a, 1.0, sched)
        return float(((q - tgt) ** 2).sum())

    v = np.random.default_rng(8).normal(0.0, 1.0, th64.shape)
    eps = 1e-5
    fd = (loss64(th64 + eps * v, CONFIG5_ALPHA) - loss64(th64 - eps * v, CONFIG5_ALPHA)) / (2 * eps)
    fda = (loss64(th64, CONFIG5_ALPHA + eps) - loss64(th64, CONFIG5_ALPHA - eps)) / (2 * eps)
    say(f"binary64 adjoint and differences {time.perf_counter() - t0:.1f} s: direction {gt @ v:.8g} vs {fd:.8g}, "
        f"alpha {ga:.8g} vs {fda:.8g}")
    assert abs(gt @ v - fd) < 1e-3 * abs(fd)
    assert abs(ga - fda) < 1e-3 * abs(fda)
