"""The differentiable tick and its hand-written adjoint on the GPU
(msim_batch_diff_* through msim.adjoint.DiffRun).

* theta = 1, alpha = 0 reproduces the plain batched run bit for bit;
* with alpha = 0 and any theta the run and its recorded decisions equal the
  oracle's binary32 restatement bit for bit;
* a replay under the recorded decisions reproduces the run;
* dL/dtheta and dL/dalpha agree with the binary64 host adjoint of the same
  frozen-decision run (relative 1e-3), with central differences (eps 1e-3) of
  its binary64 replay (1e-3) and of the GPU's own binary32 replay (1e-2: the
  binary32 loss carries rounding noise of ~1e-3 of the difference).  This is
  checked on sparse scenarios (about 60 agents per map, like the calibration
  workload's 0.09 agents per metre of road); on dense ones (1,000 agents per
  small map, several neighbours each) the pushes amplify perturbations over
  100 ticks, eps = 1e-3 differences are off by up to 1e-2 (they converge to
  the adjoint only as eps -> 0, see test_oracle_adjoint.py), and dL/dtheta is
  compared with the host adjoint alone;
* the checkpoint interval does not change the gradient;
* the carry-over surrogate through arrivals agrees with the host's binary64
  surrogate adjoint (the carry-over itself is pinned on the CPU against
  differences of the carried run, test_oracle_adjoint.py).
"""
import numpy as np
import pytest

import msim
import orc
from msim import scenarios as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from msim.adjoint import DiffRun  # noqa: E402

DEV = "cuda"


def problem(n_scen=3, scale=0.01, seed=5):
    scs = S.batch_config(n_scen, scale=scale, seed=seed)
    return scs, S.pack_scenarios(scs)


def theta_for(n_roads, seed=1, sigma=0.2):
    rng = np.random.default_rng(seed)
    return np.exp(rng.normal(0.0, sigma, n_roads)).astype(np.float32)


def schedule(batch, ticks):
    tg = np.zeros((ticks, batch.n, 2), np.float32)
    rd = np.zeros((ticks, batch.n), np.uint32)
    fl = np.zeros((ticks, batch.n), np.uint8)
    for t in range(ticks):
        tg[t], rd[t], fl[t] = batch.diff_schedule(t)
    ft, fr, ff = batch.diff_schedule(ticks)  # the decisions after the run
    assert not ff.any()
    return orc.DiffSchedule(tg, rd, fl, ft, fr)


def test_neutral_parameters_are_the_plain_tick():
    scs, _ = problem(4, seed=31)
    T = 130
    plain = msim.Batch(scs)
    plain.run_ticks(T)
    b = msim.Batch(scs)
    run = DiffRun(b, T, checkpoint_every=50)
    pos = run(torch.ones(b.n_roads, device=DEV), torch.tensor(0.0, device=DEV)).cpu().numpy()
    for s in range(len(scs)):
        d, p = b.state(s), plain.state(s)
        for k in ("pos", "tgt", "road", "rng"):
            np.testing.assert_array_equal(np.asarray(d[k]).view(np.uint32), np.asarray(p[k]).view(np.uint32),
                                          err_msg=f"scenario {s}: {k}")
        a0, a1 = b.agent_off[s], b.agent_off[s + 1]
        np.testing.assert_array_equal(pos[a0:a1].view(np.uint32), p["pos"].view(np.uint32))


def test_record_equals_oracle_without_avoidance():
    scs, (roads, conns, ents, speed, _, _, off) = problem(seed=7)
    T = 110
    theta = theta_for(roads.shape[0], seed=3)
    b = msim.Batch(scs)
    run = DiffRun(b, T, checkpoint_every=40, keep_schedule=True)
    pos = run(torch.from_numpy(theta).to(DEV), torch.tensor(0.0, device=DEV)).cpu().numpy()
    fin, sched = orc.diff_record(roads, conns, ents, speed, off, theta, 0.0, 1.0, T)
    np.testing.assert_array_equal(pos.view(np.uint32), fin.view(np.uint32))
    g = schedule(b, T)
    np.testing.assert_array_equal(g.flag, sched.flag)
    np.testing.assert_array_equal(g.road, sched.road)
    np.testing.assert_array_equal(g.tgt.view(np.uint32), sched.tgt.view(np.uint32))


def test_avoidance_run_equals_oracle_and_replay_exact():
    """With avoidance the run equals the oracle's binary32 record bit for bit
    (both sum the push terms in fixed point), decisions included."""
    scs, (roads, conns, ents, speed, _, _, off) = problem(seed=9)
    T = 100
    theta = theta_for(roads.shape[0], seed=4)
    alpha = 0.05
    b = msim.Batch(scs)
    run = DiffRun(b, T, checkpoint_every=30, keep_schedule=True)
    th = torch.from_numpy(theta).to(DEV)
    pos = run(th, torch.tensor(alpha, device=DEV)).cpu().numpy()
    fin, sched = orc.diff_record(roads, conns, ents, speed, off, theta, alpha, 1.0, T)
    assert not np.array_equal(fin, orc.diff_record(roads, conns, ents, speed, off, theta, 0.0, 1.0, T)[0])
    np.testing.assert_array_equal(pos.view(np.uint32), fin.view(np.uint32))
    g = schedule(b, T)
    np.testing.assert_array_equal(g.flag, sched.flag)
    np.testing.assert_array_equal(g.road, sched.road)
    np.testing.assert_array_equal(g.tgt.view(np.uint32), sched.tgt.view(np.uint32))
    rep = run.replay(th, alpha).cpu().numpy()
    np.testing.assert_array_equal(rep.view(np.uint32), pos.view(np.uint32))


def gradients(ck, alpha=0.05, T=100, seed=5, arrival_gradient="frozen"):
    scs, (roads, conns, ents, speed, _, _, off) = problem(seed=seed)
    theta = theta_for(roads.shape[0], seed=6)
    b = msim.Batch(scs)
    run = DiffRun(b, T, checkpoint_every=ck, keep_schedule=True, arrival_gradient=arrival_gradient)
    th = torch.from_numpy(theta).to(DEV).requires_grad_(True)
    al = torch.tensor(alpha, device=DEV, requires_grad=True)
    pos = run(th, al)
    rng = np.random.default_rng(2)
    target = pos.detach() + torch.from_numpy(rng.normal(0.0, 1.0, (b.n, 2)).astype(np.float32)).to(DEV)
    loss = ((pos - target) ** 2).sum()
    loss.backward()
    return dict(b=b, run=run, theta=theta, alpha=alpha, target=target, pos=pos.detach(), gt=th.grad.double().cpu().numpy(),
                ga=float(al.grad), ents=ents, speed=speed, off=off, T=T)


def test_adjoint_matches_fp64_host_adjoint_and_differences():
    r = gradients(ck=30, seed=5)
    sched = schedule(r["b"], r["T"])
    p0 = r["ents"]["pos"]
    th64 = r["theta"].astype(np.float64)
    gpos = 2.0 * (r["pos"].cpu().numpy().astype(np.float64) - r["target"].cpu().numpy())
    gt64, ga64 = orc.diff_adjoint64(p0, r["speed"], r["off"], th64, r["alpha"], 1.0, sched, gpos)
    rel = np.linalg.norm(r["gt"] - gt64) / np.linalg.norm(gt64)
    tgt64 = r["target"].cpu().numpy().astype(np.float64)

    def loss64(t, a):
        q = orc.diff_replay64(p0, r["speed"], r["off"], t, a, 1.0, sched)
        return float(((q - tgt64) ** 2).sum())

    v = np.random.default_rng(8).normal(0.0, 1.0, th64.shape)
    eps = 1e-3
    fd = (loss64(th64 + eps * v, r["alpha"]) - loss64(th64 - eps * v, r["alpha"])) / (2 * eps)
    fda = (loss64(th64, r["alpha"] + eps) - loss64(th64, r["alpha"] - eps)) / (2 * eps)
    print(f"dL/dtheta vs binary64 host adjoint: relative {rel:.2e}; dL/dalpha {r['ga']:.6g} vs {ga64:.6g} "
          f"(binary64 central difference {fda:.6g}); directional derivative: adjoint {r['gt'] @ v:.6g}, "
          f"host {gt64 @ v:.6g}, binary64 central difference {fd:.6g}")
    assert rel < 1e-3
    assert abs(r["ga"] - ga64) < 1e-3 * abs(ga64)
    assert abs(r["gt"] @ v - fd) < 1e-3 * abs(fd)
    assert abs(r["ga"] - fda) < 1e-3 * abs(fda)

    # central differences of the GPU's own binary32 replay
    run = r["run"]

    def loss32(t, a):
        q = run.replay(torch.from_numpy(t.astype(np.float32)).to(DEV), a).double().cpu().numpy()
        return float(((q - tgt64) ** 2).sum())

    th32 = r["theta"].astype(np.float64)
    fd32 = (loss32(th32 + eps * v, r["alpha"]) - loss32(th32 - eps * v, r["alpha"])) / (2 * eps)
    print(f"binary32 replay central difference {fd32:.6g}")
    assert abs(r["gt"] @ v - fd32) < 1e-2 * abs(fd32)


def test_dense_theta_gradient_matches_fp64_host_adjoint():
    r = gradients(ck=30, seed=11)
    sched = schedule(r["b"], r["T"])
    gpos = 2.0 * (r["pos"].cpu().numpy().astype(np.float64) - r["target"].cpu().numpy())
    gt64, _ = orc.diff_adjoint64(r["ents"]["pos"], r["speed"], r["off"], r["theta"].astype(np.float64), r["alpha"],
                                 1.0, sched, gpos)
    rel = np.linalg.norm(r["gt"] - gt64) / np.linalg.norm(gt64)
    print(f"dense: dL/dtheta vs binary64 host adjoint: relative {rel:.2e} ({r['b'].n} agents)")
    assert rel < 1e-3


def test_carry_over_gradient_matches_fp64_host():
    """The arrival surrogate (each arrival differentiated as carrying its
    leftover time onto the next road) against the binary64 host adjoint with
    the same surrogate on the same snapped run; it also reaches the roads
    agents left before the end, which the frozen gradient cannot."""
    r = gradients(ck=30, seed=5, arrival_gradient="carry")
    sched = schedule(r["b"], r["T"])
    assert (sched.flag & 1).sum() > 20
    gpos = 2.0 * (r["pos"].cpu().numpy().astype(np.float64) - r["target"].cpu().numpy())
    p0, th64 = r["ents"]["pos"], r["theta"].astype(np.float64)
    gt64, ga64 = orc.diff_adjoint64(p0, r["speed"], r["off"], th64, r["alpha"], 1.0, sched, gpos, carry=True)
    rel = np.linalg.norm(r["gt"] - gt64) / np.linalg.norm(gt64)
    print(f"carry-over: dL/dtheta vs binary64 host: relative {rel:.2e}; dL/dalpha {r['ga']:.6g} vs {ga64:.6g}")
    assert rel < 1e-3
    assert abs(r["ga"] - ga64) < 1e-3 * abs(ga64)
    frozen, _ = orc.diff_adjoint64(p0, r["speed"], r["off"], th64, r["alpha"], 1.0, sched, gpos)
    assert np.linalg.norm(gt64 - frozen) > 1e-3 * np.linalg.norm(frozen)
    last = np.unique(np.concatenate([sched.road[-1], sched.final_road]))
    assert np.count_nonzero(np.delete(r["gt"], last)) > np.count_nonzero(np.delete(frozen, last))


def test_checkpoint_interval_does_not_change_the_gradient():
    a = gradients(ck=7)
    b = gradients(ck=100)
    np.testing.assert_allclose(a["gt"], b["gt"], rtol=1e-9, atol=1e-9 * np.abs(b["gt"]).max())
    assert abs(a["ga"] - b["ga"]) <= 1e-9 * abs(b["ga"])


def test_plain_ticks_refused_after_setup():
    scs, _ = problem(2, seed=13)
    b = msim.Batch(scs)
    DiffRun(b, 10)
    with pytest.raises(msim.MsimError):
        b.run_ticks(1)


def test_forward_with_avoidance_is_bit_stable():
    """Two forward runs (and their backward passes) at alpha = 0.05 on dense
    scenarios are bit-identical over 200 ticks: the push sums are
    accumulated in fixed point, so their order does not matter."""
    scs, _ = problem(seed=11)
    T = 200
    b = msim.Batch(scs)
    theta = torch.from_numpy(theta_for(b.n_roads, seed=6)).to(DEV)
    run = DiffRun(b, T, checkpoint_every=40)
    runs = []
    for k in range(2):
        th = theta.clone().requires_grad_(True)
        al = torch.tensor(0.05, device=DEV, requires_grad=True)
        pos = run(th, al)
        (pos ** 2).sum().backward()
        runs.append((pos.detach().cpu().numpy(), th.grad.cpu().numpy(), float(al.grad)))
    np.testing.assert_array_equal(runs[0][0].view(np.uint32), runs[1][0].view(np.uint32))
    np.testing.assert_array_equal(runs[0][1].view(np.uint32), runs[1][1].view(np.uint32))
    assert runs[0][2] == runs[1][2]
    # the pushes act: without them most agents end elsewhere
    free = run(theta, torch.tensor(0.0, device=DEV)).cpu().numpy()
    moved = (free != runs[0][0]).any(axis=1).mean()
    print(f"{b.n} agents; the pushes changed {moved:.1%} of the final positions")
    assert moved > 0.2


def test_backward_of_an_older_forward_raises():
    scs, _ = problem(2, seed=13)
    b = msim.Batch(scs)
    run = DiffRun(b, 40, checkpoint_every=20)
    th = torch.ones(b.n_roads, device=DEV, requires_grad=True)
    first = run(th, torch.tensor(0.05, device=DEV))
    second = run(th * 1.1, torch.tensor(0.05, device=DEV))
    with pytest.raises(RuntimeError, match="another forward run"):
        first.sum().backward(retain_graph=True)
    th.grad = None
    second.sum().backward()
    assert torch.isfinite(th.grad).all()


def test_carry_over_with_zero_multipliers_is_finite():
    """Roads with theta = 0 hold their agents still; under the carry-over
    surrogate their arrivals fall back to the frozen adjoint (no division by
    a zero speed), as in the binary64 host adjoint."""
    scs, (roads, conns, ents, speed, _, _, off) = problem(seed=5)
    T = 100
    theta = theta_for(roads.shape[0], seed=6)
    theta[::7] = 0.0
    b = msim.Batch(scs)
    run = DiffRun(b, T, checkpoint_every=30, keep_schedule=True, arrival_gradient="carry")
    th = torch.from_numpy(theta).to(DEV).requires_grad_(True)
    al = torch.tensor(0.05, device=DEV, requires_grad=True)
    pos = run(th, al)
    target = pos.detach() + 1.0
    ((pos - target) ** 2).sum().backward()
    gt = th.grad.double().cpu().numpy()
    assert np.isfinite(gt).all() and np.isfinite(float(al.grad))
    sched = schedule(b, T)
    gpos = 2.0 * (pos.detach().double().cpu().numpy() - target.double().cpu().numpy())
    gt64, ga64 = orc.diff_adjoint64(ents["pos"], speed, off, theta.astype(np.float64), 0.05, 1.0, sched, gpos,
                                    carry=True)
    assert np.isfinite(gt64).all()
    rel = np.linalg.norm(gt - gt64) / np.linalg.norm(gt64)
    print(f"zero multipliers on {int((theta == 0).sum())} roads: relative {rel:.2e}")
    assert rel < 1e-3
    assert abs(float(al.grad) - float(ga64)) < 1e-3 * abs(float(ga64))
