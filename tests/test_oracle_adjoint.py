"""The oracle's differentiable tick and its binary64 adjoint, pinned on the
CPU: with theta = 1 and alpha = 0 the recording run is the oracle's plain tick;
the binary64 adjoint agrees with central finite differences of the binary64
replay, and the finite differences converge to it at the rate of a central
difference (error ~ eps^2)."""
import numpy as np
import pytest

import orc
from msim import scenarios as S


def packed(n_scen=3, scale=0.01, seed=5):
    scs = S.batch_config(n_scen, scale=scale, seed=seed)
    return scs, S.pack_scenarios(scs)


def test_neutral_record_is_the_plain_tick():
    scs, (roads, conns, ents, speed, _, _, off) = packed()
    theta = np.ones(roads.shape[0], np.float32)
    fin, sched = orc.diff_record(roads, conns, ents, speed, off, theta, 0.0, 1.0, 90)
    for s, sc in enumerate(scs):
        o = orc.OracleSim(sc.roads, sc.conns, sc.entities, sc.pc, speed=sc.speed)
        o.run_ticks(90)
        a, b = off[s], off[s + 1]
        np.testing.assert_array_equal(fin[a:b].view(np.uint32), o.st.pos.view(np.uint32))
    assert sched.flag[0].min() == 2 and sched.flag[1:].max() <= 1  # insert-only first tick


def _problem(alpha=0.05, ticks=80, seed=1):
    scs, (roads, conns, ents, speed, _, _, off) = packed()
    rng = np.random.default_rng(seed)
    theta = np.exp(rng.normal(0.0, 0.2, roads.shape[0])).astype(np.float32)
    fin, sched = orc.diff_record(roads, conns, ents, speed, off, theta, alpha, 1.0, ticks)
    p0 = ents["pos"].copy()
    th = theta.astype(np.float64)
    p = orc.diff_replay64(p0, speed, off, th, alpha, 1.0, sched)
    target = p + rng.normal(0.0, 1.0, p.shape)

    def loss(t, a):
        q = orc.diff_replay64(p0, speed, off, t, a, 1.0, sched)
        return float(((q - target) ** 2).sum())

    gt, ga = orc.diff_adjoint64(p0, speed, off, th, alpha, 1.0, sched, 2.0 * (p - target))
    return th, alpha, loss, gt, ga, rng


def test_fp64_adjoint_matches_central_differences():
    th, alpha, loss, gt, ga, rng = _problem()
    v = rng.normal(0.0, 1.0, th.shape)
    eps = 1e-3
    fd = (loss(th + eps * v, alpha) - loss(th - eps * v, alpha)) / (2 * eps)
    assert abs(gt @ v - fd) <= 1e-3 * abs(fd)
    fda = (loss(th, alpha + eps) - loss(th, alpha - eps)) / (2 * eps)
    assert abs(ga - fda) <= 1e-3 * abs(fda)


def test_fp64_adjoint_is_the_limit_of_central_differences():
    th, alpha, loss, gt, _, _ = _problem()
    for r in np.argsort(-np.abs(gt))[:3]:
        e = np.zeros_like(th)
        e[r] = 1.0
        errs = []
        for eps in (1e-3, 1e-4):
            fd = (loss(th + eps * e, alpha) - loss(th - eps * e, alpha)) / (2 * eps)
            errs.append(abs(fd - gt[r]) / abs(gt[r]))
        assert errs[1] < 1e-5 and (errs[0] < 1e-9 or errs[1] < 0.05 * errs[0]), (r, errs)


def test_alpha_zero_gradient_is_the_pushes_first_order_effect():
    # at alpha = 0 the alpha derivative is sum(lam . A) over the run
    th, _, loss, gt, ga, _ = _problem(alpha=0.0)
    eps = 1e-4
    fda = (loss(th, eps) - loss(th, -eps)) / (2 * eps)
    assert abs(ga - fda) <= 1e-4 * max(1.0, abs(fda))


def _carry_problem(alpha=0.05, ticks=80, seed=2):
    scs, (roads, conns, ents, speed, _, _, off) = packed()
    rng = np.random.default_rng(seed)
    theta = np.exp(rng.normal(0.0, 0.2, roads.shape[0])).astype(np.float32)
    fin, sched = orc.diff_record(roads, conns, ents, speed, off, theta, alpha, 1.0, ticks)
    p0 = ents["pos"].copy()
    th = theta.astype(np.float64)
    p = orc.diff_replay64_carry(p0, speed, off, th, alpha, 1.0, sched)
    target = p + rng.normal(0.0, 1.0, p.shape)

    def loss(t, a):
        q = orc.diff_replay64_carry(p0, speed, off, t, a, 1.0, sched)
        return float(((q - target) ** 2).sum())

    gt, ga = orc.diff_adjoint64(p0, speed, off, th, alpha, 1.0, sched, 2.0 * (p - target), carry=True,
                                carry_traj=True)
    return sched, th, alpha, loss, gt, ga, rng


def test_carry_over_adjoint_matches_differences_of_the_carried_replay():
    """The arrival surrogate differentiates the run in which arrivals carry
    their leftover time onto the next road: on that run's own trajectory it
    is its exact gradient.  That run follows the snapped run's decisions, so
    a carried agent can pass its next target before the decisions say it
    arrives, and the moves then turn sharply: the differences need a small
    step (eps 1e-6; at 1e-4 some roads are off by 2x, at 1e-6 they agree to
    1e-8 relative)."""
    sched, th, alpha, loss, gt, ga, rng = _carry_problem()
    assert (sched.flag & 1).sum() > 50  # arrivals happen
    eps = 1e-6
    for r in np.argsort(-np.abs(gt))[:8]:
        e = np.zeros_like(th)
        e[r] = eps
        fd = (loss(th + e, alpha) - loss(th - e, alpha)) / (2 * eps)
        assert abs(gt[r] - fd) <= 1e-5 * abs(fd), (r, gt[r], fd)
    v = rng.normal(0.0, 1.0, th.shape)
    fd = (loss(th + eps * v, alpha) - loss(th - eps * v, alpha)) / (2 * eps)
    assert abs(gt @ v - fd) <= 1e-4 * abs(fd)
    fda = (loss(th, alpha + eps) - loss(th, alpha - eps)) / (2 * eps)
    assert abs(ga - fda) <= 1e-4 * abs(fda)
    # the roads agents left before the end receive gradient too
    last = np.unique(np.concatenate([sched.road[-1], sched.final_road]))
    assert np.count_nonzero(np.delete(gt, last)) > 0


def test_carry_over_without_arrivals_is_the_frozen_adjoint():
    scs, (roads, conns, ents, speed, _, _, off) = packed()
    theta = np.ones(roads.shape[0], np.float32)
    _, rec = orc.diff_record(roads, conns, ents, speed, off, theta, 0.05, 1.0, 6)
    # the same decisions with every arrival taken out: nothing to carry
    sched = orc.DiffSchedule(rec.tgt, rec.road, rec.flag & 2, rec.final_tgt, rec.final_road)
    p0 = ents["pos"].copy()
    g = np.random.default_rng(4).normal(0.0, 1.0, (p0.shape[0], 2))
    a = orc.diff_adjoint64(p0, speed, off, theta.astype(np.float64), 0.05, 1.0, sched, g)
    b = orc.diff_adjoint64(p0, speed, off, theta.astype(np.float64), 0.05, 1.0, sched, g, carry=True)
    np.testing.assert_array_equal(a[0], b[0])
    assert a[1] == b[1]


def _numpy_push(pt, off, alpha, radius):
    """p = p~ + alpha A(p~), restated with numpy: brute-force in-range pairs
    of each scenario (binary32 |x|, |y| <= R boxes and sqrt(x^2 + y^2) < R),
    w = (1 - |x|^2 / R^2)^3 in binary32, each term w x rounded to a multiple
    of 2^-40 and the terms summed as integers."""
    sc = orc.diff_fix_scale(radius)
    assert sc == 2.0 ** 40
    r = np.float32(radius)
    inv_r2 = np.float32(1.0) / (r * r)
    out = np.empty_like(pt)
    for s in range(off.shape[0] - 1):
        p = pt[off[s]:off[s + 1]]
        x = p[:, None, 0] - p[None, :, 0]
        y = p[:, None, 1] - p[None, :, 1]
        near = (np.abs(x) <= r) & (np.abs(y) <= r) & (np.sqrt(x * x + y * y) < r)
        np.fill_diagonal(near, False)
        h = np.float32(1.0) - (x * x + y * y) * inv_r2
        w = h * h * h
        ix = np.where(near, np.rint((w * x).astype(np.float64) * sc), 0.0).astype(np.int64).sum(axis=1)
        iy = np.where(near, np.rint((w * y).astype(np.float64) * sc), 0.0).astype(np.int64).sum(axis=1)
        ax = (ix.astype(np.float64) / sc).astype(np.float32)
        ay = (iy.astype(np.float64) / sc).astype(np.float32)
        out[off[s]:off[s + 1], 0] = p[:, 0] + np.float32(alpha) * ax
        out[off[s]:off[s + 1], 1] = p[:, 1] + np.float32(alpha) * ay
    return out


def test_record_push_is_the_fixed_point_sum_and_thread_independent():
    """Each tick's push, from the recorded post-move positions, equals a
    brute-force numpy restatement of the fixed-point sum; the record is
    bit-identical with 1 and 4 threads."""
    scs, (roads, conns, ents, speed, _, _, off) = packed(seed=11)
    theta = np.exp(np.random.default_rng(2).normal(0.0, 0.2, roads.shape[0])).astype(np.float32)
    alpha, T = 0.05, 30
    runs = []
    try:
        for threads in (1, 4):
            orc.diff_set_threads(threads)
            runs.append(orc.diff_record(roads, conns, ents, speed, off, theta, alpha, 1.0, T, trajectory=True))
    finally:
        orc.diff_set_threads(1)
    (fin, sched, traj), (fin4, sched4, traj4) = runs
    np.testing.assert_array_equal(fin.view(np.uint32), fin4.view(np.uint32))
    np.testing.assert_array_equal(traj.view(np.uint32), traj4.view(np.uint32))
    np.testing.assert_array_equal(sched.flag, sched4.flag)
    pushed = 0
    for t in range(T):
        want = _numpy_push(traj[t, 1], off, alpha, 1.0)
        got = traj[t + 1, 0] if t + 1 < T else fin
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32), err_msg=f"tick {t}")
        pushed += int((want != traj[t, 1]).any(axis=1).sum())
    assert pushed > 100  # the pushes act


def test_adjoint_around_the_binary32_trajectory():
    """Linearised around the record's own binary32 trajectory, the adjoint
    agrees with the one around the binary64 replay to the two trajectories'
    difference on a short run, in both arrival modes (the positions differ by
    up to 3e-4 m after 10 ticks; the difference of the gradients grows with
    that drift, from 2e-6 after 2 ticks to 1e-5 after 10)."""
    scs, (roads, conns, ents, speed, _, _, off) = packed(seed=11)
    rng = np.random.default_rng(4)
    theta = np.exp(rng.normal(0.0, 0.2, roads.shape[0])).astype(np.float32)
    alpha, T = 0.05, 10
    fin, sched, traj = orc.diff_record(roads, conns, ents, speed, off, theta, alpha, 1.0, T, trajectory=True)
    p0, th = ents["pos"], theta.astype(np.float64)
    g = rng.normal(0.0, 1.0, fin.shape)
    for carry in (False, True):
        a = orc.diff_adjoint64(p0, speed, off, th, alpha, 1.0, sched, g, carry=carry)
        b = orc.diff_adjoint64(p0, speed, off, th, alpha, 1.0, sched, g, carry=carry, traj32=traj)
        rel = np.linalg.norm(a[0] - b[0]) / np.linalg.norm(a[0])
        assert 0.0 < rel < 1e-4, rel
        assert abs(a[1] - b[1]) < 1e-3 * abs(a[1])
    with pytest.raises(ValueError):
        orc.diff_adjoint64(p0, speed, off, th, alpha, 1.0, sched, g, traj32=traj[:-1])
