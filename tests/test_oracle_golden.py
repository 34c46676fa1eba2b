"""Pins the CPU oracle before it is trusted as the checker.

Each check comes from outside the oracle's own formulas: the published
xorshift128 golden vector (Marsaglia 2003), an independent exact-rational
model of IEEE binary32 rounding, closed forms, brute force, the reference's
known-answer collision test (shader_validation/src/main.cpp:1137-1199) and,
where it was built, the reference's own CPU quadtree code.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import orc
from msim import scenarios as S

# ---- independent binary32 model -----------------------------------------


def fl32(x):
    """Round an exact rational to binary32, round-half-to-even (normal range)."""
    x = Fraction(x)
    if x == 0:
        return Fraction(0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    q = a / Fraction(2) ** (e - 23)  # in [2^23, 2^24)
    m = q.numerator // q.denominator
    rem = q - m
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and m % 2 == 1):
        m += 1
    return sign * Fraction(m) * Fraction(2) ** (e - 23)


def state_emitting(u):
    """An xorshift128 state (shader order x,y,z,w) whose next output is u:
    with w = 0 the output is x ^ (x >> 19), inverted by x = u ^ (u >> 19)."""
    return np.array([u ^ (u >> 19), 7, 11, 0], np.uint32)


# ---- A4: xorshift128 ------------------------------------------------------


def test_xorshift_published_golden_vector():
    # Marsaglia's xor128 seeds (x,y,z,w) = (123456789, 362436069, 521288629,
    # 88675123); the shader keeps the newest word in .x, so the same stream
    # starts from uvec4(88675123, 521288629, 362436069, 123456789).
    s = np.array([88675123, 521288629, 362436069, 123456789], np.uint32)
    out = []
    for _ in range(4):
        v, s = orc.rng_next(s)
        out.append(v)
    assert out == [3701687786, 458299110, 2500872618, 3633119408]
    assert s.tolist() == [3633119408, 2500872618, 458299110, 3701687786]


def _marsaglia_xor128(x, y, z, w, n):
    out = []
    m = 0xFFFFFFFF
    for _ in range(n):
        t = (x ^ (x << 11)) & m
        x, y, z = y, z, w
        w = (w ^ (w >> 19) ^ (t ^ (t >> 8))) & m
        out.append(w)
    return out


def test_xorshift_matches_published_algorithm_random_seeds():
    rng = np.random.default_rng(1)
    for _ in range(200):
        x, y, z, w = (int(v) for v in rng.integers(1, 2**32, 4))
        want = _marsaglia_xor128(x, y, z, w, 50)
        s = np.array([w, z, y, x], np.uint32)
        got = []
        for _ in range(50):
            v, s = orc.rng_next(s)
            got.append(v)
        assert got == want


# ---- A5/A6: next_float, next(state, 1, cnt) --------------------------------


@pytest.mark.parametrize("u,want", [
    (0, 0.0),
    (2**31, 0.5),
    (4294967167, float(np.float32(0.99999994))),
    (4294967168, 1.0),
    (4294967295, 1.0),
])
def test_rng_float_edges(u, want):
    f, _ = orc.rng_float(state_emitting(u))
    assert float(f) == want
    assert Fraction(float(f)) == fl32(u) / 2**32


def _expected_offset(u, cnt):
    f = fl32(u) / 2**32
    prod = fl32(f * cnt)
    s = fl32(1 + prod)
    return math.ceil(s) - 1


def test_rng_range_against_exact_rounding_model():
    rng = np.random.default_rng(2)
    us = list(rng.integers(0, 2**32, 3000)) + [0, 1, 2**31, 4294967167, 4294967168, 4294967295]
    for cnt in (3, 4, 5, 6, 7, 13, 100):
        # values of u where f * cnt sits on or next to an integer
        for k in range(cnt + 1):
            c = (k * 2**32) // cnt
            us.extend([max(0, c - 300), min(c, 2**32 - 1), min(2**32 - 1, c + 300)])
        for u in us:
            u = int(u)
            got, _ = orc.rng_range(state_emitting(u), 1, cnt)
            assert got == _expected_offset(u, cnt), (u, cnt)
            assert 0 <= got <= cnt


def test_rng_range_band_structure():
    # Away from rounding boundaries, offset k <=> f in ((k-1)/cnt, k/cnt].
    rng = np.random.default_rng(3)
    for cnt in (3, 4, 6):
        for u in rng.integers(0, 2**32, 2000):
            u = int(u)
            f = Fraction(float(np.float32(u))) / 2**32
            k = math.ceil(f * cnt)
            if abs(f * cnt - round(f * cnt)) < Fraction(1, 2**20):
                continue
            got, _ = orc.rng_range(state_emitting(u), 1, cnt)
            assert got == k
    assert orc.rng_range(state_emitting(0), 1, 4)[0] == 0  # 1 + 0*cnt == 1.0f
    assert orc.rng_range(state_emitting(4294967295), 1, 4)[0] == 4  # f == 1.0


def test_rng_range_distribution_top_offset():
    # Offset cnt (the read past the group, Q1) occurs with probability ~1/cnt.
    rng = np.random.default_rng(4)
    cnt = 4
    vals = np.array([orc.rng_range(state_emitting(int(u)), 1, cnt)[0]
                     for u in rng.integers(0, 2**32, 20000)])
    freq = np.bincount(vals, minlength=cnt + 1) / vals.size
    assert freq[0] < 1e-3
    np.testing.assert_allclose(freq[1:], 1 / cnt, atol=0.015)


# ---- A7/A8: update_direction and move ------------------------------------


def test_update_direction_closed_forms():
    s = np.float32(1.4)
    assert float(s) == 1.39999997615814208984375
    np.testing.assert_array_equal(orc.update_direction([5, 0], [0, 0], s), [s, 0])
    np.testing.assert_array_equal(orc.update_direction([0, -7], [0, 0], s), [0, -s])
    np.testing.assert_array_equal(orc.update_direction([2, 3], [2, 3], s), [0, 0])  # len == 0


def test_update_direction_error_vs_exact():
    rng = np.random.default_rng(5)
    for _ in range(2000):
        p = rng.uniform(-5000, 5000, 2).astype(np.float32)
        t = (p + rng.uniform(-400, 400, 2)).astype(np.float32)
        sp = np.float32(rng.uniform(0.8, 1.6))
        d = orc.update_direction(t, p, sp)
        dx = Fraction(float(t[0])) - Fraction(float(p[0]))
        dy = Fraction(float(t[1])) - Fraction(float(p[1]))
        ln = math.sqrt(dx * dx + dy * dy)
        for got, comp in ((d[0], dx), (d[1], dy)):
            exact = float(comp) / ln * float(sp)
            ulp = np.spacing(np.float32(abs(exact))) if exact else np.float32(1e-45)
            assert abs(float(got) - exact) <= 3 * ulp


def _move(pos, tgt, speed, roads, conns, rng_state=(1, 2, 3, 4), road=0):
    st = orc.State(_ents([pos], [tgt], [road], [rng_state]), np.array([speed], np.float32))
    st.init[:] = 1
    arr = np.zeros(1, np.uint64)
    rc = orc.L.orc_move_phase(1, orc._p(st.pos), orc._p(st.tgt), orc._p(st.dir), orc._p(st.road),
                              orc._p(st.rng), orc._p(st.speed), orc._p(st.init), orc._p(roads), roads.shape[0],
                              orc._p(conns), conns.shape[0], orc._p(arr))
    assert rc == 0
    return st, int(arr[0])


def _ents(pos, tgt, road, rng):
    e = np.zeros(len(pos), S.ENTITY_DTYPE)
    e["pos"] = pos
    e["target"] = tgt
    e["roadIndex"] = road
    e["randomState"] = rng
    e["initialized"] = 1
    return e


def test_move_snap_threshold():
    roads, conns, _, _ = S.star_map()
    s = np.float32(1.4)
    # dist == speed: snaps to the target (random_move.comp:844 uses '>')
    st, arr = _move([0, 0], [float(s), 0], s, roads, conns)
    assert arr == 1 and st.pos[0].tolist() == [float(s), 0.0]
    # dist one ulp above speed: ordinary step
    d = float(np.nextafter(s, np.float32(2)))
    st, arr = _move([0, 0], [d, 0], s, roads, conns)
    assert arr == 0 and st.pos[0].tolist() == [float(s), 0.0]


def test_route_choice_star_hand_derived():
    """Agent on road 0 arriving at the centre (count 4, group [0,1,2,3] at
    index 1).  With the golden state the first draw is u = 3701687786:
    f = 0.8618..., 1 + 4f = 4.447 -> offset 4 -> connections[1 + 4] = 1, the
    first entry of road 1's start group (the read past the group).  Road 1
    starts at (95,5) != the old target (50,50), so the new target is (95,5)."""
    roads, conns, _, _ = S.star_map()
    assert conns[1:5].tolist()[0] == 0 and conns[5] == 1
    golden = [88675123, 521288629, 362436069, 123456789]
    st, arr = _move([49.5, 49.5], [50, 50], np.float32(1.4), roads, conns, golden, road=0)
    assert arr == 1
    assert st.road[0] == 1
    assert st.tgt[0].tolist() == [95.0, 5.0]
    assert st.pos[0].tolist() == [50.0, 50.0]
    assert st.rng[0].tolist() == [3701687786, 88675123, 521288629, 362436069]


def test_route_choice_dead_end_and_count_two():
    roads, conns, _, _ = S.star_map()
    # leaf endpoint (count 1): turn around, same road, no draw
    st, arr = _move([5.5, 5.5], [5, 5], np.float32(1.4), roads, conns, (9, 8, 7, 6), road=0)
    assert arr == 1 and st.road[0] == 0 and st.tgt[0].tolist() == [50.0, 50.0]
    assert st.rng[0].tolist() == [9, 8, 7, 6]
    # count 2: connections[index + 1], no draw
    nodes = np.array([[0, 0], [10, 0], [20, 0]], np.float32)
    r2, c2 = S.build_roads(nodes, np.array([[0, 1], [1, 2]]), "own_first")
    st, arr = _move([9.5, 0], [10, 0], np.float32(1.4), r2, c2, (9, 8, 7, 6), road=0)
    assert arr == 1 and st.road[0] == 1 and st.tgt[0].tolist() == [20.0, 0.0]
    assert st.rng[0].tolist() == [9, 8, 7, 6]


def test_advance_error_vs_binary64():
    sc = S.config(1)
    p, t, sp = sc.entities["pos"], sc.entities["target"], sc.speed
    exact, arr64 = orc.advance_f64(p, t, sp)
    st = orc.State(sc.entities, sp)
    st.init[:] = 1
    orc.L.orc_move_phase(st.n, orc._p(st.pos), orc._p(st.tgt), orc._p(st.dir), orc._p(st.road),
                         orc._p(st.rng), orc._p(st.speed), orc._p(st.init), orc._p(sc.roads),
                         sc.roads.shape[0], orc._p(sc.conns), sc.conns.shape[0], None)
    moved = arr64 == 0
    err = np.abs(st.pos[moved].astype(np.float64) - exact[moved])
    # one rounding of the sum (half an ulp of the new position) plus the
    # heading's error (sqrt, divide, multiply: a few ulps of |dir| <= 1.6)
    ulp = np.spacing(np.abs(st.pos[moved])).astype(np.float64)
    bound = 0.5 * ulp + 3 * float(np.spacing(np.float32(1.6)))
    assert (err <= bound).all(), float((err / bound).max())


# ---- A12/A13: neighbour pairs ----------------------------------------------


def test_in_range_boundaries():
    assert not orc.in_range([1, 0], [0, 0], 1.0)  # distance == R is not < R
    assert orc.in_range([np.nextafter(np.float32(1), np.float32(0)), 0], [0, 0], 1.0)
    assert orc.in_range([3, 4], [3, 4], 1.0)  # coincident agents are a pair
    assert not orc.in_range([0.8, 0.8], [0, 0], 1.0)  # |dx|,|dy| <= R but sqrt(1.28) > 1
    assert orc.in_range([0.7, 0.7], [0, 0], 1.0)


def _random_cloud(rng, n, span, radius):
    p = rng.uniform(-span, span, (n, 2)).astype(np.float32)
    k = n // 10
    p[:k] = p[k:2 * k]  # duplicates
    # partners at exactly R and one ulp inside, along both axes
    for i in range(2 * k, 3 * k):
        j = i + k
        if j < n:
            p[j] = p[i] + np.array([radius, 0], np.float32)
    return p


@pytest.mark.parametrize("radius", [1.0, 10.0, 0.75, 2.5])
def test_grid_equals_brute_force(radius):
    rng = np.random.default_rng(int(radius * 100))
    for n, span in ((50, 3.0), (400, 10.0), (1500, 30.0)):
        p = _random_cloud(rng, n, span, radius)
        c1 = np.zeros(n, np.uint32)
        c2 = np.zeros(n, np.uint32)
        a = orc.L.orc_collide_grid(n, orc._p(p), radius, None, orc._p(c1))
        b = orc.L.orc_collide_brute(n, orc._p(p), radius, None, orc._p(c2))
        assert a == b
        np.testing.assert_array_equal(c1, c2)
        assert int(c1.sum()) == 2 * a


def test_grid_partial_search_rule():
    # agents with initialized == 0 do not search (random_move.comp:863-867):
    # pair (i, j), i > j, counts only if agent i searches.
    rng = np.random.default_rng(9)
    n = 600
    p = _random_cloud(rng, n, 8.0, 1.0)
    init = (rng.random(n) < 0.6).astype(np.uint32)
    c1, c2 = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    i1, i2 = init.copy(), init.copy()
    a = orc.L.orc_collide_grid(n, orc._p(p), 1.0, orc._p(i1), orc._p(c1))
    b = orc.L.orc_collide_brute(n, orc._p(p), 1.0, orc._p(i2), orc._p(c2))
    assert a == b and (c1 == c2).all() and (i1 == 1).all()
    # brute force by definition
    want = 0
    for i in range(n):
        for j in range(i):
            if init[i] and orc.in_range(p[j], p[i], 1.0):
                want += 1
    assert a == want


def test_reference_known_answer_restated():
    """shader_validation/src/main.cpp:1137-1199: R = 10, world 29007.4609 wide;
    agents at (W/2-3,0), (W/2+3,0), (0,0), (W/4,0), (W/4-9,0); the final pass
    finds the pairs (1,0) and (4,3)."""
    w = np.float32(29007.4609)
    p = np.array([[w / 2 - 3, 0], [w / 2 + 3, 0], [0, 0], [w / 4, 0], [w / 4 - 9, 0]], np.float32)
    c = np.zeros(5, np.uint32)
    assert orc.L.orc_collide_grid(5, orc._p(p), 10.0, None, orc._p(c)) == 2
    assert c.tolist() == [1, 1, 0, 1, 1]


ref = orc.ref_library()


@pytest.mark.skipif(ref is None, reason="reference tree not mounted: oracle/_ref not built")
def test_reference_quadtree_known_answer_runs():
    assert ref.ref_known_answer_test() == 0


def _ref_world():
    import ctypes
    w, h, r, nmax = (ctypes.c_float(), ctypes.c_float(), ctypes.c_float(), ctypes.c_uint())
    ref.ref_world(ctypes.byref(w), ctypes.byref(h), ctypes.byref(r), ctypes.byref(nmax))
    return w.value, h.value, r.value, nmax.value


@pytest.mark.skipif(ref is None, reason="reference tree not mounted: oracle/_ref not built")
def test_oracle_pairs_equal_reference_quadtree_single_leaf():
    """Up to 100 agents the reference's tree is one leaf (NODE_ENTITY_CAP,
    main.cpp:240) and its walk visits every pair exactly once: the pair
    counts must be equal."""
    w, h, r, nmax = _ref_world()
    assert r == 10.0 and nmax == 1000
    rng = np.random.default_rng(10)
    for trial in range(30):
        n = int(rng.integers(2, 101))
        c = rng.uniform(100, [w - 100, h - 100])
        p = np.clip(c + rng.normal(0, 8 + trial, (n, 2)), 1, [w - 1, h - 1]).astype(np.float32)
        flags = np.zeros(n, np.uint32)
        want = ref.ref_quadtree_pairs(n, orc._p(p), orc._p(flags))
        cnt = np.zeros(n, np.uint32)
        got = orc.L.orc_collide_grid(n, orc._p(p), 10.0, None, orc._p(cnt))
        assert want == got, trial
        np.testing.assert_array_equal(flags, (cnt > 0).astype(np.uint32))


@pytest.mark.skipif(ref is None, reason="reference tree not mounted: oracle/_ref not built")
def test_oracle_pair_set_matches_reference_quadtree():
    """Once the reference's tree splits, its counter can count a pair more
    than once: quad_tree_check_collisions_on_node's walk
    (random_move.comp:620-666) continues past the subtree it was called on
    into later siblings, which the parent climb (:694-718) also visits.  The
    set of agents in at least one pair (recoloured by quad_tree_collision)
    is still exact, and the
    reference count can only be larger."""
    w, h, r, nmax = _ref_world()
    rng = np.random.default_rng(11)
    extra = 0
    for trial in range(12):
        n = 1000
        centres = rng.uniform(100, [w - 100, h - 100], (5, 2))
        p = (centres[rng.integers(0, 5, n)] + rng.normal(0, 25 + 20 * trial, (n, 2)))
        p = np.clip(p, 1, [w - 1, h - 1]).astype(np.float32)
        flags = np.zeros(n, np.uint32)
        want = ref.ref_quadtree_pairs(n, orc._p(p), orc._p(flags))
        cnt = np.zeros(n, np.uint32)
        got = orc.L.orc_collide_grid(n, orc._p(p), 10.0, None, orc._p(cnt))
        np.testing.assert_array_equal(flags, (cnt > 0).astype(np.uint32))
        assert want >= got
        extra += want - got
    assert extra > 0  # the duplicate counting is real for split trees
