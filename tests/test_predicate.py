"""The neighbour predicate's exact rewrite, checked with numpy's correctly
rounded binary32 arithmetic (independent of the library and the oracle).

quad_tree_in_range (random_move.comp:551-562) is
    |dx| > R -> no; |dy| > R -> no; sqrt(dx*dx + dy*dy) < R.
The kernels evaluate d2 < L with L the smallest binary32 x with
sqrtf(x) >= R, and drop the box tests, relying on fl(R*R) >= L.
"""
import numpy as np

f32 = np.float32


def limit(r):
    r = f32(r)
    lo, hi = 0, 0x7F800000
    while lo < hi:
        mid = (lo + hi) // 2
        x = np.array([mid], np.uint32).view(np.float32)[0]
        if np.sqrt(x) >= r:
            hi = mid
        else:
            lo = mid + 1
    return np.array([lo], np.uint32).view(np.float32)[0]


def test_box_tests_are_implied_for_many_radii():
    rng = np.random.default_rng(0)
    radii = [1.0, 10.0, 0.75, 2.5, 1.5, 3.0, 1e-3, 123.456, 2.0**-20, 7.0e4]
    radii += list(rng.uniform(0.01, 100, 300).astype(np.float32))
    radii += list((2.0 ** rng.uniform(-30, 60, 300)).astype(np.float32))
    for r in radii:
        r = f32(r)
        L = limit(r)
        assert f32(r * r) >= L, r
        # the limit really is the boundary of sqrt(x) < r
        below = np.nextafter(L, f32(0))
        assert np.sqrt(below) < r and np.sqrt(L) >= r


def test_rewrite_matches_reference_predicate_on_boundary_pairs():
    rng = np.random.default_rng(1)
    for r in (f32(1.0), f32(0.75), f32(10.0), f32(2.5)):
        L = limit(r)
        n = 200000
        p = rng.uniform(-3000, 3000, (n, 2)).astype(np.float32)
        ang = rng.uniform(0, 2 * np.pi, n)
        rad = (r * (1 + rng.normal(0, 1e-6, n))).astype(np.float64)
        q = (p + np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1)).astype(np.float32)
        # axis-aligned pairs at exactly r and one ulp around it
        q[:1000, 0] = p[:1000, 0] + r
        q[:1000, 1] = p[:1000, 1]
        q[1000:2000, 0] = np.nextafter(p[1000:2000, 0] + r, f32(-np.inf))
        q[1000:2000, 1] = p[1000:2000, 1]
        dx = (q[:, 0] - p[:, 0]).astype(np.float32)
        dy = (q[:, 1] - p[:, 1]).astype(np.float32)
        d2 = (dx * dx + dy * dy).astype(np.float32)
        ref = ~((np.abs(p[:, 0] - q[:, 0]) > r) | (np.abs(p[:, 1] - q[:, 1]) > r)) & (np.sqrt(d2) < r)
        new = d2 < L
        assert (ref == new).all()
        assert ref.any() and (~ref).any()
