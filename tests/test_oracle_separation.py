"""Oracle signed separation distance (restating core/pack_minkowski.py:33-266), pinned
by closed forms (squares and rectangles, a two-piece rectangle whose piece-pair zones
overlap), convex hulls against scipy, and, for the tree, penetration depths found by
casting rays from the point in 7200 directions (an independent computation: the
smallest exit distance over directions converges to the distance to the boundary)."""
import numpy as np
import pytest
from scipy.spatial import ConvexHull

from packing_amd.geometry import Geometry, reference_tree


def rect_geom(a, b):
    v = np.array([[-a / 2, -b / 2], [a / 2, -b / 2], [a / 2, b / 2], [-a / 2, b / 2]])
    return Geometry(pieces=[v], outline=v, radius=float(np.hypot(a, b) / 2), name="rect")


def two_square_geom():
    """A 2 x 1 rectangle made of two unit squares (pieces share the edge x = 0)."""
    l = np.array([[-1, -0.5], [0, -0.5], [0, 0.5], [-1, 0.5]], dtype=float)
    r = l + [1, 0]
    out = np.array([[-1, -0.5], [1, -0.5], [1, 0.5], [-1, 0.5]], dtype=float)
    return Geometry(pieces=[l, r], outline=out, radius=float(np.hypot(1, 0.5)), name="two")


def rect_sep(a, b, dx, dy):
    """Zone of two axis-aligned a x b rectangles: the 2a x 2b rectangle."""
    ex, ey = abs(dx) - a, abs(dy) - b
    if ex < 0 and ey < 0:
        return min(-ex, -ey)
    return -np.hypot(max(ex, 0), max(ey, 0))


@pytest.mark.parametrize("a,b", [(1.0, 1.0), (1.3, 0.4)])
def test_rectangles_closed_form(oracle_mod, a, b):
    og = oracle_mod.OracleGeometry(rect_geom(a, b))
    rng = np.random.default_rng(0)
    for _ in range(300):
        dx, dy = rng.uniform(-2.5 * a, 2.5 * a), rng.uniform(-2.5 * b, 2.5 * b)
        assert abs(oracle_mod.separation(og, dx, dy, 0.0) - rect_sep(a, b, dx, dy)) < 1e-12
        # a quarter turn swaps the second rectangle's sides: zone (a + b) x (a + b)
        s = rect_sep((a + b) / 2, (a + b) / 2, dx, dy)
        assert abs(oracle_mod.separation(og, dx, dy, np.pi / 2) - s) < 1e-12
        # a half turn maps a rectangle onto itself
        assert abs(oracle_mod.separation(og, dx, dy, np.pi) - rect_sep(a, b, dx, dy)) < 1e-12


def test_overlapping_piece_zones(oracle_mod):
    """The 2 x 1 rectangle as two squares: the four piece-pair zones overlap and their
    union is the 4 x 2 rectangle; edges inside the union are not boundary."""
    og = oracle_mod.OracleGeometry(two_square_geom())
    rng = np.random.default_rng(1)
    for _ in range(400):
        dx, dy = rng.uniform(-2.6, 2.6), rng.uniform(-1.4, 1.4)
        assert abs(oracle_mod.separation(og, dx, dy, 0.0) - rect_sep(2.0, 1.0, dx, dy)) < 1e-12
    # near the internal edges of the single zones, deep inside the union
    for dx in (-1.0, 0.0, 1.0, 0.5):
        assert abs(oracle_mod.separation(og, dx, 0.0, 0.0) - rect_sep(2.0, 1.0, dx, 0.0)) < 1e-12


def test_hulls_against_scipy(oracle_mod):
    g = reference_tree()
    og = oracle_mod.OracleGeometry(g)
    rng = np.random.default_rng(2)
    for th in rng.uniform(-np.pi, np.pi, 5):
        hulls = oracle_mod.minkowski_hulls(og, th)
        c, s = np.cos(th), np.sin(th)
        k = 0
        for pa in g.pieces:
            for pb in g.pieces:
                rb = pb @ np.array([[c, s], [-s, c]])
                pts = (pa[:, None, :] - rb[None, :, :]).reshape(-1, 2)
                ref = ConvexHull(pts)
                h = hulls[k]
                assert abs(ConvexHull(h).volume - ref.volume) < 1e-12
                assert len(h) == len(ref.vertices)
                k += 1


def _exit_dist(hulls, p, angles):
    """Distance from p (inside the union) along each direction to the first point outside
    every hull interior, vectorised over directions."""
    U = np.stack([np.cos(angles), np.sin(angles)], 1)
    ivs = []
    for h in hulls:
        lo = np.full(len(U), -np.inf)
        hi = np.full(len(U), np.inf)
        ok = np.ones(len(U), dtype=bool)
        for a in range(len(h)):
            w0, w1 = h[a], h[(a + 1) % len(h)]
            e = w1 - w0
            f0 = e[0] * (p[1] - w0[1]) - e[1] * (p[0] - w0[0])
            df = e[0] * U[:, 1] - e[1] * U[:, 0]
            with np.errstate(divide="ignore", invalid="ignore"):
                r = -f0 / df
            lo = np.where(df > 0, np.maximum(lo, r), lo)
            hi = np.where(df < 0, np.minimum(hi, r), hi)
            ok &= (df != 0) | (f0 > 0)
        ivs.append((np.where(ok, lo, np.inf), np.where(ok, hi, -np.inf)))
    s = np.zeros(len(U))
    for _ in range(len(hulls) + 1):           # grow the covered stretch from 0
        for lo, hi in ivs:
            cov = (lo < s) & (s < hi)
            s = np.where(cov, hi, s)
    return s


def _ray_depth(hulls, p):
    """Smallest exit distance over directions: 3600 directions, then each of the best
    ten refined on a 1e-7 rad grid around it (the exact nearest boundary point is hit by
    the direction towards it, so the minimum converges to the distance)."""
    ang = np.linspace(0, 2 * np.pi, 3600, endpoint=False)
    e = _exit_dist(hulls, p, ang)
    best = e.min()
    for k in np.argsort(e)[:10]:
        fine = ang[k] + np.linspace(-2e-3, 2e-3, 40001)
        best = min(best, _exit_dist(hulls, p, fine).min())
    return best


def test_tree_depth_against_ray_casting(oracle_mod):
    g = reference_tree()
    og = oracle_mod.OracleGeometry(g)
    rng = np.random.default_rng(3)
    n_in = n_out = 0
    while n_in < 6 or n_out < 6:
        dx, dy, th = rng.uniform(-0.7, 0.7), rng.uniform(-1.1, 1.1), rng.uniform(-np.pi, np.pi)
        d = oracle_mod.separation(og, dx, dy, th)
        hulls = oracle_mod.minkowski_hulls(og, th)
        if d > 0 and n_in < 6:
            r = _ray_depth(hulls, np.array([dx, dy]))
            assert d <= r + 1e-12 and r - d <= 1e-8 * d + 1e-12, (d, r)
            n_in += 1
        elif d < 0 and n_out < 6:
            # outside: the distance to the union is the smallest distance to a hull
            p = np.array([dx, dy])
            best = np.inf
            for h in hulls:
                for a in range(len(h)):
                    u, v = h[a], h[(a + 1) % len(h)]
                    t = np.clip(np.dot(p - u, v - u) / np.dot(v - u, v - u), 0, 1)
                    best = min(best, np.linalg.norm(p - (u + t * (v - u))))
            assert abs(-d - best) < 1e-12
            n_out += 1


def test_separation_sign_matches_overlap(oracle_mod):
    """Positive separation exactly where the exact overlap area is positive."""
    g = reference_tree()
    og = oracle_mod.OracleGeometry(g)
    rng = np.random.default_rng(4)
    for _ in range(200):
        dx, dy, th = rng.uniform(-1.2, 1.2), rng.uniform(-1.2, 1.2), rng.uniform(-np.pi, np.pi)
        d = oracle_mod.separation(og, dx, dy, th)
        a, _ = oracle_mod.overlap(og, np.array([[[0, 0, 0], [dx, dy, th]]]), False, grads=False)
        if abs(d) > 1e-9:
            assert (d > 0) == (a[0] > 1e-12), (dx, dy, th, d, a)
