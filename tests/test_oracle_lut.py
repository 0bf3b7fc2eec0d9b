"""Oracle table lookup (restating core/pack_cuda_lut.py:298-415), pinned by closed forms:
a table of an exactly trilinear function is reproduced between nodes with the matching
analytic gradient, and the half table's mirror rule holds."""
import numpy as np


def test_trilinear_closed_form(oracle_mod):
    X = np.linspace(-1, 1, 9, dtype=np.float32)
    Y = np.linspace(-1, 1, 7, dtype=np.float32)
    T = np.linspace(-np.pi, np.pi, 5, dtype=np.float32)

    def f(x, y, t):
        return 0.3 + 0.2 * x - 0.1 * y + 0.05 * t + 0.4 * x * y - 0.07 * x * y * t

    vals = f(X[:, None, None], Y[None, :, None], T[None, None, :]).astype(np.float32)
    OL = oracle_mod.OracleLut(X, Y, T, vals)
    rng = np.random.default_rng(1)
    for _ in range(50):
        x, y, t = rng.uniform(-0.99, 0.99), rng.uniform(-0.99, 0.99), rng.uniform(-3.1, 3.1)
        v, g = OL.lookup(x, y, t)
        assert abs(v - f(x, y, t)) < 1e-6
        e = 1e-6
        num = [(OL.lookup(x + e, y, t)[0] - OL.lookup(x - e, y, t)[0]) / (2 * e),
               (OL.lookup(x, y + e, t)[0] - OL.lookup(x, y - e, t)[0]) / (2 * e),
               (OL.lookup(x, y, t + e)[0] - OL.lookup(x, y, t - e)[0]) / (2 * e)]
        assert np.allclose(g, num, atol=1e-6)


def test_outside_range_is_zero(oracle_mod):
    X = np.linspace(-1, 1, 5, dtype=np.float32)
    vals = np.ones((5, 5, 5), np.float32)
    OL = oracle_mod.OracleLut(X, X, np.linspace(-np.pi, np.pi, 5), vals)
    for p in ((1.01, 0, 0), (0, -1.01, 0), (0, 0, 3.2)):
        v, g = OL.lookup(*p)
        assert v == 0 and not g.any()


def test_mirror_half_table(oracle_mod):
    """Half table: x < 0 is looked up as (-x, -theta) with the x and theta gradients
    negated (core/pack_cuda_lut.py:314-320,404-407)."""
    X = np.linspace(-0.25, 1, 6, dtype=np.float32)
    Y = np.linspace(-1, 1, 5, dtype=np.float32)
    T = np.linspace(-np.pi, np.pi, 7, dtype=np.float32)
    rng = np.random.default_rng(2)
    vals = rng.uniform(0, 1, (6, 5, 7)).astype(np.float32)
    OL = oracle_mod.OracleLut(X, Y, T, vals, mirror=True)
    for _ in range(20):
        x, y, t = rng.uniform(0.01, 0.99), rng.uniform(-0.9, 0.9), rng.uniform(-3.0, 3.0)
        v1, g1 = OL.lookup(x, y, t)
        v2, g2 = OL.lookup(-x, y, -t)
        assert v1 == v2 and g1[0] == -g2[0] and g1[1] == g2[1] and g1[2] == -g2[2]


# ---------------------------------------------------------------------------------
# or_lut_overlap (core/pack_cuda_lut.py:420-640): the pair loop, the relative pose in
# the first tree's frame and the chain rule to both trees, pinned independently of the
# restated formulas.
# ---------------------------------------------------------------------------------
def _trilinear_table(nx=17, ny=15, nt=13, R=0.6):
    """A table of a function that is trilinear over the whole grid, so that the
    interpolated value equals the function everywhere and the pair cost is smooth."""
    X = np.linspace(-2 * R, 2 * R, nx).astype(np.float32)
    Y = np.linspace(-2 * R, 2 * R, ny).astype(np.float32)
    T = np.linspace(-np.pi, np.pi, nt).astype(np.float32)

    def f(x, y, t):
        return 0.5 + 0.3 * x - 0.2 * y + 0.07 * t + 0.25 * x * y - 0.05 * y * t + 0.03 * x * y * t

    vals = f(X[:, None, None].astype(np.float64), Y[None, :, None], T[None, None, :]).astype(np.float32)
    return X, Y, T, vals, f


def _cluster(rng, n, R, spread=0.9):
    """n trees within the pair cutoff of each other's neighbours, relative angles away
    from the +-pi wrap and centre distances away from the cutoff."""
    while True:
        p = np.zeros((n, 3))
        p[:, 0] = rng.uniform(0, spread * R, n)
        p[:, 1] = rng.uniform(0, spread * R, n)
        p[:, 2] = rng.uniform(-1.2, 1.2, n) + 0.4
        d = np.sqrt((p[:, None, 0] - p[None, :, 0]) ** 2 + (p[:, None, 1] - p[None, :, 1]) ** 2)
        if np.all((np.abs(d - 2 * R) > 0.05)):
            return p


def test_lut_overlap_gradient_central_differences(oracle_mod):
    """The analytic gradient w.r.t. every tree's (x, y, theta) equals FP64 central
    differences of the oracle's own cost: a sign or frame error in the relative pose or
    in its chain rule to either tree fails here."""
    R = 0.6
    X, Y, T, vals, _ = _trilinear_table(R=R)
    OL = oracle_mod.OracleLut(X, Y, T, vals)
    rng = np.random.default_rng(3)
    for _ in range(4):
        p = _cluster(rng, 6, R)[None]
        c, g = OL.overlap(R, p)
        h = 1e-6
        num = np.zeros_like(p)
        for t in range(p.shape[1]):
            for k in range(3):
                a, b = p.copy(), p.copy()
                a[0, t, k] += h
                b[0, t, k] -= h
                num[0, t, k] = (OL.overlap(R, a, grads=False)[0][0] - OL.overlap(R, b, grads=False)[0][0]) / (2 * h)
        assert np.abs(num - g).max() <= 1e-6 * max(1.0, np.abs(g).max()), np.abs(num - g).max()
        assert np.abs(g).max() > 0.05


def test_lut_overlap_rigid_motion_invariance(oracle_mod):
    """Rotating and translating every tree together leaves each relative pose, hence the
    cost, unchanged, and rotates the translational gradient with the motion."""
    R = 0.6
    X, Y, T, vals, _ = _trilinear_table(R=R)
    OL = oracle_mod.OracleLut(X, Y, T, vals)
    rng = np.random.default_rng(4)
    for _ in range(5):
        p = _cluster(rng, 7, R)[None]
        c0, g0 = OL.overlap(R, p)
        phi = rng.uniform(-np.pi, np.pi)
        cphi, sphi = np.cos(phi), np.sin(phi)
        tx, ty = rng.uniform(-5, 5, 2)
        q = p.copy()
        q[..., 0] = cphi * p[..., 0] - sphi * p[..., 1] + tx
        q[..., 1] = sphi * p[..., 0] + cphi * p[..., 1] + ty
        q[..., 2] = p[..., 2] + phi
        c1, g1 = OL.overlap(R, q)
        assert abs(c1[0] - c0[0]) <= 1e-12 * max(1.0, abs(c0[0]))
        gx = cphi * g0[..., 0] - sphi * g0[..., 1]
        gy = sphi * g0[..., 0] + cphi * g0[..., 1]
        assert np.allclose(g1[..., 0], gx, atol=1e-11)
        assert np.allclose(g1[..., 1], gy, atol=1e-11)
        assert np.allclose(g1[..., 2], g0[..., 2], atol=1e-11)
        # translations and a common rotation carry no force or torque in total
        assert abs(g0[..., 0].sum()) < 1e-11 and abs(g0[..., 1].sum()) < 1e-11


def _exact_table(og, oracle_mod, X, Y, T):
    """Node values = exact FP64 overlap area of the polygon at the origin (theta 0) and
    at (X, Y, T), rounded to float32: the content the builder writes."""
    I = np.stack(np.meshgrid(np.arange(len(X)), np.arange(len(Y)), np.arange(len(T)), indexing="ij"),
                 -1).reshape(-1, 3)
    xyt = np.zeros((len(I), 2, 3))
    xyt[:, 1, 0] = X[I[:, 0]]
    xyt[:, 1, 1] = Y[I[:, 1]]
    xyt[:, 1, 2] = T[I[:, 2]]
    v, _ = oracle_mod.overlap(og, xyt, False, grads=False)
    return v.astype(np.float32).reshape(len(X), len(Y), len(T))


def test_lut_overlap_on_nodes_equals_exact_area(oracle_mod):
    """Two trees at arbitrary world poses whose relative pose (j in i's frame) is a grid
    node: the table cost equals the exact clipped area of the two trees to the table's
    float32 node rounding.  Both the full table and the reference's mirrored half table
    (for the mirror-symmetric tree, j on the x < 0 side) are checked; a transposed frame
    or a wrong mirror rule lands on another node."""
    from packing_amd.geometry import reference_tree
    geom = reference_tree()
    og = oracle_mod.OracleGeometry(geom)
    R = geom.radius
    nx, ny, nt = 21, 19, 16
    Xf = np.linspace(-2 * R, 2 * R, nx).astype(np.float32)
    Y = np.linspace(-2 * R, 2 * R, ny).astype(np.float32)
    T = np.linspace(-np.pi, np.pi, nt).astype(np.float32)
    Xh = Xf[nx // 2 - 1:]
    full = oracle_mod.OracleLut(Xf, Y, T, _exact_table(og, oracle_mod, Xf, Y, T))
    half = oracle_mod.OracleLut(Xh, Y, T, _exact_table(og, oracle_mod, Xh, Y, T), mirror=True)
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(300):
        a, b, k = rng.integers(1, nx - 1), rng.integers(1, ny - 1), rng.integers(1, nt - 1)
        rel = np.array([Xf[a], Y[b], T[k]], dtype=np.float64)
        if rel[0] ** 2 + rel[1] ** 2 > 4 * R * R:
            continue
        xi, yi, ti = rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(-np.pi, np.pi)
        c, s = np.cos(ti), np.sin(ti)
        p = np.array([[[xi, yi, ti], [xi + c * rel[0] - s * rel[1], yi + s * rel[0] + c * rel[1], ti + rel[2]]]])
        exact, _ = oracle_mod.overlap(og, p, False, grads=False)
        for L in (full, half):
            v, _ = L.overlap(R, p, grads=False)
            assert abs(v[0] - exact[0]) <= 2e-7 * max(1.0, exact[0]) + 1e-9, (rel, v, exact)
        checked += exact[0] > 0
    assert checked > 40
