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
