"""Oracle genetic-diversity matrix (restating core/kaggle_support.py:2014-2133), pinned by
what the definition fixes: a packing and any symmetric, permuted copy of it are at
distance 0; the Hungarian total is scipy's optimal assignment of the same float32
matrices; the min-cost-row total agrees with the reference's shortcut formula
(core/lap_batch.py:216-331, the oracle's float32 restatement) and lower-bounds it."""
import numpy as np
from scipy.optimize import linear_sum_assignment


def _pack(rng, n):
    x = np.empty((n, 3))
    x[:, :2] = rng.uniform(-3, 3, (n, 2))
    x[:, 2] = rng.uniform(-np.pi, np.pi, n)
    return x


def _apply(x, rot, mirror):
    """The transform the matrix applies to reference individuals."""
    y = x.copy()
    if mirror:
        y[:, 1] = -y[:, 1]
        y[:, 2] = np.pi - y[:, 2]
    c, s = np.cos(rot), np.sin(rot)
    y[:, 0], y[:, 1] = y[:, 0] * c - y[:, 1] * s, y[:, 0] * s + y[:, 1] * c
    y[:, 2] = y[:, 2] + rot
    return y


def test_symmetric_permuted_copies_are_identical(oracle_mod):
    rng = np.random.default_rng(0)
    n = 12
    pop = _pack(rng, n)
    refs = []
    for k, (rot, mir) in enumerate([(0.0, False), (np.pi / 2, False), (np.pi, True), (3 * np.pi / 2, True)]):
        # ref such that transforming it gives pop back (up to a permutation of the trees)
        inv = _apply(pop, -rot, False)
        if mir:
            inv[:, 1] = -inv[:, 1]
            inv[:, 2] = np.pi - inv[:, 2]
        refs.append(inv[rng.permutation(n)])
    per, best = oracle_mod.diversity_matrix(pop[None], np.stack(refs))
    assert np.all(best < 1e-4), best
    other = _pack(rng, n)
    _, b2 = oracle_mod.diversity_matrix(pop[None], other[None])
    assert b2[0, 0] > 0.5


def test_hungarian_is_optimal_and_bounds(oracle_mod):
    rng = np.random.default_rng(1)
    pop = np.stack([_pack(rng, 15) for _ in range(3)])
    ref = np.stack([_pack(rng, 15) for _ in range(2)])
    C = oracle_mod.diversity_cost_matrices(pop, ref)
    per_h, _ = oracle_mod.diversity_matrix(pop, ref, "hungarian")
    per_r, _ = oracle_mod.diversity_matrix(pop, ref, "min_cost_row")
    per_c, _ = oracle_mod.diversity_matrix(pop, ref, "min_cost_col")
    for t in range(8):
        for p in range(3):
            for r in range(2):
                m = C[t, p, r].astype(np.float64)
                ri, ci = linear_sum_assignment(m)
                assert abs(per_h[t, p, r] - m[ri, ci].sum()) < 1e-4
                assert per_r[t, p, r] <= per_h[t, p, r] + 1e-4 and per_c[t, p, r] <= per_h[t, p, r] + 1e-4
                assert abs(per_r[t, p, r] - m.min(axis=1).sum()) < 1e-4


def test_min_cost_row_matches_shortcut_formula(oracle_mod):
    rng = np.random.default_rng(2)
    pop = np.stack([_pack(rng, 20) for _ in range(3)]).astype(np.float32)
    ref = np.stack([_pack(rng, 20) for _ in range(4)]).astype(np.float32)
    per, best = oracle_mod.diversity_matrix(pop, ref, "min_cost_row")
    sc = np.full((3, 4), np.inf, dtype=np.float32)
    for rot, mir in ((0.0, False), (np.pi / 2, False), (np.pi, False), (3 * np.pi / 2, False),
                     (0.0, True), (np.pi / 2, True), (np.pi, True), (3 * np.pi / 2, True)):
        sc = np.minimum(sc, oracle_mod.diversity(pop, ref, np.float32(np.cos(rot)), np.float32(np.sin(rot)), mir))
    assert np.allclose(best, sc, rtol=1e-4, atol=1e-4)
