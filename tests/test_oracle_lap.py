"""Oracle Hungarian (restating core/lap_batch.py:62-196) pinned by what mathematics fixes:
the optimal total equals scipy.optimize.linear_sum_assignment's, the assignment is a
permutation, brute force agrees on tiny problems, and the float32 restatement equals the
integer one on integer matrices below 2^20 (exact float arithmetic)."""
import itertools

import numpy as np
import pytest
from scipy.optimize import linear_sum_assignment


@pytest.mark.parametrize("n", [1, 2, 3, 7, 31, 64, 100])
def test_int_optimal_vs_scipy(oracle_mod, n):
    rng = np.random.default_rng(n)
    C = rng.integers(0, 10 ** 6, size=(6, n, n)).astype(np.int32)
    rm, tot = oracle_mod.hungarian_i32(C)
    for b in range(6):
        assert sorted(rm[b]) == list(range(n))
        r, c = linear_sum_assignment(C[b])
        assert tot[b] == C[b][r, c].sum() == C[b][np.arange(n), rm[b]].sum()


def test_brute_force_tiny(oracle_mod):
    rng = np.random.default_rng(0)
    C = rng.integers(0, 20, size=(40, 5, 5)).astype(np.int32)   # many ties
    rm, tot = oracle_mod.hungarian_i32(C)
    for b in range(40):
        best = min(sum(C[b][i, p[i]] for i in range(5)) for p in itertools.permutations(range(5)))
        assert tot[b] == best


def test_float_equals_int_on_integer_matrices(oracle_mod):
    rng = np.random.default_rng(3)
    for n, hi in ((50, 10 ** 6), (120, 1000), (200, 10 ** 6)):
        C = rng.integers(0, hi, size=(3, n, n)).astype(np.int32)
        rmi, toti = oracle_mod.hungarian_i32(C)
        rmf, totf = oracle_mod.hungarian_f32(C.astype(np.float32))
        assert np.array_equal(rmi, rmf)


def test_diversity_closed_form(oracle_mod):
    """Identity transform and identical sets give 0; a single tree translated by t gives
    |t|; a quarter turn of a single tree at the origin gives the angle pi/2."""
    rng = np.random.default_rng(1)
    pop = rng.uniform(-2, 2, (3, 9, 3)).astype(np.float32)
    out = oracle_mod.diversity(pop, pop, 1.0, 0.0, False)
    assert np.allclose(np.diag(out), 0, atol=1e-5)   # theta passes through float wrap
    one = rng.uniform(-2, 2, (3, 1, 3)).astype(np.float32)
    shifted = one.copy()
    shifted[..., 0] += 3.0
    shifted[..., 1] -= 4.0
    assert np.allclose(np.diag(oracle_mod.diversity(one, shifted, 1.0, 0.0, False)), 5.0, rtol=1e-5)
    z = np.zeros((1, 1, 3), np.float32)
    assert np.allclose(oracle_mod.diversity(z, z, 0.0, 1.0, False), np.pi / 2, rtol=1e-6)
