"""Hungarian and diversity kernels against the oracle: assignments bit-exact (int32 and
float32 matrices), totals exact; diversity within float32 rounding."""
import numpy as np
import pytest
import torch

from gpu_util import require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


@pytest.mark.parametrize("n", [1, 2, 5, 63, 64, 65, 128, 129, 200])
def test_int32_bit_exact(oracle_mod, n):
    from packing_amd.lap_batch import solve_lap_batch
    rng = np.random.default_rng(100 + n)
    B = 24
    C = rng.integers(0, 10 ** 6, size=(B, n, n)).astype(np.int32)
    rm, tot = solve_lap_batch(torch.tensor(C, device="cuda"), exact_int_total=True)
    rm_o, tot_o = oracle_mod.hungarian_i32(C)
    assert np.array_equal(rm.cpu().numpy(), rm_o)
    assert np.array_equal(tot.cpu().numpy(), tot_o)
    _, tot32 = solve_lap_batch(torch.tensor(C, device="cuda"))
    _, f32 = oracle_mod.hungarian_f32(C.astype(np.float32))
    assert np.array_equal(tot32.cpu().numpy(), f32)


@pytest.mark.parametrize("n", [3, 40, 200])
def test_float32_bit_exact(oracle_mod, n):
    """Real-valued costs (as the diversity matrices are): same float32 arithmetic and
    tie-breaks as the reference kernel, so identical row_match and totals."""
    from packing_amd.lap_batch import solve_lap_batch
    rng = np.random.default_rng(n)
    C = np.sqrt(rng.uniform(0, 4, size=(16, n, n))).astype(np.float32)
    rm, tot = solve_lap_batch(torch.tensor(C, device="cuda"))
    rm_o, tot_o = oracle_mod.hungarian_f32(C)
    assert np.array_equal(rm.cpu().numpy(), rm_o)
    assert np.array_equal(tot.cpu().numpy(), tot_o)


def test_ties_and_config_size(oracle_mod):
    """Many ties (entries in [0, 3)) and the benchmark shape's first problems."""
    from packing_amd.lap_batch import solve_lap_batch
    rng = np.random.default_rng(5)
    C = rng.integers(0, 3, size=(32, 60, 60)).astype(np.int32)
    rm, _ = solve_lap_batch(torch.tensor(C, device="cuda"))
    assert np.array_equal(rm.cpu().numpy(), oracle_mod.hungarian_i32(C)[0])
    C = rng.integers(0, 10 ** 6, size=(512, 200, 200)).astype(np.int32)
    rm, tot = solve_lap_batch(torch.tensor(C, device="cuda"), exact_int_total=True)
    rm = rm.cpu().numpy()
    rm_o, tot_o = oracle_mod.hungarian_i32(C[:16])
    assert np.array_equal(rm[:16], rm_o)
    from scipy.optimize import linear_sum_assignment
    for b in range(16, 512, 37):
        r, c = linear_sum_assignment(C[b])
        assert int(tot[b]) == int(C[b][r, c].sum())


def test_min_cost_modes():
    from packing_amd.lap_batch import LAPConfig, solve_lap_batch
    C = torch.rand(4, 10, 10, device="cuda")
    rm, t = solve_lap_batch(C, LAPConfig("min_cost_row"))
    assert (rm == -1).all() and torch.allclose(t, C.amin(2).sum(1))
    rm, t = solve_lap_batch(C, LAPConfig("min_cost_col"))
    assert torch.allclose(t, C.amin(1).sum(1))


@pytest.mark.parametrize("mirror", [False, True])
def test_diversity_vs_oracle(oracle_mod, mirror):
    from packing_amd.lap_batch import compute_diversity_shortcut_kernel
    rng = np.random.default_rng(7)
    pop = rng.uniform(-3, 3, (5, 30, 3)).astype(np.float32)
    ref = rng.uniform(-3, 3, (4, 30, 3)).astype(np.float32)
    a = 0.7
    got = compute_diversity_shortcut_kernel(torch.tensor(pop, device="cuda"), torch.tensor(ref, device="cuda"),
                                            np.cos(a), np.sin(a), mirror).cpu().numpy()
    want = oracle_mod.diversity(pop, ref, np.float32(np.cos(a)), np.float32(np.sin(a)), mirror)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-5)


@pytest.mark.parametrize("algorithm", ["hungarian", "min_cost_row", "min_cost_col"])
@pytest.mark.parametrize("n,n_pop,n_ref", [(15, 5, 4), (50, 3, 3), (200, 1, 2)])
def test_diversity_matrix_vs_oracle(oracle_mod, algorithm, n, n_pop, n_ref):
    """compute_genetic_diversity_matrix in one launch (cost matrices built in LDS and
    solved there) equals the oracle bit for bit, per transform and after the minimum."""
    from packing_amd.diversity import compute_genetic_diversity_matrix
    from packing_amd.lap_batch import LAPConfig
    rng = np.random.default_rng(n)
    L = np.sqrt(n)
    pop = np.concatenate([rng.uniform(0, L, (n_pop, n, 2)), rng.uniform(0, 2 * np.pi, (n_pop, n, 1))], 2)
    ref = np.concatenate([rng.uniform(0, L, (n_ref, n, 2)), rng.uniform(0, 2 * np.pi, (n_ref, n, 1))], 2)
    pop, ref = pop.astype(np.float32), ref.astype(np.float32)
    best, per = compute_genetic_diversity_matrix(torch.tensor(pop, device="cuda"), torch.tensor(ref, device="cuda"),
                                                 LAPConfig(algorithm=algorithm), allow_shortcut=False,
                                                 per_transform=True)
    per_o, best_o = oracle_mod.diversity_matrix(pop, ref, algorithm)
    assert np.array_equal(per.cpu().numpy(), per_o)
    assert np.array_equal(best.cpu().numpy(), best_o)


def test_diversity_single_reference_and_shortcut():
    """compute_genetic_diversity (one reference) is the matrix's column; the shortcut path
    (min_cost_row with allow_shortcut) agrees with the full min_cost_row path to float32
    rounding of its float32 formula."""
    from packing_amd.diversity import compute_genetic_diversity, compute_genetic_diversity_matrix
    from packing_amd.lap_batch import LAPConfig
    rng = np.random.default_rng(3)
    pop = torch.tensor(rng.uniform(0, 5, (6, 25, 3)), device="cuda", dtype=torch.float32)
    ref = torch.tensor(rng.uniform(0, 5, (3, 25, 3)), device="cuda", dtype=torch.float32)
    m = compute_genetic_diversity_matrix(pop, ref)
    for r in range(3):
        assert torch.equal(compute_genetic_diversity(pop, ref[r]), m[:, r])
    cfg = LAPConfig(algorithm="min_cost_row")
    a = compute_genetic_diversity_matrix(pop, ref, cfg)
    b = compute_genetic_diversity_matrix(pop, ref, cfg, allow_shortcut=False)
    assert torch.allclose(a, b, rtol=1e-4, atol=1e-4)
