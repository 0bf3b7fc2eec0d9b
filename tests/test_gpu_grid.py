"""Candidate-pair enumeration of the cost kernels (hashed cell grid, wave-local queues).

The kernels report how many tree pairs passed the (2R)^2 cutoff; that count must equal a
brute-force count over all i < j with the same float32 differences, and the cost must
still match the FP64 oracle, also for layouts that stress the grid: every tree on one
point (one long bucket chain), trees far outside the clamped cell range, negative
coordinates, rows spaced at the cutoff distance.
"""
import numpy as np
import pytest
import torch

from gpu_util import GEOMS, assert_cost_close, dense_poses, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def _brute_pairs(x, cutoff_sq):
    """Pairs i < j with (x_j - x_i)^2 + (y_j - y_i)^2 <= cutoff^2 in float32 differences;
    also the number of pairs whose distance is within 1e-5 relative of the cutoff."""
    xs = x[:, 0].astype(np.float32)
    ys = x[:, 1].astype(np.float32)
    dx = (xs[None, :] - xs[:, None]).astype(np.float64)
    dy = (ys[None, :] - ys[:, None]).astype(np.float64)
    d2 = dx * dx + dy * dy
    iu = np.triu_indices(x.shape[0], 1)
    d2 = d2[iu]
    near = np.abs(d2 - cutoff_sq) <= 1e-5 * cutoff_sq
    return int((d2 <= cutoff_sq).sum()), int(near.sum())


def _cutoff_sq(geom):
    r = np.float32(geom.radius)
    return float(np.float32((2.0 * np.float64(r)) ** 2))


def _exact(geom, x, use_sep=False, block=0):
    from packing_amd import pack_cuda
    from packing_amd.context import set_geometry
    set_geometry(geom)
    t = torch.tensor(x, dtype=torch.float32, device="cuda")
    c = torch.empty(t.shape[0], dtype=torch.float32, device="cuda")
    g = torch.empty_like(t)
    st = torch.zeros(2, dtype=torch.int64, device="cuda")
    pack_cuda.overlap_multi_ensemble(t, t, use_sep, out_cost=c, out_grads=g, block=block, stats=st)
    torch.cuda.synchronize()
    return c.cpu().numpy(), g.cpu().numpy(), st.cpu().numpy()


def _layouts(n):
    rng = np.random.default_rng(11 + n)
    out = {}
    k = min(n, 120)    # all pairs overlap completely: keep the oracle's share small
    x = np.zeros((k, 3))
    x[:, 2] = rng.uniform(0, 2 * np.pi, k)
    out["one_point"] = x
    x = np.zeros((n, 3))
    x[:, 0] = rng.uniform(-8e4, 8e4, n)
    x[:, 1] = rng.uniform(-8e4, 8e4, n)
    x[: n // 2, 0] = 5e4 + rng.uniform(0, 2, n // 2)     # a cluster beyond the clamp
    x[: n // 2, 1] = -6e4 + rng.uniform(0, 2, n // 2)
    x[:, 2] = rng.uniform(0, 2 * np.pi, n)
    out["far"] = x
    x = dense_poses(5 + n, 1, n)[0].astype(np.float64)
    x[:, :2] -= 3.3
    out["negative"] = x
    return out


@pytest.mark.parametrize("gname", ["tree", "random"])
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 200, 700])
@pytest.mark.parametrize("block", [64, 256])
def test_pair_count_dense(gname, n, block):
    if gname == "random" and n > 400:
        n = 400    # 10 pieces: rotated vertices of 700 trees exceed the LDS of a workgroup
    geom = GEOMS[gname]()
    x = dense_poses(300 + n, 6, n)
    _, _, st = _exact(geom, x, block=block)
    total, near = 0, 0
    for e in range(x.shape[0]):
        t, nr = _brute_pairs(x[e], _cutoff_sq(geom))
        total += t
        near += nr
    assert abs(int(st[0]) - total) <= near, (st[0], total, near)


@pytest.mark.parametrize("gname", ["tree", "random"])
@pytest.mark.parametrize("n", [40, 300])
def test_grid_stress_layouts(oracle_mod, gname, n):
    geom = GEOMS[gname]()
    O = oracle_mod
    og = O.OracleGeometry(geom)
    for name, x in _layouts(n).items():
        x32 = x.astype(np.float32)[None]
        c, g, st = _exact(geom, x32)
        c64, g64 = O.overlap(og, x32.astype(np.float64), False)
        assert_cost_close(c, c64, 1e-5, atol=1e-6, what=f"{name} cost")
        total, near = _brute_pairs(x32[0], _cutoff_sq(geom))
        assert abs(int(st[0]) - total) <= near, (name, st[0], total)


def test_row_at_cutoff(oracle_mod):
    """Trees in a row spaced at exactly the cutoff distance (and just inside / outside):
    adjacent-cell pairs at the cell boundary."""
    geom = GEOMS["tree"]()
    O = oracle_mod
    og = O.OracleGeometry(geom)
    cut = np.sqrt(_cutoff_sq(geom))
    for f in (0.999, 1.0, 1.001):
        n = 50
        x = np.zeros((1, n, 3), dtype=np.float32)
        x[0, :, 0] = (np.arange(n) * cut * f).astype(np.float32)
        x[0, :, 1] = (np.arange(n) % 3 * 0.01).astype(np.float32)
        c, g, st = _exact(geom, x)
        c64, _ = O.overlap(og, x.astype(np.float64), False)
        assert_cost_close(c, c64, 1e-5, atol=1e-7, what=f"row {f}")
        total, near = _brute_pairs(x[0], _cutoff_sq(geom))
        assert abs(int(st[0]) - total) <= near, (f, st[0], total)


def test_block_invariance_bitwise():
    """Grid enumeration splits work by wave; the result must not depend on the block."""
    geom = GEOMS["random"]()
    x = dense_poses(77, 8, 300)
    ref = None
    for block in (64, 128, 256):
        c, g, st = _exact(geom, x, block=block)
        if ref is None:
            ref = (c, g, st)
        else:
            assert np.array_equal(c, ref[0]) and np.array_equal(g, ref[1])
            assert np.array_equal(st, ref[2])
