"""Exact overlap / separation / boundary kernels (pk_overlap, pk_boundary) against the FP64
oracle and the reference's recorded known answers.

Tolerances: cost within 1e-5 relative of the FP64 re-evaluation; the
gradient within 1e-5 * (|g| + max|g| of the instance) -- the second term covers elements
that are sums of cancelling contributions.  FP64 mode is held to 1e-10.
"""
import json
import os

import numpy as np
import pytest
import torch

from gpu_util import (GEOMS, assert_cost_close, assert_grad_close, dense_poses, poses,
                      require_gpu)

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def _run(geom, xyt_np, use_sep, grads=True, dtype=torch.float32, block=0, crystal=None,
         npd=2, only_self=False):
    from packing_amd import pack_cuda
    from packing_amd.context import set_geometry
    set_geometry(geom)
    x = torch.tensor(xyt_np, dtype=dtype, device="cuda")
    E, N = x.shape[:2]
    c = torch.empty(E, dtype=dtype, device="cuda")
    g = torch.empty(E, N, 3, dtype=dtype, device="cuda") if grads else None
    cr = None if crystal is None else torch.tensor(crystal, dtype=dtype, device="cuda")
    pack_cuda.overlap_multi_ensemble(x, x, use_sep, out_cost=c, out_grads=g, crystal_axes=cr,
                                     N_periodic=npd, only_self_interactions=only_self, block=block)
    torch.cuda.synchronize()
    return c.cpu().numpy(), (g.cpu().numpy() if grads else None)


def test_known_answers_gpu(oracle_mod):
    """The reference's recorded FP64 known answers (run_tests.ipynb:358-433), both precisions."""
    from test_oracle_known_answers import build
    from packing_amd.geometry import reference_tree
    geom = reference_tree()
    O = oracle_mod
    for key, use_sep in (("CollisionCostOverlappingArea", False), ("CollisionCostSeparation", True)):
        ref = KA["costs"][key]
        for i, sol in enumerate(KA["solutions"]):
            P, h, ax, npd, _, _ = build(sol, O)
            for dtype, rtol in ((torch.float32, 1e-5), (torch.float64, 1e-11)):
                xin = P.astype(np.float32) if dtype == torch.float32 else P
                if dtype == torch.float32:
                    # compare against the oracle on the float32-rounded inputs, then the record
                    o, _ = O.overlap(O.OracleGeometry(geom), xin.astype(np.float64), use_sep,
                                     crystal=None if ax is None else ax.astype(np.float32).astype(np.float64),
                                     n_periodic=npd, grads=False)
                    target = o[0]
                else:
                    target = ref["f64"][i] / ref["scaling"]
                c, _ = _run(geom, xin, use_sep, grads=False, dtype=dtype, crystal=ax, npd=npd)
                assert_cost_close(c, [target], rtol, what=f"{key}[{i}] {dtype}")
                # the float32 kernel also matches the record to the reference's own bar
                assert_cost_close(ref["scaling"] * c, [ref["f64"][i]], 1e-5 if dtype == torch.float32 else 1e-11)


@pytest.mark.parametrize("gname", ["tree", "random"])
@pytest.mark.parametrize("use_sep", [False, True])
@pytest.mark.parametrize("n", [2, 7, 50, 100])
def test_overlap_vs_oracle(oracle_mod, gname, use_sep, n):
    geom = GEOMS[gname]()
    O = oracle_mod
    og = O.OracleGeometry(geom)
    x = dense_poses(100 + n, 16, n)
    c, g = _run(geom, x, use_sep)
    c64, g64 = O.overlap(og, x.astype(np.float64), use_sep)
    assert (c64 > 0).sum() >= (8 if n > 2 else 0)
    assert_cost_close(c, c64, 1e-5, atol=1e-9, what="cost")
    assert_grad_close(g, g64, 1e-5)


@pytest.mark.parametrize("gname", ["tree", "random"])
@pytest.mark.parametrize("use_sep", [False, True])
def test_overlap_fp64_vs_oracle(oracle_mod, gname, use_sep):
    geom = GEOMS[gname]()
    O = oracle_mod
    x = dense_poses(7, 8, 40, dtype=np.float64)
    c, g = _run(geom, x, use_sep, dtype=torch.float64)
    c64, g64 = O.overlap(O.OracleGeometry(geom), x, use_sep)
    assert_cost_close(c, c64, 1e-10, atol=1e-14)
    assert_grad_close(g, g64, 1e-9)


@pytest.mark.parametrize("use_sep", [False, True])
def test_forward_only_equals_with_grad(use_sep):
    """core/pack_test.py:271 -- cost without gradients is bit-identical."""
    geom = GEOMS["tree"]()
    x = dense_poses(3, 32, 60)
    c1, _ = _run(geom, x, use_sep, grads=True)
    c2, _ = _run(geom, x, use_sep, grads=False)
    assert np.array_equal(c1, c2)


@pytest.mark.parametrize("use_sep", [False, True])
def test_batch_and_block_invariance(use_sep):
    """core/pack_test.py:372-447 -- an instance gives the same bits alone, in a batch and
    with any workgroup size."""
    geom = GEOMS["random"]()
    x = dense_poses(4, 24, 80)
    c, g = _run(geom, x, use_sep)
    for blk in (64, 128, 256):
        cb, gb = _run(geom, x, use_sep, block=blk)
        assert np.array_equal(c, cb) and np.array_equal(g, gb)
    for k in (0, 5, 23):
        c1, g1 = _run(geom, x[k:k + 1], use_sep)
        assert np.array_equal(c1[0], c[k]) and np.array_equal(g1[0], g[k])


def test_ragged_matches_uniform():
    from packing_amd import pack_cuda
    from packing_amd.context import set_geometry
    geom = GEOMS["tree"]()
    set_geometry(geom)
    sizes = [1, 2, 17, 60, 33, 100, 5]
    xs = [dense_poses(50 + k, 1, n)[0] for k, n in enumerate(sizes)]
    flat = torch.tensor(np.concatenate(xs), device="cuda")
    off = torch.tensor(np.concatenate([[0], np.cumsum(sizes)]), dtype=torch.int32, device="cuda")
    c = torch.empty(len(sizes), device="cuda")
    g = torch.empty_like(flat)
    pack_cuda.overlap_ragged(flat, off, max(sizes), False, c, g)
    torch.cuda.synchronize()
    for k, n in enumerate(sizes):
        cu, gu = _run(geom, xs[k][None], False)
        assert np.array_equal(cu[0], c[k].item()), k
        o = int(off[k])
        assert np.array_equal(gu[0], g[o:o + n].cpu().numpy()), k


def test_periodic_vs_oracle(oracle_mod):
    """Lattice images, including a tree's own images (cost only) and only_self."""
    geom = GEOMS["tree"]()
    O = oracle_mod
    og = O.OracleGeometry(geom)
    rng = np.random.default_rng(9)
    x = np.zeros((6, 3, 3))
    x[..., :2] = rng.uniform(-0.3, 0.3, (6, 3, 2))
    x[..., 2] = rng.uniform(0, 2 * np.pi, (6, 3))
    ax = np.stack([rng.uniform(0.5, 0.9, 6), np.zeros(6), rng.uniform(-0.3, 0.3, 6),
                   rng.uniform(0.6, 1.0, 6)], -1)
    for use_sep in (False, True):
        for only_self in (False, True):
            c, g = _run(geom, x, use_sep, dtype=torch.float64, crystal=ax, npd=2, only_self=only_self)
            c64, g64 = O.overlap(og, x, use_sep, crystal=ax, only_self=only_self, n_periodic=2)
            assert_cost_close(c, c64, 1e-10, atol=1e-14, what=f"sep={use_sep} self={only_self}")
            assert_grad_close(g, g64, 1e-9)


def test_boundary_vs_oracle(oracle_mod):
    from packing_amd import pack_cuda
    from packing_amd.context import set_geometry
    O = oracle_mod
    for gname in ("tree", "random"):
        geom = GEOMS[gname]()
        set_geometry(geom)
        og = O.OracleGeometry(geom)
        x = poses(11, 16, 40)
        h = np.tile(np.array([[6.0, 3.5, 3.5]]), (16, 1))
        h[:, 0] = np.linspace(2.0, 8.0, 16)
        for dtype, tol in ((torch.float32, 1e-5), (torch.float64, 1e-11)):
            xt = torch.tensor(x, dtype=dtype, device="cuda")
            ht = torch.tensor(h, dtype=dtype, device="cuda")
            c = torch.empty(16, dtype=dtype, device="cuda")
            g = torch.empty_like(xt)
            gh = torch.empty_like(ht)
            pack_cuda.boundary_distance_multi_ensemble(xt, ht, out_cost=c, out_grads=g, out_grad_h=gh)
            torch.cuda.synchronize()
            c64, g64, gh64 = O.boundary(og, x.astype(np.float64), h)
            assert_cost_close(c.cpu().numpy(), c64, tol, atol=1e-9)
            assert_grad_close(g.cpu().numpy(), g64, tol * 10)
            assert_grad_close(gh.cpu().numpy(), gh64, tol * 10)


def test_empty_and_tiny():
    from packing_amd import pack_cuda
    x = torch.empty(0, 5, 3, device="cuda")
    c = torch.empty(0, device="cuda")
    pack_cuda.overlap_multi_ensemble(x, x, False, out_cost=c)
    geom = GEOMS["tree"]()
    c1, g1 = _run(geom, np.zeros((3, 1, 3), np.float32), False)
    assert np.all(c1 == 0) and np.all(g1 == 0)
    # two identical poses overlap by exactly the polygon area
    c2, _ = _run(geom, np.zeros((1, 2, 3), np.float32), False)
    assert abs(c2[0] - geom.area) < 1e-6


def test_bad_arguments():
    from packing_amd import pack_cuda
    x = torch.zeros(2, 3, 3, device="cuda")
    with pytest.raises(ValueError):
        pack_cuda.overlap_multi_ensemble(x, x, False, out_cost=None)
    with pytest.raises(ValueError):
        pack_cuda.overlap_multi_ensemble(x, x, 1, out_cost=torch.empty(2, device="cuda"))
    with pytest.raises(ValueError):
        pack_cuda.overlap_multi_ensemble(x, x, False, out_cost=torch.empty(3, device="cuda"))
