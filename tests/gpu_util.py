"""Helpers shared by the GPU tests (inputs, tolerances, oracle views)."""
import numpy as np
import torch

import oracle as O
from packing_amd.geometry import random_polygon, reference_tree


def require_gpu():
    assert torch.cuda.is_available(), "GPU tests need a GPU (run with -m gpu on the MI355X)"
    from packing_amd import _native
    _native.lib()


def poses(seed, n_ens, n, dtype=np.float32):
    rng = np.random.default_rng(seed)
    from packing_amd.geometry import make_poses
    return make_poses(rng, n_ens, n, dtype)


def dense_poses(seed, n_ens, n, spread=0.55, dtype=np.float32):
    """Denser than the benchmark start so that most trees overlap several neighbours."""
    rng = np.random.default_rng(seed)
    L = spread * np.sqrt(n)
    x = np.empty((n_ens, n, 3))
    x[..., 0] = rng.uniform(0, L, (n_ens, n))
    x[..., 1] = rng.uniform(0, L, (n_ens, n))
    x[..., 2] = rng.uniform(0, 2 * np.pi, (n_ens, n))
    return x.astype(dtype)


def assert_cost_close(got, ref, rtol, atol=0.0, what="cost"):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    err = np.abs(got - ref)
    bad = err > rtol * np.abs(ref) + atol
    assert not bad.any(), f"{what}: max rel err {np.max(err / (np.abs(ref) + 1e-300)):.3e} at {np.argmax(bad)}"


def assert_grad_close(got, ref, rtol, what="grad"):
    """|g - g64| <= rtol * (|g64| + max |g64| of the instance)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1).reshape(-1, *([1] * (ref.ndim - 1)))
    err = np.abs(got - ref)
    bad = err > rtol * (np.abs(ref) + scale)
    assert not bad.any(), (f"{what}: {bad.sum()} elements off, worst "
                           f"{np.max(err / (np.abs(ref) + scale + 1e-300)):.3e}")


GEOMS = {"tree": reference_tree, "random": lambda: random_polygon(0)}
