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


def lut_face_guard(lut, xyt, radius, band_cells=1e-4):
    """(E, N) mask of trees in a pair whose FP64 relative pose lies within band_cells of a
    cell face of the table (or of the mirror plane x = 0 of a half table, or of the table's
    range).  The trilinear gradient jumps across a face, and rounding the relative pose to
    float32 moves it by at most ~3e-5 cells at these sizes (|dx| <= 2R, coordinates <= 8,
    a few ulps of each; 2pi/(nt-1) per theta cell), so only such pairs may see the
    neighbouring cell's gradient.  band_cells = 1e-4 is 3x that bound."""
    x = np.asarray(xyt, dtype=np.float64)
    E, N = x.shape[:2]
    guard = np.zeros((E, N), dtype=bool)
    cut = 4.0 * radius * radius
    nx, ny, nt = lut.N_x, lut.N_y, lut.N_theta
    for e in range(E):
        X = x[e]
        dx = X[None, :, 0] - X[:, None, 0]
        dy = X[None, :, 1] - X[:, None, 1]
        c, s = np.cos(X[:, 2])[:, None], np.sin(X[:, 2])[:, None]
        xl, yl = c * dx + s * dy, -s * dx + c * dy
        th = np.mod(X[None, :, 2] - X[:, None, 2] + np.pi, 2 * np.pi) - np.pi
        if lut.mirror_x:
            near_mirror = np.abs(xl) < band_cells * lut.grid_dx
            th = np.where(xl < 0, -th, th)
            xl = np.abs(xl)
        else:
            near_mirror = np.zeros_like(xl, dtype=bool)
        fx = (xl - lut.X_min) / (lut.X_max - lut.X_min) * (nx - 1)
        fy = (yl - lut.Y_min) / (lut.Y_max - lut.Y_min) * (ny - 1)
        ft = (th - lut.theta_min) / (lut.theta_max - lut.theta_min) * (nt - 1)
        near = near_mirror.copy()
        for f in (fx, fy, ft):
            fr = f - np.floor(f)
            near |= np.minimum(fr, 1 - fr) < band_cells
        near |= np.abs(np.abs(th) - np.pi) < band_cells * lut.grid_dtheta   # the +-pi wrap
        iu = np.triu(np.ones((N, N), dtype=bool), 1) & (dx * dx + dy * dy <= cut)
        bad = near & iu
        guard[e] = bad.any(axis=1) | bad.any(axis=0)
    return guard
