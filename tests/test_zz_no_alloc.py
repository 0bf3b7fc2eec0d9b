"""Allocation-free relaxation (pk_relax_run_noalloc): the same bits as pk_relax_step over
fresh slots, on two calls in a row (the second smaller, so leftovers of the first in the
shared scratch would show), and a clear refusal of batches larger than the scratch."""
import ctypes as C

import numpy as np
import pytest
import torch

from gpu_util import GEOMS, dense_poses, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def _batch(geom, sizes, seed, **kw):
    from packing_amd.relax import RelaxBatch, RelaxOptions
    xs = [dense_poses(seed + k, 1, n)[0] for k, n in enumerate(sizes)]
    hs = np.array([[1.2 * np.sqrt(n), 0.6 * np.sqrt(n), 0.6 * np.sqrt(n)] for n in sizes])
    return RelaxBatch(xs, hs, RelaxOptions(**kw), geometry=geom)


@pytest.mark.parametrize("mode", ["area", "lut"])
def test_noalloc_equals_relax_step(mode):
    from packing_amd.context import set_geometry
    from packing_amd.pack_cuda_lut import LookupTable
    geom = GEOMS["random"]()
    set_geometry(geom)
    kw = dict(mode=mode, max_iter=60, history=8, max_step=0.1)
    if mode == "lut":
        kw["lut"] = LookupTable.build_exact(96, 96, 64, full_x=True)
    for call, (sizes, seed) in enumerate((([12, 40, 77, 150, 33, 200, 5], 3), ([9, 20, 31], 50))):
        ref = _batch(geom, sizes, seed, **kw)
        ref.run(iters_per_launch=25)
        got = _batch(geom, sizes, seed, **kw)
        got.run_noalloc()
        torch.cuda.synchronize()
        assert torch.equal(got.x, ref.x), call
        assert torch.equal(got.loss, ref.loss), call
        assert torch.equal(got.status, ref.status), call
        assert not got.state.any()          # the batch's own records were not used


def test_noalloc_refuses_oversized_batch():
    from packing_amd import _native as N
    geom = GEOMS["tree"]()
    b = _batch(geom, [10, 10], 1, mode="area", max_iter=5, history=16, max_step=0.1)
    p = b._params(5)
    p.g = p.s_hist = p.y_hist = p.state = None
    n_state = N.PK_RELAX_SCRATCH_BYTES // 4 // 33 + 1        # (1 + 2 * 16) floats each
    rc = N.lib().pk_relax_run_noalloc(b._dg.handle, C.byref(p), n_state, N.stream_handle(None))
    assert rc != 0 and b"scratch" in N.lib().pk_last_error()
    p.g = b.g.data_ptr()
    rc = N.lib().pk_relax_run_noalloc(b._dg.handle, C.byref(p), b.n_state, N.stream_handle(None))
    assert rc != 0 and b"NULL" in N.lib().pk_last_error()
