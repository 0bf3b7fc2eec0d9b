"""Fused relaxation engine (pk_relax_step) against the FP64 oracle's per-instance L-BFGS,
plus the invariances the batched and served paths rely on (all bitwise)."""
import numpy as np
import pytest
import torch

from gpu_util import GEOMS, dense_poses, poses, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def batch(geom, xs, hs, **kw):
    from packing_amd.relax import RelaxBatch, RelaxOptions
    opts = RelaxOptions(**kw)
    return RelaxBatch(xs, hs, opts, geometry=geom)


def ragged(seed, sizes, dense=True):
    xs = [(dense_poses if dense else poses)(seed + k, 1, n)[0] for k, n in enumerate(sizes)]
    hs = np.array([[1.2 * np.sqrt(n), 0.6 * np.sqrt(n), 0.6 * np.sqrt(n)] for n in sizes])
    return xs, hs


@pytest.mark.parametrize("gname", ["tree", "random"])
@pytest.mark.parametrize("mode", ["area", "sep"])
@pytest.mark.parametrize("iters", [1, 2, 4])
def test_first_iterations_vs_oracle(oracle_mod, gname, mode, iters):
    """Early iterates agree with the FP64 oracle (same algorithm, float32 arithmetic)."""
    geom = GEOMS[gname]()
    sizes = [5, 12, 30, 41]
    xs, hs = ragged(3, sizes)
    b = batch(geom, xs, hs, mode=mode, max_iter=iters, history=8, max_step=0.1)
    b.run()
    torch.cuda.synchronize()
    og = oracle_mod.OracleGeometry(geom)
    for i, n in enumerate(sizes):
        ref = oracle_mod.relax(og, xs[i][None].astype(np.float64), hs[i][None], use_sep=mode == "sep",
                               max_iter=iters, m=8, max_step=0.1)[0]
        got = b.xyt(i).cpu().numpy()
        # each step moves a tree by <= max_step; float32 vs float64 agree to 1e-4 of that
        assert np.abs(got - ref).max() < 2e-5, (i, np.abs(got - ref).max())


@pytest.mark.parametrize("gname,full_x", [("tree", False), ("random", True)])
@pytest.mark.parametrize("iters", [1, 2, 4])
def test_first_iterations_vs_oracle_table(oracle_mod, gname, full_x, iters):
    """Table mode (the benchmark's cost): early iterates agree with the FP64 oracle's
    L-BFGS over the same float32 table (mirrored half table for the symmetric tree)."""
    from packing_amd.context import set_geometry
    from packing_amd.pack_cuda_lut import LookupTable
    geom = GEOMS[gname]()
    set_geometry(geom)
    lut = LookupTable.build_exact(96, 96, 64, full_x=full_x)
    OL = oracle_mod.OracleLut(lut.X, lut.Y, lut.theta, lut.lut_d.cpu().numpy(), mirror=lut.mirror_x)
    sizes = [5, 12, 30, 41]
    xs, hs = ragged(3, sizes)
    b = batch(geom, xs, hs, mode="lut", lut=lut, max_iter=iters, history=8, max_step=0.1)
    b.run()
    torch.cuda.synchronize()
    og = oracle_mod.OracleGeometry(geom)
    for i, n in enumerate(sizes):
        ref = oracle_mod.relax(og, xs[i][None].astype(np.float64), hs[i][None], max_iter=iters, m=8,
                               max_step=0.1, lut=OL)[0]
        got = b.xyt(i).cpu().numpy()
        assert np.abs(got - ref).max() < 2e-5, (i, np.abs(got - ref).max())
        if iters == 4:
            assert np.abs(got - xs[i]).max() > 1e-3   # the poses did move


def test_invariances():
    """Alone vs in a batch, plain vs batched, chunked vs one launch, any block size,
    graph replay: bitwise the same poses."""
    geom = GEOMS["tree"]()
    sizes = [3, 17, 40, 64, 9, 25]
    xs, hs = ragged(11, sizes)
    kw = dict(mode="area", max_iter=30, history=8, max_step=0.1)
    ref = batch(geom, xs, hs, **kw)
    ref.run()
    R = [ref.xyt(i).cpu().numpy() for i in range(len(sizes))]
    for blk in (1, 64, 128, 256):
        b = batch(geom, xs, hs, block=blk, **kw)
        b.run(iters_per_launch=7)
        for i in range(len(sizes)):
            assert np.array_equal(b.xyt(i).cpu().numpy(), R[i]), (blk, i)
    p = batch(geom, xs, hs, **kw)
    p.run_plain()
    for i in range(len(sizes)):
        assert np.array_equal(p.xyt(i).cpu().numpy(), R[i]), ("plain", i)
    for i in (1, 3):
        a = batch(geom, [xs[i]], hs[i:i + 1], **kw)
        a.run()
        assert np.array_equal(a.xyt(0).cpu().numpy(), R[i]), ("alone", i)
    g = batch(geom, xs, hs, **kw)
    graph = g.capture_graph(10)
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
    for i in range(len(sizes)):
        assert np.array_equal(g.xyt(i).cpu().numpy(), R[i]), ("graph", i)


@pytest.mark.parametrize("mode", ["area", "sep"])
def test_relaxation_reaches_zero_overlap(oracle_mod, mode):
    """Sparse start (the benchmark's pose distribution): 300 iterations remove every
    overlap; the FP64 oracle confirms the final packings."""
    geom = GEOMS["random"]()
    sizes = [10, 20, 50, 100]
    xs, hs = ragged(5, sizes, dense=False)
    b = batch(geom, xs, hs, mode=mode, max_iter=300, history=8, max_step=0.1)
    b.run()
    torch.cuda.synchronize()
    og = oracle_mod.OracleGeometry(geom)
    for i in range(len(sizes)):
        x = b.xyt(i).cpu().numpy().astype(np.float64)[None]
        c0, _ = oracle_mod.overlap(og, xs[i][None].astype(np.float64), False, grads=False)
        c1, _ = oracle_mod.overlap(og, x, False, grads=False)
        if mode == "area":
            assert c1[0] < 1e-5 * max(c0[0], 1e-3), (i, c0, c1)
        else:
            # squared SAT depth is flat near contact: most, not all, overlap goes
            assert c1[0] < 0.1 * max(c0[0], 1e-3), (i, c0, c1)
    rec = b.slot_records()
    assert (rec[:, 3] == 2).all()   # every slot finished


def test_lut_mode_relaxes():
    from packing_amd.pack_cuda_lut import LookupTable
    from packing_amd.context import set_geometry
    geom = GEOMS["tree"]()
    set_geometry(geom)
    lut = LookupTable.build_exact(96, 96, 64)
    sizes = [8, 30, 60]
    xs, hs = ragged(2, sizes)
    b = batch(geom, xs, hs, mode="lut", lut=lut, max_iter=50, history=8, max_step=0.1)
    l0 = None
    b.launch(1)
    torch.cuda.synchronize()
    l0 = b.loss.cpu().numpy().copy()
    b.run()
    torch.cuda.synchronize()
    assert (b.loss.cpu().numpy() < l0).all()


@pytest.mark.parametrize("mode", ["lut", "area"])
def test_wave_kernel_matches_lds_kernel(mode):
    """The one-wave kernel (state and history in registers) and the LDS kernel give the
    same bits, also when a run switches kernels between launches (the stored history is
    read back by either), near the one-wave size limit and with a short history."""
    from packing_amd.context import set_geometry
    from packing_amd.pack_cuda_lut import LookupTable
    geom = GEOMS["random"]()
    set_geometry(geom)
    kw = dict(mode=mode, max_iter=40, history=8, max_step=0.1)
    if mode == "lut":
        kw["lut"] = LookupTable.build_exact(96, 96, 64, full_x=True)
    sizes = [212, 150, 21, 22, 43, 5] if mode == "lut" else [60, 21, 22, 43, 5]
    xs, hs = ragged(7, sizes)
    for hist in (8, 3):
        kw["history"] = hist
        ref = batch(geom, xs, hs, block=128, **kw)
        ref.run()
        one = batch(geom, xs, hs, block=1, **kw)
        one.run(iters_per_launch=13)
        mix = batch(geom, xs, hs, block=1, **kw)
        for k, blk in enumerate((1, 256, 1, 64)):
            mix.opts.block = blk
            mix.launch(10)
        for i in range(len(sizes)):
            r = ref.xyt(i).cpu().numpy()
            assert np.array_equal(one.xyt(i).cpu().numpy(), r), ("one", hist, i)
            assert np.array_equal(mix.xyt(i).cpu().numpy(), r), ("mix", hist, i)
        assert np.array_equal(one.slot_records(), ref.slot_records())


def test_wave_kernel_limits():
    """Block 1 takes up to 212 trees and history 8; beyond that the library's own choice
    falls back to the LDS kernel and an explicit block 1 is rejected."""
    geom = GEOMS["tree"]()
    xs, hs = ragged(2, [213])
    b = batch(geom, xs, hs, mode="area", max_iter=3, history=8, max_step=0.1, block=1)
    with pytest.raises(ValueError):
        b.run()
    a = batch(geom, xs, hs, mode="area", max_iter=3, history=8, max_step=0.1)
    a.run()
    c = batch(geom, xs, hs, mode="area", max_iter=3, history=8, max_step=0.1, block=256)
    c.run()
    assert np.array_equal(a.xyt(0).cpu().numpy(), c.xyt(0).cpu().numpy())
    h = batch(geom, *ragged(3, [30]), mode="area", max_iter=3, history=9, max_step=0.1, block=1)
    with pytest.raises(ValueError):
        h.run()
