"""Continuous batching: requests admitted while others run, HBM-budget admission, and
every served result bitwise equal to relaxing that request alone."""
import numpy as np
import pytest
import torch

from gpu_util import GEOMS, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


@pytest.mark.parametrize("mode", ["area", "lut"])
def test_served_equals_alone(mode):
    from packing_amd.context import set_geometry
    from packing_amd.pack_cuda_lut import LookupTable
    from packing_amd.relax import RelaxBatch, RelaxOptions
    from packing_amd.serving import RelaxServer, poisson_requests
    geom = GEOMS["random"]()
    set_geometry(geom)
    kw = {"lut": LookupTable.build_exact(96, 96, 64, full_x=True)} if mode == "lut" else {}
    opts = RelaxOptions(mode=mode, max_iter=60, history=8, max_step=0.1, **kw)
    reqs = poisson_requests(40, rate=2000.0, n_lo=10, n_hi=60)
    srv = RelaxServer(opts, n_slots=16, n_max=60, chunk=20)
    rep = srv.serve(reqs)
    assert rep["requests"] == 40 and rep["launches"] > 3
    # requests that arrived while others were running were admitted later than t=0
    assert max(r.t_admit for r in reqs) > 0
    for r in reqs[::7]:
        b = RelaxBatch([r.xyt], r.h[None], opts, geometry=geom)
        b.run()
        assert torch.equal(b.xyt(0).cpu(), r.result), r.rid


def test_budget_limits_concurrency():
    from packing_amd.context import set_geometry
    from packing_amd.relax import RelaxOptions
    from packing_amd.serving import RelaxServer, poisson_requests, slot_bytes
    set_geometry(GEOMS["tree"]())
    opts = RelaxOptions(mode="area", max_iter=30, history=4, max_step=0.1)
    per = slot_bytes(40, 4)
    srv = RelaxServer(opts, n_slots=64, n_max=40, chunk=10, hbm_budget_bytes=3 * per)
    assert srv.S == 3
    seen = []
    orig = srv._launch

    def spy():
        seen.append(len(srv.busy))
        orig()

    srv._launch = spy
    rep = srv.serve(poisson_requests(12, rate=5000.0, n_lo=10, n_hi=40), realtime=False)
    assert rep["requests"] == 12 and max(seen) <= 3
    with pytest.raises(MemoryError):
        RelaxServer(opts, n_slots=4, n_max=40, hbm_budget_bytes=per - 1)


def test_mixed_sizes_saturation_and_one_step_in_flight():
    """Requests of every size class at once (saturation), served with two steps in flight
    and with one: the same results, each bitwise the result of relaxing it alone, and
    every request's latency covers its host-to-device copy and its result's return."""
    from packing_amd.context import set_geometry
    from packing_amd.pack_cuda_lut import LookupTable
    from packing_amd.relax import RelaxBatch, RelaxOptions
    from packing_amd.serving import RelaxServer, poisson_requests, size_class
    geom = GEOMS["random"]()
    set_geometry(geom)
    lut = LookupTable.build_exact(96, 96, 64, full_x=True)
    opts = RelaxOptions(mode="lut", lut=lut, max_iter=50, history=8, max_step=0.1)
    reqs = poisson_requests(30, rate=1e6, n_lo=10, n_hi=200, seed_requests=11)
    assert len({size_class(r.xyt.shape[0]) for r in reqs}) >= 3
    out = {}
    for fl in (2, 1):
        rq = poisson_requests(30, rate=1e6, n_lo=10, n_hi=200, seed_requests=11)
        srv = RelaxServer(opts, n_slots=12, n_max=200, chunk=20, in_flight=fl)
        rep = srv.serve(rq, realtime=False)
        assert rep["requests"] == 30 and rep["service_p50_s"] > 0
        assert all(r.result.device.type == "cpu" and r.t_done > r.t_admit for r in rq)
        out[fl] = {r.rid: r.result for r in rq}
    for r in reqs[::4]:
        b = RelaxBatch([r.xyt], r.h[None], opts, geometry=geom)
        b.run()
        a = b.xyt(0).cpu()
        assert torch.equal(a, out[2][r.rid]) and torch.equal(a, out[1][r.rid]), r.rid


def test_admission_buffers_survive_the_next_step():
    """Each step's admitted slots and tree counts are staged in a host buffer of their own:
    the step's non-blocking copy may run after the host has queued the next step, so a
    shared buffer could hand it the next step's slots (and slot ids as tree counts)."""
    from packing_amd.context import set_geometry
    from packing_amd.relax import RelaxOptions
    from packing_amd.serving import RelaxServer, poisson_requests
    set_geometry(GEOMS["tree"]())
    opts = RelaxOptions(mode="area", max_iter=20, history=4, max_step=0.1)
    srv = RelaxServer(opts, n_slots=8, n_max=40, chunk=5, in_flight=2)
    reqs = poisson_requests(6, rate=1e6, n_lo=10, n_hi=40)
    srv._admit(reqs[0], 0.0)
    srv._admit(reqs[1], 0.0)
    srv._launch()
    first = srv.adm_h[0, :4].tolist()
    srv._admit(reqs[2], 0.0)
    srv._launch()
    # the first step's staging is untouched by the second step's admission
    assert srv.adm_h[0, :4].tolist() == first
    assert srv.adm_h[1, 0].item() == reqs[2].slot and srv.adm_h[1, 1].item() == reqs[2].xyt.shape[0]
    torch.cuda.synchronize()
    nt = srv.ntree.cpu().numpy()
    for r in reqs[:3]:
        assert nt[r.slot] == r.xyt.shape[0]


def test_step_refuses_tree_count_above_bound():
    """A slot whose tree count exceeds the launch's n_tree_max is not run: status -1, its
    state untouched (the kernel's LDS and rows are sized for n_tree_max)."""
    import ctypes as C
    from packing_amd import _native as N
    from packing_amd.context import device_geometry, set_geometry
    from packing_amd.relax import RelaxOptions
    from packing_amd.serving import RelaxServer, poisson_requests
    geom = GEOMS["tree"]()
    set_geometry(geom)
    opts = RelaxOptions(mode="area", max_iter=20, history=4, max_step=0.1)
    srv = RelaxServer(opts, n_slots=4, n_max=40, chunk=5, in_flight=1)
    reqs = poisson_requests(2, rate=1e6, n_lo=30, n_hi=40)
    for r in reqs:
        srv._admit(r, 0.0)
    srv._reset_pending()
    torch.cuda.synchronize()
    x0 = srv.x.clone()
    order = torch.tensor([r.slot for r in reqs], dtype=torch.int32, device=srv.x.device)
    p = srv._params(order.data_ptr(), 2, 20)   # bound below both tree counts
    st = torch.cuda.current_stream()
    N.check(N.lib().pk_relax_step(device_geometry(srv.x.device, geom).handle, C.byref(p), st.cuda_stream),
            "pk_relax_step")
    torch.cuda.synchronize()
    s = srv.status.cpu().numpy()
    assert all(s[r.slot] == -1 for r in reqs)
    assert torch.equal(srv.x, x0)
