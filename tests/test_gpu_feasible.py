"""FP64 feasibility kernel (pk_feasible) against the oracle's restatement of the metric's
collision rule, including exact touching (allowed) and one-ulp penetration."""
import numpy as np
import pytest
import torch

from gpu_util import GEOMS, dense_poses, poses, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


@pytest.mark.parametrize("gname", ["tree", "random"])
def test_counts_match_oracle(oracle_mod, gname):
    from packing_amd.context import set_geometry
    from packing_amd.feasibility import overlapping_pairs
    geom = GEOMS[gname]()
    set_geometry(geom)
    og = oracle_mod.OracleGeometry(geom)
    for x in (dense_poses(1, 12, 60, dtype=np.float64), poses(2, 12, 60, dtype=np.float64)):
        got = overlapping_pairs(torch.tensor(x, device="cuda")).cpu().numpy()
        ref = oracle_mod.feasible_pairs(og, x)
        assert np.array_equal(got, ref), (got, ref)
    # float32 input is widened to double exactly
    x32 = dense_poses(3, 8, 40)
    got = overlapping_pairs(torch.tensor(x32, device="cuda")).cpu().numpy()
    assert np.array_equal(got, oracle_mod.feasible_pairs(og, x32.astype(np.float64)))


def test_touching_allowed_and_ulp_penetration(oracle_mod):
    """Tree j stacked on tree i: j's trunk bottom edge exactly on i's tip is touching
    (feasible); one ulp lower the tip is inside (one overlapping pair)."""
    from packing_amd.context import set_geometry
    from packing_amd.feasibility import overlapping_pairs
    geom = GEOMS["tree"]()
    set_geometry(geom)
    og = oracle_mod.OracleGeometry(geom)
    tip = geom.outline[0, 1]
    bottom = geom.outline[7, 1]
    y = tip - bottom
    while y + bottom < tip:
        y = np.nextafter(y, np.inf)
    while y + bottom > tip:
        y = np.nextafter(y, -np.inf)
    assert y + bottom == tip
    for yj, expect in ((y, 0), (np.nextafter(y, -np.inf), 1), (y + 1e-9, 0), (y - 1e-9, 1)):
        x = np.array([[[0.0, 0.0, 0.0], [0.0, yj, 0.0]]])
        got = int(overlapping_pairs(torch.tensor(x, device="cuda"))[0])
        ref = int(oracle_mod.feasible_pairs(og, x)[0])
        assert got == ref == expect, (yj, got, ref, expect)
