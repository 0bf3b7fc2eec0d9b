"""Table cache file format (no GPU): a round trip keeps every array bit for bit, and a
file written for other parameters or another polygon reads as absent."""
import numpy as np

from packing_amd import lut_cache
from packing_amd.pack_cuda_lut import LookupTable


def test_round_trip_and_key(tmp_path):
    rng = np.random.default_rng(0)
    X = np.linspace(-0.1, 1.0, 7, dtype=np.float32)
    Y = np.linspace(-1.0, 1.0, 5, dtype=np.float32)
    T = np.linspace(-np.pi, np.pi, 9, dtype=np.float32)
    vals = rng.normal(size=(7, 5, 9)).astype(np.float32)
    lut = LookupTable(X=X, Y=Y, theta=T, vals=vals, apply_quadratic_transform=True, mirror_x=True)
    path = str(tmp_path / "t.pklut")
    lut_cache.save(path, "exact_sep_key_A", lut)
    X2, Y2, T2, v2, quad, mirror = lut_cache.load(path, "exact_sep_key_A")
    for a, b in ((X, X2), (Y, Y2), (T, T2), (vals, v2)):
        assert a.tobytes() == b.tobytes()
    assert quad and mirror
    assert lut_cache.load(path, "exact_sep_key_B") is None
    assert lut_cache.load(str(tmp_path / "missing.pklut"), "exact_sep_key_A") is None
    with open(path, "r+b") as f:        # a corrupted magic reads as absent
        f.write(b"XXXX")
    assert lut_cache.load(path, "exact_sep_key_A") is None
