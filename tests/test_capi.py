"""The C-ABI library loads and exports every symbol include/packing.h declares (no GPU)."""
import os
import re

from packing_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "packing.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pk_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header():
    L = _native.lib()
    names = declared()
    assert names, "no declarations parsed"
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_native.exported_symbols()) == declared()


def test_version():
    assert _native.lib().pk_version() == 1


def test_lut_create_rejects_oversized_table():
    """Tables of 2^32 nodes or more are refused before any device work (no GPU needed);
    the reference's default 900^3 grid is below the limit."""
    import ctypes as C
    L = _native.lib()
    h = C.c_void_p()
    rc = L.pk_lut_create(C.c_void_p(16), 2048, 2048, 1024, -1.0, 1.0, -1.0, 1.0, -3.0, 3.0, 0, 0, C.byref(h))
    assert rc != 0
    assert b"2^32" in L.pk_last_error()
    assert 900 ** 3 < 2 ** 32


def test_lut_create_rejects_asymmetric_mirror_table():
    """A mirrored (half) table looks x < 0 up as (-x, -theta); with a theta range that is
    not symmetric about 0 that index could fall below the first node, so such a table is
    refused before any device work (no GPU needed)."""
    import ctypes as C
    L = _native.lib()
    h = C.c_void_p()
    rc = L.pk_lut_create(C.c_void_p(16), 8, 8, 8, -0.1, 1.0, -1.0, 1.0, -3.0, 3.1, 0, 1, C.byref(h))
    assert rc != 0
    assert b"symmetric" in L.pk_last_error()


def test_geometry_create_rejects_far_vertices():
    """Pair terms are summed in fixed point; the bound on them needs every piece vertex
    within 8 of the tree origin, checked before any device work (no GPU needed)."""
    import ctypes as C
    L = _native.lib()
    h = C.c_void_p()
    sq = (C.c_double * 8)(*[9.0, 0.0, 10.0, 0.0, 10.0, 1.0, 9.0, 1.0])
    nv = (C.c_int32 * 1)(4)
    rc = L.pk_geometry_create(sq, nv, 1, 4, sq, 4, 10.1, C.byref(h))
    assert rc != 0
    assert b"within 8" in L.pk_last_error()
