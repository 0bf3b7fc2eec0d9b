"""Process-wide polygon selection and its device handles.

The reference reads one global polygon (kgs.convex_breakdown / kgs.tree_vertices,
core/kaggle_support.py:327-354) into constant memory on first use
(core/pack_cuda.py:852-1032).  Here the active polygon is a Geometry object and each
device gets its own handle (pk_geometry_create), created on first use.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _native as N
from .geometry import Geometry, MAX_VERTS, reference_tree

_lock = threading.Lock()
_active: Geometry | None = None
_handles: dict = {}


def set_geometry(g: Geometry) -> None:
    """Select the polygon every kernel entry point uses from now on."""
    global _active
    g.validate()
    with _lock:
        _active = g


def get_geometry() -> Geometry:
    global _active
    with _lock:
        if _active is None:
            _active = reference_tree()
        return _active


class DeviceGeometry:
    def __init__(self, g: Geometry, device: int):
        self.geometry = g
        self.device = device
        xy, nv = g.piece_table(MAX_VERTS)
        self._xy = np.ascontiguousarray(xy, dtype=np.float64)
        self._nv = np.ascontiguousarray(nv, dtype=np.int32)
        self._ol = np.ascontiguousarray(g.outline, dtype=np.float64)
        h = C.c_void_p()
        import torch
        with torch.cuda.device(device):
            N.check(N.lib().pk_geometry_create(
                self._xy.ctypes.data_as(C.POINTER(C.c_double)),
                self._nv.ctypes.data_as(C.POINTER(C.c_int32)), g.n_pieces, MAX_VERTS,
                self._ol.ctypes.data_as(C.POINTER(C.c_double)), len(self._ol), float(g.radius),
                C.byref(h)), "pk_geometry_create")
        self.handle = h.value

    def __del__(self):
        try:
            if self.handle:
                N.lib().pk_geometry_destroy(self.handle)
        except Exception:
            pass


def device_geometry(device=None, g: Geometry | None = None) -> DeviceGeometry:
    import torch
    if device is None:
        device = torch.cuda.current_device()
    elif isinstance(device, torch.device):
        device = device.index if device.index is not None else torch.cuda.current_device()
    g = g or get_geometry()
    key = (id(g), g.fingerprint(), int(device))
    with _lock:
        dg = _handles.get(key)
        if dg is None:
            dg = DeviceGeometry(g, int(device))
            _handles[key] = dg
        return dg
