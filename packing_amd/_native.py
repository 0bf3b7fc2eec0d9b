"""ctypes binding of libpacking.so (C-ABI declared in include/packing.h).

The library must have been built (packing_amd/_build.py); there is no fallback: every
entry point raises if the native library cannot be loaded.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpacking.so")

PK_F32, PK_F64, PK_I32 = 0, 1, 2
PK_MODE_AREA, PK_MODE_SEP, PK_MODE_LUT = 0, 1, 2
PK_MAX_HISTORY = 16
PK_LBFGS_SLOT_BYTES = 184
PK_LBFGS_SYS_BYTES = 48
PK_LBFGS_LS_BYTES = 160
PK_RELAX_SCRATCH_BYTES = 128 << 20

_lib = None
_lock = threading.Lock()

vp = C.c_void_p
i32 = C.c_int


# name -> (restype, argtypes)
_SIGS = {
    "pk_last_error": (C.c_char_p, []),
    "pk_version": (i32, []),
    "pk_device_info": (i32, [C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
    "pk_geometry_create": (i32, [C.POINTER(C.c_double), C.POINTER(C.c_int32), i32, i32,
                                 C.POINTER(C.c_double), i32, C.c_double, C.POINTER(vp)]),
    "pk_geometry_destroy": (i32, [vp]),
    "pk_overlap": (i32, [vp, i32, vp, vp, i32, i32, i32, vp, i32, i32, vp, vp, i32, vp, vp]),
    "pk_boundary": (i32, [vp, i32, vp, vp, i32, i32, vp, vp, vp, vp, vp]),
    "pk_lut_create": (i32, [vp, i32, i32, i32, C.c_double, C.c_double, C.c_double, C.c_double,
                            C.c_double, C.c_double, i32, i32, C.POINTER(vp)]),
    "pk_lut_destroy": (i32, [vp]),
    "pk_lut_build": (i32, [vp, vp, i32, vp, i32, vp, i32, vp, vp]),
    "pk_sep_build": (i32, [vp, vp, i32, vp, i32, vp, i32, vp, vp]),
    "pk_lut_overlap": (i32, [vp, vp, vp, vp, i32, i32, vp, vp, i32, vp, vp]),
    "pk_relax_step": (i32, [vp, vp, vp]),
    "pk_relax_run_noalloc": (i32, [vp, vp, i32, vp]),
    "pk_relax_slot_bytes": (i32, []),
    "pk_relax_lds_bytes": (i32, [vp, i32, i32, i32, i32, i32]),
    "pk_relax_occupancy": (i32, [vp, i32, i32, i32, i32, i32, C.POINTER(i32)]),
    "pk_feasible": (i32, [vp, i32, vp, vp, i32, i32, vp, vp]),
    "pk_lbfgs_sys_bytes": (i32, []),
    "pk_lbfgs_begin": (i32, [vp, vp]),
    "pk_lbfgs_direction": (i32, [vp, i32, vp]),
    "pk_lbfgs_step": (i32, [vp, i32, vp]),
    "pk_lbfgs_ls_begin": (i32, [vp, vp, vp]),
    "pk_lbfgs_ls_gtd": (i32, [vp, vp, vp]),
    "pk_lbfgs_ls_bracket": (i32, [vp, vp, i32, vp]),
    "pk_lbfgs_ls_zoom_begin": (i32, [vp, vp, i32, vp]),
    "pk_lbfgs_ls_zoom_trial": (i32, [vp, vp, vp]),
    "pk_lbfgs_ls_zoom_update": (i32, [vp, vp, vp]),
    "pk_lbfgs_ls_end": (i32, [vp, vp, vp]),
    "pk_lap_hungarian": (i32, [i32, vp, i32, i32, vp, vp, vp, vp]),
    "pk_diversity": (i32, [vp, vp, i32, i32, i32, C.c_float, C.c_float, i32, vp, vp]),
    "pk_diversity_matrix": (i32, [vp, vp, i32, i32, i32, i32, i32, vp, vp]),
}


class RelaxParams(C.Structure):
    """ctypes mirror of pk_relax_params (include/packing.h)."""
    _fields_ = [
        ("n_slots", i32), ("xoff", vp), ("ntree", vp), ("x", vp), ("g", vp), ("s_hist", vp),
        ("y_hist", vp), ("state", vp), ("order", vp), ("n_launch", i32), ("n_tree_max", i32),
        ("mode", i32), ("w_overlap", C.c_float), ("w_boundary", C.c_float), ("w_area", C.c_float),
        ("fixed_h", i32), ("lut", vp), ("history", i32), ("max_iter", i32), ("max_eval", i32),
        ("lr", C.c_float), ("max_step", C.c_float), ("tol_grad", C.c_float),
        ("tol_change", C.c_float), ("tol_rel", C.c_float), ("stop_on_increase", i32),
        ("iters", i32), ("block", i32), ("status_out", vp), ("loss_out", vp), ("stats", vp),
        ("hist_lds", i32),
    ]


class LbfgsBatch(C.Structure):
    """ctypes mirror of pk_lbfgs_batch (include/packing.h)."""
    _fields_ = [("M", i32), ("n", i32), ("history", i32), ("dtype", i32), ("x", vp), ("g", vp),
                ("loss", vp), ("d", vp), ("g_prev", vp), ("s_hist", vp), ("y_hist", vp), ("ro", vp),
                ("al", vp), ("sys", vp), ("counters", vp), ("lr", C.c_double), ("tol_grad", C.c_double),
                ("tol_change", C.c_double), ("tol_rel", C.c_double), ("max_step", C.c_double),
                ("max_eval", i32), ("stop_on_increase", i32), ("line_search", i32)]


class LbfgsLs(C.Structure):
    """ctypes mirror of pk_lbfgs_ls (include/packing.h)."""
    _fields_ = [("f", vp), ("g", vp), ("x", vp), ("g_prev", vp), ("bg", vp), ("sys", vp), ("flag", vp),
                ("count", vp)]


class PackingError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            # torch ships its own HIP runtime under the same soname (libamdhip64.so.7) as
            # the one this library links; loading torch first makes both use that one
            # runtime (loading ours first leaves torch and this library on different
            # runtime instances, and device allocations then fail)
            import torch  # noqa: F401
            if not os.path.exists(LIB_PATH):
                raise PackingError(
                    f"{LIB_PATH} is missing: build it with `python -m packing_amd._build` "
                    "(or __graft_entry__.build())")
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(rc: int, what: str = "") -> None:
    if rc == 0:
        return
    msg = lib().pk_last_error().decode(errors="replace")
    if rc == -1:
        raise ValueError(f"{what}: {msg}")
    if rc == -3:
        raise MemoryError(f"{what}: {msg}")
    if rc == -4:
        raise ValueError(f"{what}: {msg}")
    raise PackingError(f"{what}: {msg} (status {rc})")


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_handle(stream=None) -> int | None:
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    h = getattr(stream, "cuda_stream", None)
    if h is None:
        h = int(stream)
    return h or None


def dtype_code(dtype) -> int:
    import torch
    if dtype == torch.float32:
        return PK_F32
    if dtype == torch.float64:
        return PK_F64
    raise ValueError(f"unsupported dtype {dtype}; float32 or float64")


def device_info():
    a, b, c = i32(), i32(), i32()
    check(lib().pk_device_info(C.byref(a), C.byref(b), C.byref(c)), "pk_device_info")
    return {"max_lds_per_block": a.value, "n_cu": b.value, "clock_khz": c.value}
