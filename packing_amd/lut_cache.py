"""On-disk cache of pose-space tables (the role of the pickle cache of
core/pack_cost.py:1131-1168, in a documented binary format that executes nothing when
read).

File layout, little-endian:
    8 bytes   magic b"PKLUT\\x00\\x01\\x00"
    uint32    format version (1)
    uint32    N_x, N_y, N_theta
    uint32    flags: bit 0 apply_quadratic_transform, bit 1 mirrored half table
    uint32    key length L, then L bytes of UTF-8 key (the parameters and the polygon's
              fingerprint the table was built for)
    float32   X[N_x], Y[N_y], theta[N_theta]
    float32   vals[N_x][N_y][N_theta] (theta fastest)
A file whose key differs from the requested one is treated as absent.
"""
from __future__ import annotations

import os
import struct

import numpy as np

MAGIC = b"PKLUT\x00\x01\x00"
VERSION = 1


def cache_dir() -> str:
    """$PACKING_AMD_CACHE, else ~/.cache/packing_amd, plus lut_cache/."""
    base = os.environ.get("PACKING_AMD_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "packing_amd")
    return os.path.join(base, "lut_cache")


def save(path: str, key: str, lut) -> None:
    import torch
    vals = lut.vals.detach().cpu().numpy() if isinstance(lut.vals, torch.Tensor) else np.asarray(lut.vals)
    vals = np.ascontiguousarray(vals, dtype="<f4")
    kb = key.encode()
    flags = (1 if lut.apply_quadratic_transform else 0) | (2 if lut.mirror_x else 0)
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<6I", VERSION, lut.N_x, lut.N_y, lut.N_theta, flags, len(kb)))
        f.write(kb)
        for a in (lut.X, lut.Y, lut.theta):
            f.write(np.ascontiguousarray(a, dtype="<f4").tobytes())
        f.write(vals.tobytes())
    os.replace(tmp, path)


def load(path: str, key: str):
    """(X, Y, theta, vals, quad, mirror) from a cache file written for key, else None."""
    if not os.path.exists(path):
        return None
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            return None
        hdr = f.read(24)
        if len(hdr) != 24:
            return None
        ver, nx, ny, nt, flags, kl = struct.unpack("<6I", hdr)
        if ver != VERSION or f.read(kl).decode(errors="replace") != key:
            return None
        X = np.frombuffer(f.read(4 * nx), dtype="<f4")
        Y = np.frombuffer(f.read(4 * ny), dtype="<f4")
        T = np.frombuffer(f.read(4 * nt), dtype="<f4")
        vals = np.frombuffer(f.read(4 * nx * ny * nt), dtype="<f4")
    if vals.size != nx * ny * nt or len(X) != nx or len(Y) != ny or len(T) != nt:
        return None
    return X.copy(), Y.copy(), T.copy(), vals.reshape(nx, ny, nt), bool(flags & 1), bool(flags & 2)
