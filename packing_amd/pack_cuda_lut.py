"""Drop-in for the reference module ``pack_cuda_lut`` (core/pack_cuda_lut.py) on MI355X.

``LookupTable`` keeps the reference's fields and constructors; its values live in a
float32 torch tensor on the device and each device gets a pk_lut handle (instead of
the reference's constant-memory copy cached by ``id(lut)``, core/pack_cuda_lut.py:
811-815).  ``build_exact`` fills a table on the GPU from the exact clipping kernel
(pk_lut_build); ``build_from_function`` keeps the reference's host-callback form.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from . import _native as N
from .context import device_geometry, get_geometry


@dataclass
class LookupTable:
    X: np.ndarray
    Y: np.ndarray
    theta: np.ndarray
    vals: object                      # (N_x, N_y, N_theta) numpy array or torch tensor
    apply_quadratic_transform: bool = False
    mirror_x: bool | None = None      # None: mirror iff X starts just below 0 (half table)
    _dev: dict = field(default_factory=dict, repr=False, compare=False)

    def __post_init__(self):
        vals = self.vals
        n_x, n_y, n_theta = tuple(vals.shape)
        if len(self.X) != n_x or len(self.Y) != n_y or len(self.theta) != n_theta:
            raise ValueError(f"LUT shape mismatch: vals={tuple(vals.shape)}, X={len(self.X)}, "
                             f"Y={len(self.Y)}, theta={len(self.theta)}")
        self.N_x, self.N_y, self.N_theta = n_x, n_y, n_theta
        self.X_min, self.X_max = float(self.X[0]), float(self.X[-1])
        self.Y_min, self.Y_max = float(self.Y[0]), float(self.Y[-1])
        self.theta_min, self.theta_max = float(self.theta[0]), float(self.theta[-1])
        self.grid_dx = (self.X_max - self.X_min) / (n_x - 1) if n_x > 1 else 0.0
        self.grid_dy = (self.Y_max - self.Y_min) / (n_y - 1) if n_y > 1 else 0.0
        self.grid_dtheta = (self.theta_max - self.theta_min) / (n_theta - 1) if n_theta > 1 else 0.0
        if self.mirror_x is None:
            # the reference's default grid keeps one negative X node (core/pack_cuda_lut.py:128-131)
            self.mirror_x = self.X_min < 0 and abs(self.X_min) < 0.5 * abs(self.X_max)

    # -------------------------------------------------------------- device side
    @property
    def lut_d(self) -> torch.Tensor:
        return self.device_values(torch.device("cuda", torch.cuda.current_device()))

    def device_values(self, device) -> torch.Tensor:
        device = torch.device(device)
        if device.index is None:
            device = torch.device("cuda", torch.cuda.current_device())
        key = ("vals", device.index)
        t = self._dev.get(key)
        if t is None:
            v = self.vals
            t = (v.to(device=device, dtype=torch.float32) if isinstance(v, torch.Tensor)
                 else torch.tensor(np.asarray(v, dtype=np.float32), device=device)).contiguous()
            self._dev[key] = t
        return t

    def handle(self, device) -> int:
        device = torch.device(device)
        if device.index is None:
            device = torch.device("cuda", torch.cuda.current_device())
        key = ("h", device.index)
        h = self._dev.get(key)
        if h is None:
            vals = self.device_values(device)
            hv = C.c_void_p()
            N.check(N.lib().pk_lut_create(vals.data_ptr(), self.N_x, self.N_y, self.N_theta,
                                          self.X_min, self.X_max, self.Y_min, self.Y_max,
                                          self.theta_min, self.theta_max,
                                          int(self.apply_quadratic_transform), int(self.mirror_x),
                                          C.byref(hv)), "pk_lut_create")
            h = hv.value
            self._dev[key] = h
        return h

    def __del__(self):
        try:
            for k, v in list(self._dev.items()):
                if k[0] == "h" and v:
                    N.lib().pk_lut_destroy(v)
        except Exception:
            pass

    def __getstate__(self):
        s = self.__dict__.copy()
        s["_dev"] = {}
        if isinstance(s["vals"], torch.Tensor):
            s["vals"] = s["vals"].cpu().numpy()
        return s

    # -------------------------------------------------------------- builders
    @classmethod
    def default_grid(cls, N_x=400, N_y=400, N_theta=400, full_x=True, radius=None):
        R = get_geometry().radius if radius is None else radius
        X = np.linspace(-2 * R, 2 * R, N_x, dtype=np.float32)
        if not full_x:
            X = X[N_x // 2 - 1:]
        Y = np.linspace(-2 * R, 2 * R, N_y, dtype=np.float32)
        T = np.linspace(-np.pi, np.pi, N_theta, dtype=np.float32)
        return X, Y, T

    @classmethod
    def build_exact(cls, N_x=256, N_y=256, N_theta=128, X=None, Y=None, theta=None,
                    full_x=True, trim_zeros=False, device="cuda", verbose=False) -> "LookupTable":
        """Overlap-area table from the exact clipping kernel, on the GPU."""
        dX, dY, dT = cls.default_grid(N_x, N_y, N_theta, full_x)
        X = dX if X is None else np.asarray(X, np.float32)
        Y = dY if Y is None else np.asarray(Y, np.float32)
        T = dT if theta is None else np.asarray(theta, np.float32)
        dev = torch.device(device)
        Xt, Yt, Tt = (torch.tensor(a, device=dev) for a in (X, Y, T))
        vals = torch.empty(len(X), len(Y), len(T), dtype=torch.float32, device=dev)
        dg = device_geometry(dev)
        N.check(N.lib().pk_lut_build(dg.handle, Xt.data_ptr(), len(X), Yt.data_ptr(), len(Y),
                                     Tt.data_ptr(), len(T), vals.data_ptr(),
                                     N.stream_handle(None)), "pk_lut_build")
        if trim_zeros:
            Xn, Yn, Tn, vn = cls._trim_zero_edges(X, Y, T, vals, verbose=verbose)
            return cls(X=Xn, Y=Yn, theta=Tn, vals=vn.contiguous(), mirror_x=not full_x)
        return cls(X=X, Y=Y, theta=T, vals=vals, mirror_x=not full_x)

    @classmethod
    def build_separation(cls, N_x=900, N_y=900, N_theta=900, full_x=True, trim_zeros=True, device="cuda",
                         verbose=False) -> "LookupTable":
        """Signed separation-distance table (positive = penetration depth, negative =
        clearance) computed on the GPU in double by pk_sep_build, with the quadratic
        transform max(0, v)^2 applied per pair: the table of
        CollisionCostExactSeparation (core/pack_cost.py:1170-1207)."""
        X, Y, T = cls.default_grid(N_x, N_y, N_theta, full_x)
        dev = torch.device(device)
        Xt, Yt, Tt = (torch.tensor(a, device=dev) for a in (X, Y, T))
        vals = torch.empty(len(X), len(Y), len(T), dtype=torch.float32, device=dev)
        dg = device_geometry(dev)
        N.check(N.lib().pk_sep_build(dg.handle, Xt.data_ptr(), len(X), Yt.data_ptr(), len(Y),
                                     Tt.data_ptr(), len(T), vals.data_ptr(),
                                     N.stream_handle(None)), "pk_sep_build")
        if trim_zeros:
            X, Y, T, vals = cls._trim_zero_edges(X, Y, T, vals, verbose=verbose)
            vals = vals.contiguous()
        return cls(X=X, Y=Y, theta=T, vals=vals, apply_quadratic_transform=True, mirror_x=not full_x)

    @classmethod
    def build_from_function(cls, eval_fn, X: Optional[np.ndarray] = None, Y: Optional[np.ndarray] = None,
                            theta: Optional[np.ndarray] = None, N_x: int = 400, N_y: int = 400,
                            N_theta: int = 400, trim_zeros: bool = True, verbose: bool = True):
        """Reference form (core/pack_cuda_lut.py:91-169): eval_fn(dx, dy, theta) -> values."""
        dX, dY, dT = cls.default_grid(N_x, N_y, N_theta, full_x=False)
        X = dX if X is None else X
        Y = dY if Y is None else Y
        theta = dT if theta is None else theta
        vals = np.zeros((len(X), len(Y), len(theta)), dtype=np.float32)
        gx, gy = np.meshgrid(X, Y, indexing="ij")
        for k, t in enumerate(theta):
            vals[:, :, k] = np.asarray(eval_fn(gx.ravel(), gy.ravel(), t)).reshape(len(X), len(Y))
        if trim_zeros:
            X, Y, theta, vals = cls._trim_zero_edges(X, Y, theta, vals, verbose=verbose)
        return cls(X=X, Y=Y, theta=theta, vals=vals)

    @staticmethod
    def _trim_zero_edges(X, Y, theta, vals, verbose=True):
        """Drop leading/trailing X and Y slices with no positive value, keeping one
        non-positive slice on each side, and theta slices with none (no margin)
        (core/pack_cuda_lut.py:172-254)."""
        n_x, n_y, n_t = vals.shape

        def span(mx, margin):
            pos = np.where(mx > 0)[0]
            if len(pos) == 0:
                return 0, len(mx) - 1
            return max(0, pos[0] - margin), min(len(mx) - 1, pos[-1] + margin)

        if isinstance(vals, torch.Tensor):
            # device tables are reduced where they live (the reference default is ~1.5 GB)
            mx = [torch.amax(vals, dim=d).cpu().numpy() for d in ((1, 2), (0, 2), (0, 1))]
        else:
            mx = [np.max(vals, axis=d) for d in ((1, 2), (0, 2), (0, 1))]
        x0, x1 = span(mx[0], 1)
        y0, y1 = span(mx[1], 1)
        t0, t1 = span(mx[2], 0)
        if verbose:
            print(f"trim: X {n_x}->{x1 - x0 + 1}, Y {n_y}->{y1 - y0 + 1}, theta {n_t}->{t1 - t0 + 1}")
        return X[x0:x1 + 1], Y[y0:y1 + 1], theta[t0:t1 + 1], vals[x0:x1 + 1, y0:y1 + 1, t0:t1 + 1]


def overlap_multi_ensemble(xyt, out_cost, lut: LookupTable, out_grads=None, stream=None, block=0,
                           stats=None):
    """Table-based total overlap per ensemble (core/pack_cuda_lut.py:783-862)."""
    if xyt.ndim != 3 or xyt.shape[2] != 3:
        raise AssertionError(f"xyt shape: {tuple(xyt.shape)}")
    n_ens, n_trees = xyt.shape[0], xyt.shape[1]
    if tuple(out_cost.shape) != (n_ens,):
        raise AssertionError(f"out_cost shape: {tuple(out_cost.shape)}")
    if not xyt.is_contiguous():
        raise AssertionError("xyt must be C-contiguous")
    if xyt.dtype != torch.float32:
        raise ValueError("the table kernel is float32")
    if out_grads is not None and (tuple(out_grads.shape) != (n_ens, n_trees, 3) or not out_grads.is_contiguous()):
        raise AssertionError("out_grads must be (n_ens, n_trees, 3) and C-contiguous")
    if n_ens == 0:
        return
    dg = device_geometry(xyt.device)
    N.check(N.lib().pk_lut_overlap(dg.handle, lut.handle(xyt.device), xyt.data_ptr(), None, n_ens,
                                   n_trees, out_cost.data_ptr(), N.ptr(out_grads), int(block),
                                   N.ptr(stats), N.stream_handle(stream)), "pk_lut_overlap")
