"""Scarborough time-step statistics (src/tstepper/mod_tstepper_scarborough.f90)
on the GPU: mean_cc (get_mean_cc_9d / get_mean_cc_5d), dt0 and dt
(tstepper_dt0_scarborough / tstepper_dt_scarborough).  Arrays: numpy arrays
or CUDA tensors (all of one kind), coefficient arrays (9, N) or (5, N)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import SolverError, last_status


def _ptr(x, n):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            if not (x.dtype == torch.float64 and x.is_contiguous() and x.numel() == n):
                raise TypeError("contiguous float64 tensor of %d values expected" % n)
            return C.c_void_p(x.data_ptr())
    except ImportError:  # pragma: no cover
        pass
    x = np.ascontiguousarray(x, dtype=np.float64)
    if x.size != n:
        raise ValueError("%d values expected" % n)
    _ptr.keep.append(x)
    return C.c_void_p(x.ctypes.data)


_ptr.keep = []


def _done(v, where):
    st = last_status()
    _ptr.keep.clear()
    if st.status:
        raise SolverError(st.status, where)
    return v


def _i(v):
    return C.byref(C.c_int(int(v)))


def mean_cc(a, nx, ny):
    n = nx * ny
    size = a.numel() if hasattr(a, "numel") else np.size(a)
    nd = 9 if size == 9 * n else 5
    f = _lib.load().m2d_get_mean_cc_9d if nd == 9 else _lib.load().m2d_get_mean_cc_5d
    return _done(f(_i(nx), _i(ny), _ptr(a, nd * n)), "mean_cc")


def dt0(Jp, u, v, nx, ny):
    n = nx * ny
    return _done(_lib.load().m2d_tstepper_dt0_scarborough(_i(nx), _i(ny), _ptr(Jp, n),
                                                         _ptr(u, n), _ptr(v, n)), "dt0")


def dt(dt_old, au, av, at, ap, nx, ny, h0, mincc, maxcc):
    n = nx * ny
    d = lambda x: C.byref(C.c_double(float(x)))  # noqa: E731
    return _done(_lib.load().m2d_tstepper_dt_scarborough(
        _i(nx), _i(ny), d(dt_old), _ptr(au, 9 * n), _ptr(av, 9 * n), _ptr(at, 9 * n),
        _ptr(ap, 5 * n), d(h0), d(mincc), d(maxcc)), "dt")
