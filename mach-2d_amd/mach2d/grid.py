"""Grid centroids, metric terms and radii on the GPU: get_real_centroids_xy,
get_metrics and get_radius of src/grid/mod_grid.f90 (:237-300, :304-558,
:562-723), the steps of grid_create after the corner coordinates.

Arrays are numpy arrays or CUDA float64 tensors (all of one kind) of
N = nx*ny values in the reference layout (flat index nx*(j-1) + i).  With
tensors the kernels are enqueued on the assembly stream and the outputs are
new tensors on the same device; with numpy arrays the call returns numpy
arrays."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import SolverError, last_status

METRICS = ("xe", "ye", "xen", "yen", "xk", "yk", "xke", "yke", "Jp", "Je", "Jn", "alphae",
           "gamman", "betae", "betan")


def _i(v):
    return C.byref(C.c_int(int(v)))


def _arrays(ins, n_out, n):
    """(input pointers, output arrays, output pointers, keep-alive list)"""
    keep = []
    try:
        import torch
        if isinstance(ins[0], torch.Tensor):
            for t in ins:
                if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
                        and t.is_contiguous() and t.numel() == n):
                    raise TypeError("contiguous CUDA float64 tensors of %d values expected" % n)
            outs = [torch.empty(n, dtype=torch.float64, device=ins[0].device)
                    for _ in range(n_out)]
            return ([C.c_void_p(t.data_ptr()) for t in ins], outs,
                    [C.c_void_p(t.data_ptr()) for t in outs], keep)
    except ImportError:  # pragma: no cover
        pass
    arr = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1)) for a in ins]
    for a in arr:
        if a.size != n:
            raise ValueError("%d values expected" % n)
    keep.extend(arr)
    outs = [np.empty(n) for _ in range(n_out)]
    return ([C.c_void_p(a.ctypes.data) for a in arr], outs,
            [C.c_void_p(o.ctypes.data) for o in outs], keep)


def _check(where):
    st = last_status()
    if st.status:
        raise SolverError(st.status, where)


def centroids(x, y, nx, ny, opt=1):
    """get_real_centroids_xy -> (xp, yp)"""
    pin, outs, pout, keep = _arrays([x, y], 2, nx * ny)
    _lib.load().m2d_get_real_centroids_xy(_i(opt), _i(nx), _i(ny), *pin, *pout)
    _check("get_real_centroids_xy")
    return tuple(outs)


def metrics(x, y, xp, yp, nx, ny):
    """get_metrics -> {name: array} over METRICS"""
    pin, outs, pout, keep = _arrays([x, y, xp, yp], len(METRICS), nx * ny)
    _lib.load().m2d_get_metrics(_i(nx), _i(ny), *pin, *pout)
    _check("get_metrics")
    return dict(zip(METRICS, outs))


def radius(coord, y, yp, nx, ny):
    """get_radius -> (radius, re, rn, rp)"""
    pin, outs, pout, keep = _arrays([y, yp], 4, nx * ny)
    _lib.load().m2d_get_radius(_i(coord), _i(nx), _i(ny), *pin, *pout)
    _check("get_radius")
    return tuple(outs)


def grid_create(x, y, nx, ny, kcm=1, coord=0):
    """The tail of grid_create (mod_grid.f90:174-187): centroids, metrics
    and radii from the corner coordinates, in one dict."""
    xp, yp = centroids(x, y, nx, ny, kcm)
    out = {"x": x, "y": y, "xp": xp, "yp": yp}
    out.update(metrics(x, y, xp, yp, nx, ny))
    out.update(zip(("radius", "re", "rn", "rp"), radius(coord, y, yp, nx, ny)))
    return out
