"""Device-resident pressure-correction loop of one SIMPLEC iteration, internal
flow (src/mach2d_main.f90:257-309), through m2d_intflow_pl_loop.

    dev = DeviceMSI(nx, ny, "fp64", ndiag=5)
    itp, normpl = intflow_pl_loop(dev, fields, modvis=0, dt=..., beta=...,
                                  p_avg=..., tolm=..., itpmax=20, nitm_p=20,
                                  tol_p=1e-2)

`fields` maps the names of m2d_pl_loop_fields (include/mach2d_hip.h) to
contiguous float64 CUDA tensors: N-vectors, plin of ny values, ap of 5 N.
The loop updates ue, ve, un, vn, Uce, Vcn, roe, ron, ro, p, pl, ap and bp in
place, as the reference updates its module arrays.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import SolverError

IN = ("x", "xp", "xe", "ye", "xk", "yk", "rp", "re", "rn", "Jp",
      "due", "dve", "dun", "dvn", "de", "dn", "g", "roa", "u", "v", "plin")
INOUT = ("ue", "ve", "un", "vn", "Uce", "Vcn", "roe", "ron", "ro", "p", "pl")
WORK = ("ap", "bp")


class PlLoopFields(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in IN + INOUT + WORK]


def _fields(f, nx, ny):
    import torch
    N = nx * ny
    out = PlLoopFields()
    for k in IN + INOUT + WORK:
        t = f[k]
        n = ny if k == "plin" else (5 * N if k == "ap" else N)
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
                and t.is_contiguous() and t.numel() == n):
            raise TypeError("%s must be a contiguous float64 CUDA tensor of %d values" % (k, n))
        setattr(out, k, t.data_ptr())
    return out


def intflow_pl_loop(dev, f, modvis, dt, beta, p_avg, tolm, itpmax=20, nitm_p=20, tol_p=1e-2):
    """Run the loop on the 5-diagonal FP64 handle `dev`; returns (itp, normpl)."""
    flds = _fields(f, dev.nx, dev.ny)
    itp, normpl = C.c_int(), C.c_double()
    rc = _lib.load().m2d_intflow_pl_loop(dev.h, C.c_void_p(C.addressof(flds)), int(modvis),
                                         float(dt),
                                         float(beta), float(p_avg), float(tolm), int(itpmax),
                                         int(nitm_p), float(tol_p), C.byref(itp),
                                         C.byref(normpl))
    if rc:
        raise SolverError(rc, "m2d_intflow_pl_loop")
    return itp.value, normpl.value
