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
import weakref

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


def set_pl_graph(dev, use_graph):
    """1: replay the loop's batches from captured HIP graphs (the default);
    0: enqueue the same kernels directly (m2d_set_pl_graph)."""
    rc = _lib.load().m2d_set_pl_graph(dev.h, int(use_graph))
    if rc:
        raise SolverError(rc, "m2d_set_pl_graph")


def pl_loop_info(dev):
    """(host reads, batches) of the last intflow_pl_loop call on `dev`."""
    reads, batches = C.c_int(), C.c_int()
    rc = _lib.load().m2d_pl_loop_info(dev.h, C.byref(reads), C.byref(batches))
    if rc:
        raise SolverError(rc, "m2d_pl_loop_info")
    return reads.value, batches.value


def set_pl_solve_cap(dev, cap):
    """MSI iterations enqueued per p' solve of the loop (m2d_set_pl_solve_cap):
    0 from the previous call's longest solve (the default), > 0 that many,
    < 0 all nitm_p; a solve that needs more makes the call run again with
    all of them, so the results do not depend on it."""
    rc = _lib.load().m2d_set_pl_solve_cap(dev.h, int(cap))
    if rc:
        raise SolverError(rc, "m2d_set_pl_solve_cap")


def pl_loop_cap_info(dev):
    """(iterations enqueued per p' solve, reruns) of the last call on `dev`."""
    cap, reruns = C.c_int(), C.c_int()
    rc = _lib.load().m2d_pl_loop_cap_info(dev.h, C.byref(cap), C.byref(reruns))
    if rc:
        raise SolverError(rc, "m2d_pl_loop_cap_info")
    return cap.value, reruns.value


# ---------------------------------------------------------------------------
# One internal-flow time step (src/mach2d_main.f90:94-404), m2d_intflow_time_step
# ---------------------------------------------------------------------------
TS_CONST = ("x", "xp", "xe", "ye", "xk", "yk", "xke", "yke", "xen", "yen", "rp", "re", "rn",
            "Jp", "Je", "Jn", "alphae", "betae", "betan", "gamman",
            "vle", "vln", "vlp", "cp", "ke", "kn", "fbe", "fbn")
TS_STATE = ("u", "v", "ue", "ve", "un", "vn", "Uce", "Vcn", "p", "T", "ro", "roe", "ron", "pl")
TS_LEVEL = ("ua", "va", "uea", "vea", "una", "vna", "pa", "Ta", "roa")
TS_WORK = ("au", "av", "at", "ap", "bu", "bv", "bt", "bp", "cup", "sup", "cvp", "svp", "due",
           "dve", "dun", "dvn", "de", "dn", "g")
TS_INFLOW = ("uin", "vin", "pin", "Tin", "Mw", "plin", "pina", "Tbe", "Tbw", "Tbn", "Tbs")
TS_ALL = TS_CONST + TS_STATE + TS_LEVEL + TS_WORK + TS_INFLOW
TS_OPTIONAL = ("cp", "ke", "kn", "fbe", "fbn")   # used by the T equation only


class TimeStepFields(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in TS_ALL]


class TimeStepParams(C.Structure):
    _fields_ = ([(k, C.c_int) for k in ("modvis", "coord", "solver9d", "itmmax", "itpmax",
                                          "nitm_u", "nitm_p")] +
                [(k, C.c_double) for k in ("tol_u", "tol_p", "tolm", "beta", "Rg", "gamma", "P0",
                                             "T0", "Cpref", "Href", "Twall", "h0", "mincc",
                                             "maxcc")])


class TimeStepResult(C.Structure):
    _fields_ = [("norm", C.c_double), ("rmass", C.c_double), ("normpl", C.c_double),
                ("itp", C.c_int), ("stage", C.c_int)]

STAGES = {1: "time levels", 2: "inflow update", 3: "p average and g", 4: "u, v assembly",
          5: "u solve", 6: "v solve", 7: "face velocities", 8: "p' loop",
          9: "node velocities", 10: "mass check", 11: "temperature", 12: "densities",
          13: "next dt"}


def ts_size(k, nx, ny):
    N = nx * ny
    if k in ("uin", "vin", "pin", "Tin", "Mw", "plin", "pina", "Tbe", "Tbw"):
        return ny
    if k in ("Tbn", "Tbs"):
        return nx
    return 9 * N if k in ("au", "av", "at") else 5 * N if k == "ap" else N


# the last checked field set per grid size: weak references to its tensors,
# their data pointers and the filled structure (a step's Python overhead was
# mostly these checks)
_TS_LAST = {}


def _ts_fields(f, nx, ny):
    import torch
    ts = [f.get(k) for k in TS_ALL]
    last = _TS_LAST.get((nx, ny))
    if last is not None:
        refs, ptrs, flds = last
        if all((r is None) if t is None else (r is not None and r() is t)
               for r, t in zip(refs, ts)) and \
                ptrs == tuple(t.data_ptr() for t in ts if t is not None):
            return flds
    flds = TimeStepFields()
    for k, t in zip(TS_ALL, ts):
        if t is None and k in TS_OPTIONAL:
            continue
        n = ts_size(k, nx, ny)
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
                and t.is_contiguous() and t.numel() == n):
            raise TypeError("%s must be a contiguous float64 CUDA tensor of %d values" % (k, n))
        setattr(flds, k, t.data_ptr())
    _TS_LAST[(nx, ny)] = (tuple(None if t is None else weakref.ref(t) for t in ts),
                          tuple(t.data_ptr() for t in ts if t is not None), flds)
    return flds


def intflow_time_step(dev9, dev5, f, params, dt):
    """One time step on device tensors `f` (names of TS_ALL; TS_OPTIONAL may
    be missing for Euler) with the FP64 handles dev9 (9 diagonals; None for
    solver9d = 1) and dev5 (5 diagonals).  Returns (next dt, result dict)."""
    flds = _ts_fields(f, dev5.nx, dev5.ny)
    prm = TimeStepParams(**params)
    d = C.c_double(dt)
    res = TimeStepResult()
    rc = _lib.load().m2d_intflow_time_step(dev9.h if dev9 is not None else None, dev5.h,
                                           C.byref(flds), C.byref(prm), C.byref(d),
                                           C.byref(res))
    if rc:
        raise SolverError(rc, "m2d_intflow_time_step (%s)" % STAGES.get(res.stage, "?"))
    return d.value, {"norm": res.norm, "rmass": res.rmass, "normpl": res.normpl,
                     "itp": res.itp}
