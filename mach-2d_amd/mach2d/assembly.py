"""Coefficient assembly on the GPU: the routines of module `coefficients`
(src/equations/mach2d_coef.f90) through the C ABI (m2d_get_*).

ROUTINES lists each routine's arguments in the reference's order: scalars
(name, 'i' | 'd'), input fields, outputs (name, number of N-vectors).
call() takes the fields as numpy arrays (copied to the GPU and back) or as
float64 CUDA tensors (processed in place on the device)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import SolverError

ROUTINES = {
    # mach2d_coef.f90:6
    "density_at_nodes": ([("Rg", "d")], ["p", "T"], [("ro", 1)]),
    # :20
    "density_at_faces": ([("beta", "d")], ["ro", "Uce", "Vcn"], [("roe", 1), ("ron", 1)]),
    # :63
    "u_coefficients": ([("modvis", "i"), ("dt", "d")],
                       ["rp", "re", "rn", "Jp", "Je", "Jn", "ye", "yk", "alphae", "betae",
                        "betan", "gamman", "vle", "vln", "roe", "ron", "roa", "Uce", "Vcn"],
                       [("a", 9)]),
    # :454
    "v_coefficients": ([("coord", "i"), ("modvis", "i"), ("dt", "d")],
                       ["rp", "re", "rn", "Jp", "Je", "Jn", "xe", "xk", "alphae", "betae",
                        "betan", "gamman", "vle", "vln", "vlp", "roe", "ron", "roa", "Uce",
                        "Vcn"],
                       [("a", 9)]),
    # :221
    "u_source": ([("modvis", "i"), ("beta", "d"), ("dt", "d")],
                 ["rp", "re", "rn", "xe", "ye", "xk", "yk", "xke", "yke", "xen", "yen", "Jp",
                  "Je", "Jn", "roe", "ron", "roa", "p", "vle", "vln", "Uce", "Vcn", "ua", "u",
                  "v"],
                 [("cup", 1), ("sup", 1), ("b", 1)]),
    # :616
    "v_source": ([("coord", "i"), ("modvis", "i"), ("beta", "d"), ("dt", "d")],
                 ["rp", "re", "rn", "xe", "ye", "xk", "yk", "xke", "yke", "xen", "yen", "Jp",
                  "Je", "Jn", "roe", "ron", "roa", "p", "vle", "vln", "Uce", "Vcn", "va", "u",
                  "v"],
                 [("cvp", 1), ("svp", 1), ("b", 1)]),
    # :787
    "T_coefficients_and_source": (
        [("modvis", "i"), ("beta", "d"), ("dt", "d")],
        ["rp", "re", "rn", "xe", "ye", "xk", "yk", "xke", "yke", "xen", "yen", "alphae",
         "betae", "betan", "gamman", "Jp", "Je", "Jn", "roe", "ron", "ro", "roa", "p", "pa",
         "cp", "vle", "vln", "ke", "kn", "Uce", "Vcn", "u", "v", "ue", "ve", "un", "vn", "ua",
         "va", "T", "Ta", "fbe", "fbn"],
        [("a", 9), ("b", 1)]),
    # :1207
    "p_coefficients": ([("dt", "d")],
                       ["rp", "re", "rn", "Jp", "Uce", "Vcn", "roe", "ron", "g", "de", "dn"],
                       [("a", 5)]),
    # :1281
    "p_source": ([("dt", "d")],
                 ["rp", "re", "rn", "Jp", "roe", "ron", "ro", "roa", "Uce", "Vcn", "de", "dn",
                  "g", "pl"],
                 [("b", 1)]),
    # SIMPLEC glue: in/out arrays are outputs whose values the caller
    # supplies; ("au", 9) inputs are 9-diagonal arrays; size 0 = scalar output
    # :1350
    "velocities_at_internal_faces": (
        [("dt", "d")],
        ["rp", "re", "rn", "xe", "ye", "xk", "yk", "xen", "yen", "xke", "yke", "Jp", "cup", "cvp",
         "sup", "svp", ("au", 9), ("av", 9), "roa", "p", "u", "v", "uea", "vea", "una", "vna"],
        [("ue", 1), ("ve", 1), ("un", 1), ("vn", 1), ("Uce", 1), ("Vcn", 1)]),
    # :1504
    "internal_simplec_coefficients": (
        [], ["re", "rn", "xe", "ye", "xk", "yk", ("au", 9), ("av", 9)],
        [("due", 1), ("dve", 1), ("dun", 1), ("dvn", 1), ("de", 1), ("dn", 1)]),
    # :1602, :1614, :1626, :1679, :1764, :1833
    "pressure_correction_with_pl": ([], ["pl"], [("p", 1)]),
    "density_correction_with_pl": ([], ["pl", "g"], [("ro", 1)]),
    "u_v_at_real_nodes_with_pl": ([], ["xe", "ye", "xk", "yk", "rp", "pl", ("au", 9), ("av", 9)],
                                  [("u", 1), ("v", 1)]),
    "velocity_correction_at_faces_with_pl": (
        [], ["xe", "ye", "xk", "yk", "due", "dve", "dun", "dvn", "de", "dn", "pl"],
        [("ue", 1), ("ve", 1), ("un", 1), ("vn", 1), ("Uce", 1), ("Vcn", 1)]),
    "boundary_faces_flags": ([], [], [("fbe", 1), ("fbn", 1)]),
    "mass_conservation": ([("dt", "d")],
                          ["rp", "re", "rn", "Jp", "ro", "roa", "ron", "roe", "Uce", "Vcn"],
                          [("res", 0)]),
    # fictitious-row writers, src/user/internal_flow/mod_intflow.f90:214, :305,
    # :396, :464; ("plin", "ny") is iflow%plin (ny values), Twall iflow%Twall
    "intflow_set_bcu": ([("modvis", "i")], ["x", "xp", "xk", "yk", "u", "v"],
                        [("a", 9), ("b", 1)]),
    "intflow_set_bcv": ([("modvis", "i")], ["x", "xp", "xk", "yk", "u", "v"],
                        [("a", 9), ("b", 1)]),
    "intflow_set_bcpl": ([], ["x", "xp", "pl", ("plin", "ny")], [("a", 5), ("b", 1)]),
    "intflow_set_bcT": ([("Twall", "d")], ["x", "xp", "T", ("Tin", "ny")], [("a", 9), ("b", 1)]),
    # mod_intflow.f90:692 (outputs are in/out: only the boundary faces change)
    "intflow_velocities_at_boundary_faces": (
        [("modvis", "i")], ["x", "xp", "xe", "ye", "xk", "yk", "u", "v"],
        [("ue", 1), ("ve", 1), ("un", 1), ("vn", 1), ("Uce", 1), ("Vcn", 1)]),
    # mod_intflow.f90:791, :881, :557, :1106 (in/out arrays as outputs whose
    # initial values `out` supplies; ("Tbe", "ny") an output of ny values)
    "intflow_boundary_simplec": ([("modvis", "i")], [],
                                 [("due", 1), ("dve", 1), ("dun", 1), ("dvn", 1), ("de", 1),
                                  ("dn", 1)]),
    "intflow_extrapolate_u_v_to_fictitious": ([("modvis", "i")], ["x", "xp", "xk", "yk"],
                                              [("u", 1), ("v", 1)]),
    "intflow_boundary_T_from_H_conservation": (
        [("Cpref", "d"), ("Href", "d")], [("Tin", "ny"), "u", "ue", "un", "v", "ve", "vn"],
        [("T", 1), ("Tbe", "ny"), ("Tbw", "ny"), ("Tbn", "nx"), ("Tbs", "nx")]),
    "intflow_time_cycle_update": (
        [("gamma", "d"), ("Rg", "d"), ("P0", "d"), ("T0", "d")], ["x", "xp", "u"],
        [("uin", "ny"), ("vin", "ny"), ("pin", "ny"), ("Tin", "ny"), ("Mw", "ny"),
         ("plin", "ny"), ("pina", "ny"), ("p", 1)]),
}

# C entry of each routine (the reference's subroutine name, prefixed)
ENTRY = {name: "m2d_get_" + name for name in ROUTINES}
ENTRY["boundary_faces_flags"] = "m2d_initialize_boundary_faces_flags"
ENTRY["mass_conservation"] = "m2d_check_mass_conservation"
for _k in ("intflow_set_bcu", "intflow_set_bcv", "intflow_set_bcpl", "intflow_set_bcT",
           "intflow_velocities_at_boundary_faces", "intflow_boundary_simplec",
           "intflow_extrapolate_u_v_to_fictitious", "intflow_boundary_T_from_H_conservation",
           "intflow_time_cycle_update"):
    ENTRY[_k] = "m2d_" + _k


def _ptr(x, n, name):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            if not (x.dtype == torch.float64 and x.is_contiguous() and x.numel() == n):
                raise TypeError("%s: contiguous float64 tensor of %d values expected" % (name, n))
            return C.c_void_p(x.data_ptr()), x
    except ImportError:  # pragma: no cover
        pass
    if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous
            and x.size == n):
        raise TypeError("%s: contiguous float64 array of %d values expected" % (name, n))
    return C.c_void_p(x.ctypes.data), x


def call(name, nx, ny, scalars, fields, out=None):
    """Run mach2d_coef.f90's get_<name>.  Outputs: `out` supplies arrays to
    write into (host arrays, or CUDA tensors when the fields are CUDA
    tensors); missing ones are allocated as zeros like the fields.  Returns
    {name: array} ((k, N) coefficient arrays, (N,) vectors)."""
    scal, names, outs = ROUTINES[name]
    n = nx * ny
    L = _lib.load()
    fn = getattr(L, ENTRY[name])
    keep = []
    args = [C.byref(C.c_int(nx)), C.byref(C.c_int(ny))]
    for k, t in scal:
        v = C.c_int(int(scalars[k])) if t == "i" else C.c_double(float(scalars[k]))
        keep.append(v)
        args.append(C.byref(v))
    first = None
    for k in names:
        key, m = k if isinstance(k, tuple) else (k, 1)
        p, obj = _ptr(fields[key].reshape(-1), ny if m == "ny" else m * n, key)
        keep.append(obj)
        args.append(p)
        first = obj if first is None else first
    res = {}
    for k, m in outs:
        size = ny if m == "ny" else nx if m == "nx" else max(m, 1) * n if m else 1
        arr = None if out is None else out.get(k)
        if arr is None:
            if first is None or isinstance(first, np.ndarray):
                arr = np.zeros(size)
            else:
                import torch
                arr = torch.zeros(size, dtype=torch.float64, device=first.device)
        p, obj = _ptr(arr, size, k)
        keep.append(obj)
        args.append(p)
        res[k] = arr
    fn(*args)
    st = _lib.last_status()
    if st.status:
        raise SolverError(st.status, ENTRY[name])
    ret = {}
    for (k, m), v in zip(outs, res.values()):
        if m == 0:
            ret[k] = float(v[0]) if isinstance(v, np.ndarray) else float(v.cpu()[0])
        else:
            ret[k] = v.reshape(m, n) if isinstance(m, int) and m > 1 else v
    return ret


def get_average(nx, ny, F):
    """get_average (src/util/mod_field_operations.f90:115-142), bit for bit"""
    p, obj = _ptr(F.reshape(-1), nx * ny, "F")
    v = _lib.load().m2d_get_average(C.byref(C.c_int(nx)), C.byref(C.c_int(ny)), p)
    st = _lib.last_status()
    if st.status:
        raise SolverError(st.status, "m2d_get_average")
    return v


def set_stream(stream=None):
    """Stream (hipStream_t as int) for device-array calls; None = default."""
    _lib.load().m2d_assembly_set_stream(C.c_void_p(stream) if stream else None)
