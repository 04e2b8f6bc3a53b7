"""ctypes bindings for the oracle (TEST INFRASTRUCTURE ONLY).

Two libraries live in this directory:

* ``liboracle.so``   -- the C restatement of the reference hot path
                        (msi2d9.c, msi2d5.c, tdma2d9.c, coef.c);
* ``_ref/libm2dref.so`` -- the reference's own Fortran sources compiled in
                        place by ``make ref`` (only where the reference tree is
                        present; the compiled file travels with the repo).

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module.  The product path (the HIP library and
its Python host mirror) never does.

All arrays are numpy float64, diagonal-major: a coefficient array for an
nx*ny grid has shape (9, nx*ny) (or (5, nx*ny)), which is the memory image of
the reference's column-major ``a(nx*ny, 9)``.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libm2dref.so")
# the same reference sources built with -O3 (amdflang), for CPU timings only
REF_O3_SO = os.path.join(HERE, "_ref", "libm2dref_o3.so")
REF_FLAGS = {REF_SO: "amdflang -O2 -ffp-contract=off", REF_O3_SO: "amdflang -O3"}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = C.POINTER(C.c_int)
_dptr = C.POINTER(C.c_double)

_lib = None
_ref = None
_ref_o3 = None


def build():
    """Compile liboracle.so (and the reference library when the reference
    sources are present)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
    if os.path.isdir(os.environ.get("MACH2D_REFERENCE", "/root/reference")):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        L.orc_lu2d9.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_lu2d5.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_double]
        for f in (L.orc_fb2d9, L.orc_fb2d5):
            f.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, C.c_double,
                          C.c_int, _dp, _ip, _dptr, C.c_void_p, C.c_int, C.c_int]
        L.orc_tdma2d9.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_norm_l1_9d.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dptr]
        L.orc_norm_l1_9d_relative.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dptr]
        L.orc_norm_l1_b.argtypes = [C.c_int, C.c_int, _dp]
        L.orc_norm_l1_b.restype = C.c_double
        L.orc_residual9.argtypes = [_dp, C.c_int, C.c_int, _dp, _dp, _dp]
        _bind_coef(L)
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_timing():
    """The -O3 build of the reference (CPU-baseline timings), else the
    fixture build; returns (library, path)."""
    global _ref_o3
    if _ref_o3 is None and os.path.exists(REF_O3_SO):
        _ref_o3 = _bind_ref(C.CDLL(REF_O3_SO))
    if _ref_o3 is not None:
        return _ref_o3, REF_O3_SO
    return ref(), REF_SO


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(REF_SO + " (run `make -C oracle ref` where the reference exists)")
        _ref = _bind_ref(C.CDLL(REF_SO))
    return _ref


def _bind_ref(R):
    R.ref_lu2d9.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int]
    R.ref_lu2d5.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int]
    for f in (R.ref_fb2d9, R.ref_fb2d5):
        f.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, C.c_double, C.c_int, _dp]
    R.ref_tdma2d9.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
    R.ref_solver_msi9d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp]
    R.ref_norm_l1_9d.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dptr]
    R.ref_norm_l1_9d_relative.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dptr]
    R.ref_norm_l1_b.argtypes = [C.c_int, C.c_int, _dp]
    R.ref_norm_l1_b.restype = C.c_double
    _bind_ref_coef(R)
    return R


# --------------------------------------------------------------------------
# Coefficient assembly (src/equations/mach2d_coef.f90): argument lists in the
# reference's order.  name -> (scalars [(name, 'i'|'d')], input fields,
# outputs [(name, number of N-vectors)]).  The C restatement is orc_<name>,
# the compiled reference ref_<name>; both take (nx, ny, scalars..., fields...,
# outputs...).
# --------------------------------------------------------------------------
COEF = {
    "density_at_nodes": ([("Rg", "d")], ["p", "T"], [("ro", 1)]),
    "density_at_faces": ([("beta", "d")], ["ro", "Uce", "Vcn"], [("roe", 1), ("ron", 1)]),
    "u_coefficients": ([("modvis", "i"), ("dt", "d")],
                       ["rp", "re", "rn", "Jp", "Je", "Jn", "ye", "yk", "alphae", "betae",
                        "betan", "gamman", "vle", "vln", "roe", "ron", "roa", "Uce", "Vcn"],
                       [("a", 9)]),
    "v_coefficients": ([("coord", "i"), ("modvis", "i"), ("dt", "d")],
                       ["rp", "re", "rn", "Jp", "Je", "Jn", "xe", "xk", "alphae", "betae",
                        "betan", "gamman", "vle", "vln", "vlp", "roe", "ron", "roa", "Uce",
                        "Vcn"],
                       [("a", 9)]),
    "u_source": ([("modvis", "i"), ("beta", "d"), ("dt", "d")],
                 ["rp", "re", "rn", "xe", "ye", "xk", "yk", "xke", "yke", "xen", "yen", "Jp",
                  "Je", "Jn", "roe", "ron", "roa", "p", "vle", "vln", "Uce", "Vcn", "ua", "u",
                  "v"],
                 [("cup", 1), ("sup", 1), ("b", 1)]),
    "v_source": ([("coord", "i"), ("modvis", "i"), ("beta", "d"), ("dt", "d")],
                 ["rp", "re", "rn", "xe", "ye", "xk", "yk", "xke", "yke", "xen", "yen", "Jp",
                  "Je", "Jn", "roe", "ron", "roa", "p", "vle", "vln", "Uce", "Vcn", "va", "u",
                  "v"],
                 [("cvp", 1), ("svp", 1), ("b", 1)]),
    "T_coefficients_and_source": (
        [("modvis", "i"), ("beta", "d"), ("dt", "d")],
        ["rp", "re", "rn", "xe", "ye", "xk", "yk", "xke", "yke", "xen", "yen", "alphae",
         "betae", "betan", "gamman", "Jp", "Je", "Jn", "roe", "ron", "ro", "roa", "p", "pa",
         "cp", "vle", "vln", "ke", "kn", "Uce", "Vcn", "u", "v", "ue", "ve", "un", "vn", "ua",
         "va", "T", "Ta", "fbe", "fbn"],
        [("a", 9), ("b", 1)]),
    "p_coefficients": ([("dt", "d")],
                       ["rp", "re", "rn", "Jp", "Uce", "Vcn", "roe", "ron", "g", "de", "dn"],
                       [("a", 5)]),
    # SIMPLEC glue (in/out arrays are listed as outputs whose initial values
    # the caller supplies; ("name", 9) inputs are 9-diagonal arrays; an
    # output of size 0 is a scalar)
    "velocities_at_internal_faces": (
        [("dt", "d")],
        ["rp", "re", "rn", "xe", "ye", "xk", "yk", "xen", "yen", "xke", "yke", "Jp", "cup", "cvp",
         "sup", "svp", ("au", 9), ("av", 9), "roa", "p", "u", "v", "uea", "vea", "una", "vna"],
        [("ue", 1), ("ve", 1), ("un", 1), ("vn", 1), ("Uce", 1), ("Vcn", 1)]),
    "internal_simplec_coefficients": (
        [], ["re", "rn", "xe", "ye", "xk", "yk", ("au", 9), ("av", 9)],
        [("due", 1), ("dve", 1), ("dun", 1), ("dvn", 1), ("de", 1), ("dn", 1)]),
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
    "p_source": ([("dt", "d")],
                 ["rp", "re", "rn", "Jp", "roe", "ron", "ro", "roa", "Uce", "Vcn", "de", "dn",
                  "g", "pl"],
                 [("b", 1)]),
    # fictitious-row writers of src/user/internal_flow/mod_intflow.f90 (only
    # boundary rows are written; ("plin", "ny") is a vector of ny values)
    "intflow_set_bcu": ([("modvis", "i")], ["x", "xp", "xk", "yk", "u", "v"],
                        [("a", 9), ("b", 1)]),
    "intflow_set_bcv": ([("modvis", "i")], ["x", "xp", "xk", "yk", "u", "v"],
                        [("a", 9), ("b", 1)]),
    "intflow_set_bcpl": ([], ["x", "xp", "pl", ("plin", "ny")], [("a", 5), ("b", 1)]),
    "intflow_set_bcT": ([("Twall", "d")], ["x", "xp", "T", ("Tin", "ny")], [("a", 9), ("b", 1)]),
    "intflow_velocities_at_boundary_faces": (
        [("modvis", "i")], ["x", "xp", "xe", "ye", "xk", "yk", "u", "v"],
        [("ue", 1), ("ve", 1), ("un", 1), ("vn", 1), ("Uce", 1), ("Vcn", 1)]),
    # the rest of one internal-flow time step (in/out arrays as outputs;
    # ("Tbe", "ny") an output of ny values)
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


def _coef_argtypes(name):
    scal, fields, outs = COEF[name]
    t = [C.c_int, C.c_int]
    t += [C.c_int if k == "i" else C.c_double for _, k in scal]
    t += [_dp] * (len(fields) + len(outs))
    return t


def _bind_coef(L):
    for name in COEF:
        f = getattr(L, "orc_" + name, None)
        if f is not None:
            f.argtypes = _coef_argtypes(name)
            f.restype = None


def _bind_ref_coef(R):
    for name in COEF:
        f = getattr(R, "ref_" + name, None)
        if f is not None:
            f.argtypes = _coef_argtypes(name)
            f.restype = None


def _field_name(k):
    return k[0] if isinstance(k, tuple) else k


def _coef_call(fn, name, nx, ny, scalars, fields, out):
    scal, names, outs = COEF[name]
    n = nx * ny
    args = [nx, ny] + [int(scalars[k]) if t == "i" else float(scalars[k]) for k, t in scal]
    args += [_c(fields[_field_name(k)]) for k in names]
    res = {}
    for k, m in outs:
        init = None if out is None else out.get(k)
        shape = ((1, ny) if m == "ny" else (1, nx) if m == "nx" else (1, 1) if m == 0
                 else (m, n))
        arr = np.zeros(shape) if init is None else _c(init).reshape(shape).copy()
        res[k] = arr
        args.append(arr)
    fn(*args)
    return {k: (float(v[0, 0]) if m == 0 else (v[0] if v.shape[0] == 1 else v))
            for (k, m), v in zip(outs, res.values())}


def coef(name, nx, ny, scalars, fields, out=None):
    """C restatement of mach2d_coef.f90's `get_<name>`: returns the outputs
    ((k, N) for coefficient arrays, (N,) for vectors).  Only interior points
    are written; `out` gives initial output values (default zeros)."""
    return _coef_call(getattr(lib(), "orc_" + name), name, nx, ny, scalars, fields, out)


def ref_coef(name, nx, ny, scalars, fields, out=None):
    """The compiled reference routine, same convention as coef()."""
    return _coef_call(getattr(ref(), "ref_" + name), name, nx, ny, scalars, fields, out)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
# C restatement
# --------------------------------------------------------------------------
def lu2d9(a, nx, ny, alpha0=0.5):
    """src/solvers/msi2d9-v03.f90:277-432 -> (dl (5,N), du (4,N))."""
    a = _c(a)
    n = nx * ny
    dl = np.zeros((5, n))
    du = np.zeros((4, n))
    lib().orc_lu2d9(a, dl, du, n, nx, ny, alpha0)
    return dl, du


def fb2d9(a, dl, du, nx, ny, var, gamma, nitm, rhs, hist_cap=None, max_iter_cap=0):
    """src/solvers/msi2d9-v03.f90:42-262.  Returns (var, nit, r0, hist)."""
    var = _c(var).copy()
    n = nx * ny
    cap = hist_cap if hist_cap is not None else max(nitm, max_iter_cap, 1)
    hist = np.zeros(max(cap, 1))
    nit = C.c_int(0)
    r0 = C.c_double(0.0)
    lib().orc_fb2d9(_c(a), _c(dl), _c(du), ny, nx, n, var, gamma, nitm, _c(rhs),
                    C.byref(nit), C.byref(r0), hist.ctypes.data_as(C.c_void_p), cap,
                    max_iter_cap)
    k = nit.value
    return var, k, r0.value, hist[:min(k, cap)].copy()


def solve_msi9d(a, rhs, var, nx, ny, nitm, tol, max_iter_cap=0):
    """class_solver_msi9d%solve: lu2d9 then fb2d9 (mod_solver_msi.f90:119-133)."""
    dl, du = lu2d9(a, nx, ny)
    return fb2d9(a, dl, du, nx, ny, var, tol, nitm, rhs, max_iter_cap=max_iter_cap)


def lu2d5(a, nx, ny, alpha0=0.5):
    a = _c(a)
    n = nx * ny
    dl = np.zeros((4, n))
    du = np.zeros((3, n))
    lib().orc_lu2d5(a, dl, du, n, nx, ny, alpha0)
    return dl, du


def fb2d5(a, dl, du, nx, ny, var, gamma, nitm, rhs, hist_cap=None, max_iter_cap=0):
    var = _c(var).copy()
    n = nx * ny
    cap = hist_cap if hist_cap is not None else max(nitm, max_iter_cap, 1)
    hist = np.zeros(max(cap, 1))
    nit = C.c_int(0)
    r0 = C.c_double(0.0)
    lib().orc_fb2d5(_c(a), _c(dl), _c(du), ny, nx, n, var, gamma, nitm, _c(rhs),
                    C.byref(nit), C.byref(r0), hist.ctypes.data_as(C.c_void_p), cap,
                    max_iter_cap)
    k = nit.value
    return var, k, r0.value, hist[:min(k, cap)].copy()


def tdma2d9(a, b, x, nx, ny, nitm):
    x = _c(x).copy()
    lib().orc_tdma2d9(nx, ny, nitm, _c(a), _c(b), x)
    return x


def residual9(a, var, rhs, nx, ny):
    r = np.zeros(nx * ny)
    lib().orc_residual9(_c(a), nx, ny, _c(var), _c(rhs), r)
    return r


def norm_l1_9d(a, b, var, nx, ny, norm=0.0):
    v = C.c_double(norm)
    lib().orc_norm_l1_9d(nx, ny, _c(var), _c(b), _c(a), C.byref(v))
    return v.value


def norm_l1_9d_relative(a, b, var, nx, ny, norm=0.0):
    v = C.c_double(norm)
    lib().orc_norm_l1_9d_relative(nx, ny, _c(var), _c(b), _c(a), C.byref(v))
    return v.value


def norm_l1_b(b, nx, ny):
    return lib().orc_norm_l1_b(nx, ny, _c(b))


# --------------------------------------------------------------------------
# The reference itself (compiled Fortran).  fb2d9 keeps its automatic arrays
# on the stack (3*N doubles), so calls run on a thread with a large stack.
# --------------------------------------------------------------------------
def _big_stack_call(fn, stack_mb=512):
    out = {}

    def run():
        try:
            out["v"] = fn()
        except BaseException as e:  # pragma: no cover - re-raised below
            out["e"] = e

    old = threading.stack_size()
    threading.stack_size(stack_mb * 1024 * 1024)
    try:
        t = threading.Thread(target=run)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if "e" in out:
        raise out["e"]
    return out.get("v")


def ref_lu2d9(a, nx, ny, lib=None):
    a = _c(a)
    n = nx * ny
    dl = np.zeros((5, n))
    du = np.zeros((4, n))
    (lib or ref()).ref_lu2d9(a, dl, du, n, nx, ny)
    return dl, du


def ref_fb2d9(a, dl, du, nx, ny, var, gamma, nitm, rhs, lib=None):
    var = _c(var).copy()
    n = nx * ny
    R = lib or ref()
    _big_stack_call(lambda: R.ref_fb2d9(_c(a), _c(dl), _c(du), ny, nx, n, var, gamma, nitm,
                                        _c(rhs)))
    return var


def ref_lu2d5(a, nx, ny):
    a = _c(a)
    n = nx * ny
    dl = np.zeros((4, n))
    du = np.zeros((3, n))
    ref().ref_lu2d5(a, dl, du, n, nx, ny)
    return dl, du


def ref_fb2d5(a, dl, du, nx, ny, var, gamma, nitm, rhs):
    var = _c(var).copy()
    n = nx * ny
    _big_stack_call(lambda: ref().ref_fb2d5(_c(a), _c(dl), _c(du), ny, nx, n, var,
                                            gamma, nitm, _c(rhs)))
    return var


def ref_ifile_get(fname, fs, varname, kind):
    """class_ifile (src/util/class_ifile.f90) of the compiled reference: load
    `fname` and read `varname` as int (kind int), real(8) (float) or string."""
    R = ref()
    f = R.ref_ifile_get
    f.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_int),
                  C.POINTER(C.c_double), C.c_char_p, C.c_int]
    f.restype = None
    iv, rv = C.c_int(), C.c_double()
    buf = C.create_string_buffer(600)
    k = {int: 0, float: 1, str: 2}[kind]
    f(fname.encode(), fs.encode(), varname.encode(), k, C.byref(iv), C.byref(rv), buf, 500)
    if kind is int:
        return iv.value
    if kind is float:
        return rv.value
    return buf.raw[:500].decode("latin-1").rstrip(" \x00")


def ref_format_residual(it, dt, norm, normpl, rmass):
    """The residual file's row and header lines as the Fortran runtime formats
    them (src/mach2d_main.f90:495-517 edit descriptors)."""
    f = ref().ref_format_residual
    f.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_char_p,
                  C.c_char_p, C.c_int]
    f.restype = None
    row, head = C.create_string_buffer(201), C.create_string_buffer(201)
    f(it, dt, norm, normpl, rmass, row, head, 200)
    return (row.raw[:200].decode().rstrip(" \x00"), head.raw[:200].decode().rstrip(" \x00"))


METRICS = ("xe", "ye", "xen", "yen", "xk", "yk", "xke", "yke", "Jp", "Je", "Jn", "alphae",
           "gamman", "betae", "betan")


def centroids(x, y, nx, ny, opt=1):
    """get_real_centroids_xy (src/grid/mod_grid.f90:237-300) -> (xp, yp)."""
    f = lib().orc_get_real_centroids_xy
    f.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
    f.restype = None
    n = nx * ny
    xp, yp = np.zeros(n), np.zeros(n)
    f(opt, nx, ny, _c(x), _c(y), xp, yp)
    return xp, yp


def grid_metrics(x, y, xp, yp, nx, ny):
    """get_metrics (mod_grid.f90:304-558) -> {name: N-vector} (METRICS)."""
    f = lib().orc_get_metrics
    f.argtypes = [C.c_int, C.c_int] + [_dp] * 19
    f.restype = None
    n = nx * ny
    out = {k: np.zeros(n) for k in METRICS}
    f(nx, ny, _c(x), _c(y), _c(xp), _c(yp), *[out[k] for k in METRICS])
    return out


def radius(coord, y, yp, nx, ny):
    """get_radius (mod_grid.f90:562-723) -> (radius, re, rn, rp)."""
    f = lib().orc_get_radius
    f.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    f.restype = None
    n = nx * ny
    out = [np.zeros(n) for _ in range(4)]
    f(coord, nx, ny, _c(y), _c(yp), *out)
    return tuple(out)


def _ref_grid(name):
    f = getattr(ref(), "_QMmod_gridP" + name)
    f.restype = None
    return f


def ref_centroids(x, y, nx, ny, opt=1):
    """The reference's own (module-private) get_real_centroids_xy
    (mod_grid.f90:237-300), called through its module symbol."""
    f = _ref_grid("get_real_centroids_xy")
    f.argtypes = [_ip, _ip, _ip, _dp, _dp, _dp, _dp]
    n = nx * ny
    xp, yp = np.zeros(n), np.zeros(n)
    f(C.byref(C.c_int(opt)), C.byref(C.c_int(nx)), C.byref(C.c_int(ny)), _c(x), _c(y), xp, yp)
    return xp, yp


def ref_grid_metrics(x, y, xp, yp, nx, ny):
    """The reference's own get_metrics (mod_grid.f90:304-558) -> {name: N-vector}."""
    f = _ref_grid("get_metrics")
    f.argtypes = [_ip, _ip] + [_dp] * 19
    n = nx * ny
    out = {k: np.zeros(n) for k in METRICS}
    f(C.byref(C.c_int(nx)), C.byref(C.c_int(ny)), _c(x), _c(y), _c(xp), _c(yp),
      *[out[k] for k in METRICS])
    return out


def ref_radius(coord, y, yp, nx, ny):
    """The reference's own get_radius (mod_grid.f90:562-723) -> (radius, re, rn, rp)."""
    f = _ref_grid("get_radius")
    f.argtypes = [_ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp]
    n = nx * ny
    out = [np.zeros(n) for _ in range(4)]
    f(C.byref(C.c_int(coord)), C.byref(C.c_int(nx)), C.byref(C.c_int(ny)), _c(y), _c(yp), *out)
    return tuple(out)


def norm_l1_5d(nx, ny, var, b, a, norm=0.0, ref_=False):
    """norm_l1_5d (mod_solvers.f90:91-125): norm + sum |A var - b| over real
    volumes, a (5, N)."""
    L = ref() if ref_ else lib()
    f = L.ref_norm_l1_5d if ref_ else L.orc_norm_l1_5d
    f.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dptr]
    f.restype = None
    v = C.c_double(norm)
    f(nx, ny, _c(var), _c(b), _c(a), C.byref(v))
    return v.value


def get_average(nx, ny, F, ref_=False):
    """get_average (mod_field_operations.f90:115-142)"""
    L = ref() if ref_ else lib()
    f = L.ref_get_average if ref_ else L.orc_get_average
    f.argtypes = [C.c_int, C.c_int, _dp]
    f.restype = C.c_double
    return f(nx, ny, _c(F))


def ref_time_cycle_update(nx, ny, scalars, fields, out):
    """intflow_time_cycle_update (mod_intflow.f90:1106-1134) composed from the
    reference's own module procedures get_uin_vin_pin_Tin_Mw and
    get_plin_and_p_fictitious (private: called by their module symbols),
    with gamma = thermomodel%gm(T0) and Rg given."""
    R = ref()
    d = C.c_double
    f1 = getattr(R, "_QMmod_intflowPget_uin_vin_pin_tin_mw")
    f1.argtypes = [_ip, _ip, _dptr, _dptr, _dptr, _dptr, _dp, _dp, _dp, _dp, _dp, _dp]
    f1.restype = None
    f2 = getattr(R, "_QMmod_intflowPget_plin_and_p_fictitious")
    f2.argtypes = [_ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp]
    f2.restype = None
    res = {k: _c(out[k]).copy() for k in ("uin", "vin", "pin", "Tin", "Mw", "plin", "pina", "p")}
    res["pina"] = res["pin"].copy()
    inx, iny = C.c_int(nx), C.c_int(ny)
    f1(C.byref(inx), C.byref(iny), C.byref(d(scalars["gamma"])), C.byref(d(scalars["Rg"])),
       C.byref(d(scalars["P0"])), C.byref(d(scalars["T0"])), _c(fields["u"]), res["uin"],
       res["vin"], res["pin"], res["Tin"], res["Mw"])
    f2(C.byref(inx), C.byref(iny), _c(fields["x"]), _c(fields["xp"]), res["pina"], res["pin"],
       res["plin"], res["p"])
    return res


def mean_cc(a, nx, ny):
    """get_mean_cc_9d / get_mean_cc_5d (mod_tstepper_scarborough.f90:260-320)
    of a (9, N) or (5, N) coefficient array."""
    a = _c(a)
    f = lib().orc_get_mean_cc_9d if a.shape[0] == 9 else lib().orc_get_mean_cc_5d
    f.argtypes = [C.c_int, C.c_int, _dp]
    f.restype = C.c_double
    return f(nx, ny, a)


def dt0_scarborough(Jp, u, v, nx, ny):
    f = lib().orc_tstepper_dt0_scarborough
    f.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
    f.restype = C.c_double
    return f(nx, ny, _c(Jp), _c(u), _c(v))


def dt_scarborough(dt, au, av, at, ap, nx, ny, h0, mincc, maxcc):
    f = lib().orc_tstepper_dt_scarborough
    f.argtypes = [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                  C.c_double]
    f.restype = C.c_double
    return f(nx, ny, dt, _c(au), _c(av), _c(at), _c(ap), h0, mincc, maxcc)


def ref_mean_cc(a, nx, ny):
    """The reference's own (module-private) get_mean_cc_9d / get_mean_cc_5d,
    called through its symbol with Fortran by-reference arguments."""
    a = _c(a)
    sym = ("_QMmod_tstepper_scarboroughPget_mean_cc_9d" if a.shape[0] == 9
           else "_QMmod_tstepper_scarboroughPget_mean_cc_5d")
    f = getattr(ref(), sym)
    f.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int), _dp]
    f.restype = C.c_double
    return f(C.byref(C.c_int(nx)), C.byref(C.c_int(ny)), a)


def ref_dt0_scarborough(Jp, u, v, nx, ny):
    f = ref().ref_tstepper_dt0_scarborough
    f.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
    f.restype = C.c_double
    return f(nx, ny, _c(Jp), _c(u), _c(v))


def ref_dt_scarborough(paramfile, dt, au, av, at, ap, nx, ny):
    """tstepper_scarborough_init from `paramfile` (h0, mincc, maxcc), then
    tstepper_dt_scarborough."""
    f = ref().ref_tstepper_dt_scarborough
    f.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
    f.restype = C.c_double
    return f(paramfile.encode(), nx, ny, dt, _c(au), _c(av), _c(at), _c(ap))


def ref_tdma2d9(a, b, x, nx, ny, nitm):
    x = _c(x).copy()
    _big_stack_call(lambda: ref().ref_tdma2d9(nx, ny, nitm, _c(a), _c(b), x))
    return x


def ref_solver_msi9d(a, b, sol, nx, ny, nitm, tol):
    sol = _c(sol).copy()
    _big_stack_call(lambda: ref().ref_solver_msi9d(nx, ny, nitm, tol, _c(a), _c(b), sol))
    return sol


def ref_norm_l1_9d(a, b, var, nx, ny, norm=0.0):
    v = C.c_double(norm)
    ref().ref_norm_l1_9d(nx, ny, _c(var), _c(b), _c(a), C.byref(v))
    return v.value


def ref_norm_l1_9d_relative(a, b, var, nx, ny, norm=0.0):
    v = C.c_double(norm)
    ref().ref_norm_l1_9d_relative(nx, ny, _c(var), _c(b), _c(a), C.byref(v))
    return v.value


def ref_norm_l1_b(b, nx, ny):
    return ref().ref_norm_l1_b(nx, ny, _c(b))
