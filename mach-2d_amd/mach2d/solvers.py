"""Host-side mirror of the reference's linear-solver plugin interface,
backed by the gfx950 C ABI (include/mach2d_hip.h).

Reference interface (src/solvers/):
  * class_solver_abstract%solve(a, b, sol)       mod_solver_abstract.f90:12-31
  * class_solver_msi9d%init(nx, ny, nitm, tol)   mod_solver_msi.f90:96-115
    class_solver_msi9d%solve(a, b, sol)          mod_solver_msi.f90:119-133
  * class_solver_msi5d%init / solve              mod_solver_msi.f90:56-93
  * class_solver_tdma9d%init(nx, ny, nitm)       mod_solver_tdma.f90:76-100
  * class_solver_tdma5d%init / solve             mod_solver_tdma.f90:48-72
  * solvers_init(nx, ny, ifile) -> (solver5d, solver9d)   mod_solvers.f90:17-87
  * norm_l1_9d / norm_l1_9d_relative / norm_l1_b          mod_solvers.f90:128-276
  * lu2d9, fb2d9 (msi2d9-v03.f90:277, :42), tdma2d9 (tdma2d9.f90:38)
  * lu2d5, fb2d5 (msi2d5-v03.f90:247, :42)

Arrays: a coefficient array for an nx*ny grid is either a C-contiguous
(9, nx*ny) float64 array or the Fortran-ordered (nx*ny, 9) array of the
reference (the same memory).  `sol`/`var`/`x` are updated in place, like the
reference's intent(inout) dummies.  numpy arrays take the synchronous
host-array entry points; torch tensors on the GPU take the device-resident
handle API without host copies.

5-diagonal systems (slots S, W, P, E, N): MSI has its own entry points
(m2d_lu2d5 / m2d_fb2d5, and handles created with 5 diagonals), whose
residual and sweeps read the 5 + 4 + 3 arrays of msi2d5; the factorization
is that of the zero-corner embedding into the 9-diagonal slots 2, 4, 5, 6, 8,
whose terms reduce to lu2d5's.  TDMA embeds the 5-point system the same way
(tdma2d9 on it performs tdma2d5's operations plus exact zero terms).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import SolverError, Status, last_status

# 5-point slots (S, W, P, E, N) -> 9-point slots, 0-based
EMBED_5_IN_9 = (1, 3, 4, 5, 7)


def _coef(a, n, k):
    """Return a C-contiguous (k, n) float64 view/copy of a coefficient array."""
    a = np.asarray(a)
    if a.shape == (k, n):
        return np.ascontiguousarray(a, dtype=np.float64)
    if a.shape == (n, k):           # Fortran a(n, k): the transpose is the memory image
        return np.ascontiguousarray(a.T, dtype=np.float64)
    raise ValueError("coefficient array must have shape (%d, %d) or (%d, %d), got %s"
                     % (k, n, n, k, a.shape))


def _vec(v, n, name):
    if not isinstance(v, np.ndarray) or v.dtype != np.float64 or not v.flags.c_contiguous:
        raise TypeError("%s must be a C-contiguous float64 numpy array" % name)
    if v.size != n:
        raise ValueError("%s has %d entries, expected %d" % (name, v.size, n))
    return v


def _p(arr):
    return arr.ctypes.data_as(C.c_void_p)


def _i(v):
    return C.byref(C.c_int(int(v)))


def _d(v):
    return C.byref(C.c_double(float(v)))


def embed5(a5, n):
    """(5, n) or (n, 5) 5-point coefficients -> (9, n) with zero corners."""
    a5 = _coef(a5, n, 5)
    a9 = np.zeros((9, n))
    a9[list(EMBED_5_IN_9)] = a5
    return a9


def _is_torch(x):
    return type(x).__module__.startswith("torch")


# ---------------------------------------------------------------------------
# reference-signature functions
# ---------------------------------------------------------------------------
def lu2d9(b, dl, du, nij, ni, nj):
    """msi2d9-v03.f90:277 lu2d9(b, dl, du, nij, ni, nj); dl (5,nij), du (4,nij) out."""
    bb = _coef(b, nij, 9)
    _vec(dl, 5 * nij, "dl")
    _vec(du, 4 * nij, "du")
    _lib.load().m2d_lu2d9(_p(bb), _p(dl), _p(du), _i(nij), _i(ni), _i(nj))
    return last_status()


def fb2d9(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs):
    """msi2d9-v03.f90:42 fb2d9(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs); var inout."""
    bb = _coef(b, nij, 9)
    dl = _coef(dl, nij, 5)
    du = _coef(du, nij, 4)
    _vec(var, nij, "var")
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    _lib.load().m2d_fb2d9(_p(bb), _p(dl), _p(du), _i(nj), _i(ni), _i(nij), _p(var),
                          _d(gamma), _i(nitm), _p(rhs))
    return last_status()


def lu2d5(b, dl, du, nij, ni, nj):
    """msi2d5-v03.f90:247 lu2d5(b, dl, du, nij, ni, nj); b (5,nij), dl (4,nij), du (3,nij) out."""
    bb = _coef(b, nij, 5)
    _vec(dl, 4 * nij, "dl")
    _vec(du, 3 * nij, "du")
    _lib.load().m2d_lu2d5(_p(bb), _p(dl), _p(du), _i(nij), _i(ni), _i(nj))
    return last_status()


def fb2d5(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs):
    """msi2d5-v03.f90:42 fb2d5(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs); var inout."""
    bb = _coef(b, nij, 5)
    dl = _coef(dl, nij, 4)
    du = _coef(du, nij, 3)
    _vec(var, nij, "var")
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    _lib.load().m2d_fb2d5(_p(bb), _p(dl), _p(du), _i(nj), _i(ni), _i(nij), _p(var),
                          _d(gamma), _i(nitm), _p(rhs))
    return last_status()


def tdma2d9(nx, ny, nitm, a, b, x):
    """tdma2d9.f90:38 tdma2d9(nx, ny, nitm, a, b, x); x inout."""
    n = nx * ny
    aa = _coef(a, n, 9)
    b = np.ascontiguousarray(b, dtype=np.float64)
    _vec(x, n, "x")
    _lib.load().m2d_tdma2d9(_i(nx), _i(ny), _i(nitm), _p(aa), _p(b), _p(x))
    return last_status()


def norm_l1_9d(nx, ny, var, b, a, norm=0.0):
    """mod_solvers.f90:128; returns norm + sum over real volumes |A.var - b|."""
    v = C.c_double(norm)
    n = nx * ny
    aa = _coef(a, n, 9)
    _lib.load().m2d_norm_l1_9d(_i(nx), _i(ny), _p(np.ascontiguousarray(var, np.float64)),
                               _p(np.ascontiguousarray(b, np.float64)), _p(aa), C.byref(v))
    _check_last("norm_l1_9d")
    return v.value


def norm_l1_5d(nx, ny, var, b, a, norm=0.0):
    """mod_solvers.f90:91; returns norm + sum over real volumes |A.var - b|
    (a: 5 diagonals S, W, P, E, N)."""
    v = C.c_double(norm)
    n = nx * ny
    aa = _coef(a, n, 5)
    _lib.load().m2d_norm_l1_5d(_i(nx), _i(ny), _p(np.ascontiguousarray(var, np.float64)),
                               _p(np.ascontiguousarray(b, np.float64)), _p(aa), C.byref(v))
    _check_last("norm_l1_5d")
    return v.value


def norm_l1_9d_relative(nx, ny, var, b, a, norm=0.0):
    """mod_solvers.f90:176; relative to ||b||_1 when ||b||_1 > 10 eps."""
    v = C.c_double(norm)
    n = nx * ny
    aa = _coef(a, n, 9)
    _lib.load().m2d_norm_l1_9d_relative(_i(nx), _i(ny),
                                        _p(np.ascontiguousarray(var, np.float64)),
                                        _p(np.ascontiguousarray(b, np.float64)), _p(aa),
                                        C.byref(v))
    _check_last("norm_l1_9d_relative")
    return v.value


def norm_l1_b(nx, ny, b):
    """mod_solvers.f90:250."""
    r = _lib.load().m2d_norm_l1_b(_i(nx), _i(ny), _p(np.ascontiguousarray(b, np.float64)))
    _check_last("norm_l1_b")
    return r


def _check_last(where):
    st = last_status()
    if st.status not in (_lib.M2D_OK,):
        raise SolverError(st.status, where)
    return st


# ---------------------------------------------------------------------------
# plugin classes
# ---------------------------------------------------------------------------
class ClassSolverAbstract:
    """mod_solver_abstract.f90:12-31: solve(a, b, sol) with sol inout."""

    def solve(self, a, b, sol):  # pragma: no cover - interface
        raise NotImplementedError


class ClassSolverMSI9d(ClassSolverAbstract):
    """MSI for 9-diagonal systems (mod_solver_msi.f90:35-50,96-133).  Every solve
    factors then iterates, as the reference (lu2d9 + fb2d9)."""

    def init(self, nx, ny, nitm, tol):
        self.nx, self.ny, self.nxy, self.nitm, self.tol = nx, ny, nx * ny, nitm, tol
        self._dev = None
        return self

    def solve(self, a, b, sol):
        if _is_torch(sol):
            return self._solve_device(a, b, sol)
        aa = _coef(a, self.nxy, 9)
        _vec(sol, self.nxy, "sol")
        b = np.ascontiguousarray(b, dtype=np.float64)
        _lib.load().m2d_solve_msi9d(_i(self.nx), _i(self.ny), _i(self.nitm), _d(self.tol),
                                    _p(aa), _p(b), _p(sol))
        self.status = last_status()
        return self.status

    def _solve_device(self, a, b, sol):
        if self._dev is None:
            self._dev = DeviceMSI(self.nx, self.ny)
        self._dev.factor(a)
        self.status = self._dev.solve(a, b, sol, self.tol, self.nitm)
        return self.status


def solve_msi9d_noalloc(nx, ny, nitm, tol, a, b, sol, stream=None):
    """class_solver_msi9d%solve on CUDA tensors (a: 9*nx*ny, b and sol:
    nx*ny, float64) through m2d_solve_msi9d_noalloc, which allocates no
    device memory (its buffers live in an array the library holds per GPU,
    for grids up to about 550x270).  Returns the Status of the call."""
    import torch
    n = nx * ny
    for t, m, name in ((a, 9 * n, "a"), (b, n, "b"), (sol, n, "sol")):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
                and t.is_contiguous() and t.numel() == m):
            raise TypeError("%s must be a contiguous float64 CUDA tensor of %d values" % (name, m))
    rc = _lib.load().m2d_solve_msi9d_noalloc(_i(nx), _i(ny), _i(nitm), _d(tol),
                                             C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()),
                                             C.c_void_p(sol.data_ptr()),
                                             C.c_void_p(stream) if stream else None)
    st = last_status()
    st.status = rc
    return st


class ClassSolverMSI5d(ClassSolverAbstract):
    """MSI for 5-diagonal systems (mod_solver_msi.f90:17-31,56-93): every
    solve factors then iterates (lu2d5 + fb2d5)."""

    def init(self, nx, ny, nitm, tol):
        self.nx, self.ny, self.nxy, self.nitm, self.tol = nx, ny, nx * ny, nitm, tol
        self._dev = None
        return self

    def solve(self, a, b, sol):
        if _is_torch(sol):
            if self._dev is None:
                self._dev = DeviceMSI(self.nx, self.ny, ndiag=5)
            self._dev.factor(a)
            self.status = self._dev.solve(a, b, sol, self.tol, self.nitm)
            return self.status
        aa = _coef(a, self.nxy, 5)
        _vec(sol, self.nxy, "sol")
        b = np.ascontiguousarray(b, dtype=np.float64)
        _lib.load().m2d_solve_msi5d(_i(self.nx), _i(self.ny), _i(self.nitm), _d(self.tol),
                                    _p(aa), _p(b), _p(sol))
        self.status = last_status()
        return self.status


class ClassSolverTDMA9d(ClassSolverAbstract):
    """mod_solver_tdma.f90:31-44,76-100: nitm alternating line sweeps."""

    def init(self, nx, ny, nitm):
        self.nx, self.ny, self.nitm = nx, ny, nitm
        return self

    def solve(self, a, b, sol):
        if _is_torch(sol):
            rc = _lib.load().m2d_tdma2d9_device(self.nx, self.ny, self.nitm, a.data_ptr(),
                                                b.data_ptr(), sol.data_ptr(), None)
            if rc:
                raise SolverError(rc, "tdma2d9")
            return rc
        self.status = tdma2d9(self.nx, self.ny, self.nitm, a, b, sol)
        return self.status


def tdma2d9_device_many(nx, ny, nitm, a_list, b_list, x_list, stream=None):
    """tdma2d9 on several systems of one grid size (CUDA float64 tensors:
    a 9*nx*ny, b and x nx*ny each) in the same launches
    (m2d_tdma2d9_device_many); x updated in place."""
    import torch
    n = nx * ny
    if not (len(a_list) == len(b_list) == len(x_list) >= 1):
        raise ValueError("a, b and x lists must have the same nonzero length")
    for t, m, name in [(t, 9 * n, "a") for t in a_list] + [(t, n, "b") for t in b_list] + \
            [(t, n, "x") for t in x_list]:
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
                and t.is_contiguous() and t.numel() == m):
            raise TypeError("%s must be a contiguous float64 CUDA tensor of %d values" % (name, m))
    cnt = len(a_list)
    arr = lambda ts: (C.c_void_p * cnt)(*[t.data_ptr() for t in ts])
    rc = _lib.load().m2d_tdma2d9_device_many(cnt, nx, ny, nitm, arr(a_list), arr(b_list),
                                             arr(x_list), C.c_void_p(stream) if stream else None)
    if rc:
        raise SolverError(rc, "m2d_tdma2d9_device_many")
    return rc


def set_tdma_mode(mode):
    """Line solver of the device TDMA for this thread's later calls
    (m2d_set_tdma_mode): 0 the reference's Thomas recurrence (bit for bit,
    the default), 1 parallel cyclic reduction (grids up to 400 x 400; the
    same iteration up to rounding)."""
    rc = _lib.load().m2d_set_tdma_mode(int(mode))
    if rc:
        raise SolverError(rc, "m2d_set_tdma_mode")


class ClassSolverTDMA5d(ClassSolverAbstract):
    """mod_solver_tdma.f90:17-29,48-72 through the zero-corner embedding."""

    def init(self, nx, ny, nitm):
        self.inner = ClassSolverTDMA9d().init(nx, ny, nitm)
        self.nxy = nx * ny
        return self

    def solve(self, a, b, sol):
        return self.inner.solve(embed5(a, self.nxy), b, sol)


def solvers_init(nx, ny, ifile):
    """mod_solvers.f90:17-87.  `ifile` maps the keys solver, nitm_u, nitm_p,
    tol_u, tol_p (a dict, or anything with get_value(name)).  Returns
    (solver5d, solver9d).  An unknown solver name stops, as the reference
    does (here: raises SystemExit)."""
    if hasattr(ifile, "get_value"):   # formats.ClassIfile (or anything with get_value)
        try:
            ksolver = str(ifile.get_value("solver", str)).strip()
            nitm_u, nitm_p = ifile.get_value("nitm_u", int), ifile.get_value("nitm_p", int)
            tol_u, tol_p = ifile.get_value("tol_u", float), ifile.get_value("tol_p", float)
        except TypeError:   # get_value(name) without a kind
            get = ifile.get_value
            ksolver = str(get("solver")).strip()
            nitm_u, nitm_p = int(get("nitm_u")), int(get("nitm_p"))
            tol_u, tol_p = float(get("tol_u")), float(get("tol_p"))
    else:
        get = ifile.__getitem__
        ksolver = str(get("solver")).strip()
        nitm_u, nitm_p = int(get("nitm_u")), int(get("nitm_p"))
        tol_u, tol_p = float(get("tol_u")), float(get("tol_p"))
    if ksolver == "MSI":
        return (ClassSolverMSI5d().init(nx, ny, nitm_p, tol_p),
                ClassSolverMSI9d().init(nx, ny, nitm_u, tol_u))
    if ksolver == "TDMA":
        return (ClassSolverTDMA5d().init(nx, ny, nitm_p),
                ClassSolverTDMA9d().init(nx, ny, nitm_u))
    raise SystemExit("solvers_init:\nUnkown option: %s\nStopping..." % ksolver)


# ---------------------------------------------------------------------------
# device-resident handle (torch tensors on the GPU)
# ---------------------------------------------------------------------------
class DeviceMSI:
    """One solver handle for an nx*ny grid on the current GPU.

    precision: "fp64" (parity mode) or "fp32_ir" (FP32 factor and sweeps
    inside FP64 iterative refinement).  ndiag: 9 (msi2d9) or 5 (msi2d5: A
    holds the slots S, W, P, E, N; the factor is dl 4*N, du 3*N)."""

    def __init__(self, nx, ny, precision="fp64", stream=None, ndiag=9):
        L = _lib.load()
        self.nx, self.ny, self.N = nx, ny, nx * ny
        self.precision = precision
        self.nd = int(ndiag)
        prec = {"fp64": _lib.M2D_FP64, "fp32_ir": _lib.M2D_FP32_IR}[precision]
        h = C.c_void_p()
        rc = L.m2d_create_stencil(nx, ny, prec, self.nd, C.byref(h))
        if rc:
            raise SolverError(rc, "m2d_create")
        self.h = h
        if stream is not None:
            L.m2d_set_stream(self.h, C.c_void_p(stream))

    def close(self):
        if getattr(self, "h", None):
            _lib.load().m2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, t, numel, name):
        import torch
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
                and t.is_contiguous() and t.numel() == numel):
            raise TypeError("%s must be a contiguous float64 CUDA tensor of %d values"
                            % (name, numel))
        return t.data_ptr()

    def set_tuning(self, publish_every=32, xcd_stride=1):
        rc = _lib.load().m2d_set_tuning(self.h, publish_every, xcd_stride)
        if rc:
            raise SolverError(rc, "m2d_set_tuning")

    # "sweeps": both sweeps of an iteration (one launch), or the forward
    # sweep alone when the sweeps are staged (then "backward" times the other)
    KERNELS = {"factor": 0, "sweeps": 1, "backward": 2, "residual": 3,
               "factor32": 4, "sweeps32": 5, "backward32": 6, "residual32": 7}

    def enable_kernel_timing(self, on=True):
        _lib.load().m2d_enable_kernel_timing(self.h, 1 if on else 0)

    def kernel_timing(self):
        """{kernel: (total_ms, launches)} of the launches timed so far."""
        out = {}
        for name, kid in self.KERNELS.items():
            ms, n = C.c_double(), C.c_long()
            _lib.load().m2d_kernel_timing(self.h, kid, C.byref(ms), C.byref(n))
            if n.value:
                out[name] = (ms.value, n.value)
        return out

    def set_staging(self, staged=True):
        rc = _lib.load().m2d_set_staging(self.h, 1 if staged else 0)
        if rc:
            raise SolverError(rc, "m2d_set_staging")

    def set_inner(self, inner_gamma=1e-4, inner_nitm=25):
        rc = _lib.load().m2d_set_inner(self.h, inner_gamma, inner_nitm)
        if rc:
            raise SolverError(rc, "m2d_set_inner")

    def factor(self, a):
        rc = _lib.load().m2d_factor(self.h, C.c_void_p(self._chk(a, self.nd * self.N, "a")))
        if rc and rc != _lib.M2D_ERR_NONFINITE:
            raise SolverError(rc, "m2d_factor")
        return rc

    def get_factor(self, dl, du):
        nl, nu = (5, 4) if self.nd == 9 else (4, 3)
        rc = _lib.load().m2d_get_factor(self.h, C.c_void_p(self._chk(dl, nl * self.N, "dl")),
                                        C.c_void_p(self._chk(du, nu * self.N, "du")))
        if rc:
            raise SolverError(rc, "m2d_get_factor")

    def residual_norm(self, a, rhs, var) -> float:
        """||A var - rhs||_2 in FP64."""
        out = C.c_double()
        rc = _lib.load().m2d_residual_norm(self.h, C.c_void_p(self._chk(a, self.nd * self.N, "a")),
                                           C.c_void_p(self._chk(rhs, self.N, "rhs")),
                                           C.c_void_p(self._chk(var, self.N, "var")),
                                           C.byref(out))
        if rc:
            raise SolverError(rc, "m2d_residual_norm")
        return out.value

    def guess_scale(self, a, rhs, g, energy=False) -> float:
        """Scale g in place by the residual-minimising alpha (energy: the
        energy-minimising one, (g.rhs)/(g.A g)); returns alpha."""
        alpha = C.c_double()
        fn = _lib.load().m2d_guess_scale_energy if energy else _lib.load().m2d_guess_scale
        rc = fn(self.h, C.c_void_p(self._chk(a, self.nd * self.N, "a")),
                C.c_void_p(self._chk(rhs, self.N, "rhs")), C.c_void_p(self._chk(g, self.N, "g")),
                C.byref(alpha))
        if rc:
            raise SolverError(rc, "m2d_guess_scale")
        return alpha.value

    def guess_prepare_many(self, a_list, rhs_list, g_list, energy=True):
        """guess_scale for several systems of this grid with two host reads in
        all; returns (alphas, ||rhs - A g|| of each scaled guess)."""
        n = len(g_list)
        arr = lambda xs, k, name: (C.c_void_p * n)(
            *[C.c_void_p(self._chk(x, k, name)) for x in xs])
        pa = arr(a_list, self.nd * self.N, "a")
        pr = arr(rhs_list, self.N, "rhs")
        pg = arr(g_list, self.N, "g")
        alphas, rn = (C.c_double * n)(), (C.c_double * n)()
        rc = _lib.load().m2d_guess_prepare_many(self.h, n, pa, pr, pg, 1 if energy else 0,
                                                alphas, rn)
        if rc:
            raise SolverError(rc, "m2d_guess_prepare_many")
        return list(alphas), list(rn)

    def solve(self, a, rhs, var, gamma, nitm, hist=None) -> Status:
        st = Status()
        hp, cap = (None, 0)
        if hist is not None:
            hp, cap = C.c_void_p(self._chk(hist, hist.numel(), "hist")), hist.numel()
        _lib.load().m2d_solve(self.h, C.c_void_p(self._chk(a, self.nd * self.N, "a")),
                              C.c_void_p(self._chk(rhs, self.N, "rhs")),
                              C.c_void_p(self._chk(var, self.N, "var")), float(gamma),
                              int(nitm), hp, cap, C.byref(st))
        return st


class BatchMSI:
    """Many independent 9-diagonal systems of uneven size solved together
    (m2d_batch_*): one launch per phase for the whole batch, iterations
    replayed from a captured hipGraph, placement chosen by an autotuner.
    Each system follows fb2d9's semantics (as DeviceMSI.solve)."""

    LAYOUTS = {0: "in order", 1: "one XCD per system", 2: "largest first",
               3: "in order, LDS-staged", 4: "one XCD per system, LDS-staged",
               5: "largest first, LDS-staged"}

    def __init__(self, sizes, precision="fp64", stream=None):
        L = _lib.load()
        self.sizes = [(int(nx), int(ny)) for nx, ny in sizes]
        self.n = len(self.sizes)
        nx = (C.c_int * self.n)(*[s[0] for s in self.sizes])
        ny = (C.c_int * self.n)(*[s[1] for s in self.sizes])
        prec = {"fp64": _lib.M2D_FP64, "fp32_ir": _lib.M2D_FP32_IR}[precision]
        self.precision = precision
        h = C.c_void_p()
        rc = L.m2d_batch_create(self.n, nx, ny, prec, C.byref(h))
        if rc:
            raise SolverError(rc, "m2d_batch_create")
        self.h = h
        if stream is not None:  # hipStream_t as an integer (torch: stream.cuda_stream)
            L.m2d_batch_set_stream(self.h, C.c_void_p(stream))

    def close(self):
        if getattr(self, "h", None):
            _lib.load().m2d_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ptrs(self, tensors, per_point, name):
        import torch
        if len(tensors) != self.n:
            raise ValueError("%s: %d tensors for %d systems" % (name, len(tensors), self.n))
        arr = (C.c_void_p * self.n)()
        for q, (t, (nx, ny)) in enumerate(zip(tensors, self.sizes)):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
                    and t.is_contiguous() and t.numel() == per_point * nx * ny):
                raise TypeError("%s[%d] must be a contiguous float64 CUDA tensor of %d values"
                                % (name, q, per_point * nx * ny))
            arr[q] = t.data_ptr()
        return arr

    def factor(self, a):
        rc = _lib.load().m2d_batch_factor(self.h, self._ptrs(a, 9, "a"))
        if rc and rc != _lib.M2D_ERR_NONFINITE:
            raise SolverError(rc, "m2d_batch_factor")
        return rc

    def solve(self, a, rhs, var, gamma, nitm):
        """Returns (first non-zero status, [Status per system]); gamma is one
        value for every system or a sequence of one per system."""
        st = (Status * self.n)()
        if np.ndim(gamma):
            g = (C.c_double * self.n)(*[float(x) for x in gamma])
            rc = _lib.load().m2d_batch_solve_each(self.h, self._ptrs(a, 9, "a"),
                                                  self._ptrs(rhs, 1, "rhs"),
                                                  self._ptrs(var, 1, "var"), g, int(nitm), st)
        else:
            rc = _lib.load().m2d_batch_solve(self.h, self._ptrs(a, 9, "a"),
                                             self._ptrs(rhs, 1, "rhs"), self._ptrs(var, 1, "var"),
                                             float(gamma), int(nitm), st)
        return rc, list(st)

    def set_inner(self, inner_gamma=1e-4, inner_nitm=50):
        rc = _lib.load().m2d_batch_set_inner(self.h, float(inner_gamma), int(inner_nitm))
        if rc:
            raise SolverError(rc, "m2d_batch_set_inner")

    def set_tuning(self, layout=0, chunk_iters=8):
        rc = _lib.load().m2d_batch_set_tuning(self.h, int(layout), int(chunk_iters))
        if rc:
            raise SolverError(rc, "m2d_batch_set_tuning")

    def get_tuning(self):
        lay, ch = C.c_int(), C.c_int()
        _lib.load().m2d_batch_get_tuning(self.h, C.byref(lay), C.byref(ch))
        return lay.value, ch.value

    def tune(self, a, rhs, iters=4):
        """Autotune the placement; returns (layout, [ms per iteration per layout])."""
        lay = C.c_int()
        ms = (C.c_double * 6)()
        rc = _lib.load().m2d_batch_tune(self.h, self._ptrs(a, 9, "a"), self._ptrs(rhs, 1, "rhs"),
                                        int(iters), C.byref(lay), ms)
        if rc:
            raise SolverError(rc, "m2d_batch_tune")
        return lay.value, list(ms)
