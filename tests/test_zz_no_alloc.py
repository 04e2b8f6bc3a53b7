"""m2d_solve_msi9d_noalloc (the solve that allocates no device memory)
against m2d_solve_msi9d, the entry point it is a variant of, through the C
ABI: same iterate bit for bit (FP64), same iteration count and residual
norms, on two calls in a row with the second on a smaller grid (the
library's array keeps the first call's contents), plus the size limit."""
import numpy as np
import pytest

from mach2d import problems as P

pytestmark = pytest.mark.gpu


def _host_solve(gpu, s, nitm, tol):
    sol = np.zeros(s.N)
    st = gpu.ClassSolverMSI9d().init(s.nx, s.ny, nitm, tol).solve(s.a, s.rhs, sol)
    return sol, st


def test_noalloc_matches_solve_msi9d_two_calls(gpu):
    import torch
    cases = [(258, 130, 11, 400, 1e-10), (130, 66, 12, 500, 1e-10)]
    for nx, ny, seed, nitm, tol in cases:
        s = P.pcorr_system(nx, ny, seed=seed, rhs="gauss")
        print("noalloc: %dx%d seed %d nitm %d tol %g" % (nx, ny, seed, nitm, tol), flush=True)
        ref, rst = _host_solve(gpu, s, nitm, tol)
        assert rst.status == 0
        ta = torch.tensor(s.a.reshape(-1), device="cuda")
        tb = torch.tensor(s.rhs, device="cuda")
        tv = torch.zeros(s.N, dtype=torch.float64, device="cuda")
        st = gpu.solve_msi9d_noalloc(nx, ny, nitm, tol, ta, tb, tv)
        assert st.status == 0, st.as_dict()
        assert st.nit == rst.nit and st.r0 == rst.r0 and st.rn == rst.rn
        assert np.array_equal(tv.cpu().numpy(), ref)


def test_noalloc_rejects_grids_beyond_its_array(gpu):
    import torch
    nx, ny = 1026, 514
    n = nx * ny
    ta = torch.zeros(9 * n, dtype=torch.float64, device="cuda")
    tb = torch.zeros(n, dtype=torch.float64, device="cuda")
    tv = torch.zeros(n, dtype=torch.float64, device="cuda")
    st = gpu.solve_msi9d_noalloc(nx, ny, 5, 1e-10, ta, tb, tv)
    assert st.status == 4     # M2D_ERR_ARG, before any launch
    # the entry point still works after the rejection
    s = P.pcorr_system(66, 34, seed=13, rhs="gauss")
    ref, _ = _host_solve(gpu, s, 300, 1e-10)
    ta = torch.tensor(s.a.reshape(-1), device="cuda")
    tb = torch.tensor(s.rhs, device="cuda")
    tv = torch.zeros(s.N, dtype=torch.float64, device="cuda")
    st = gpu.solve_msi9d_noalloc(66, 34, 300, 1e-10, ta, tb, tv)
    assert st.status == 0 and np.array_equal(tv.cpu().numpy(), ref)
