"""Pins of the oracle that do not depend on running the reference: dense
brute force on tiny grids, the ILU(0) identity at alpha = 0, closed forms
and the special cases that reduce to textbook routines.  Each check is
chosen so that a dropped term, a wrong neighbour index or a transposed
operand fails it."""
import numpy as np
import pytest
from scipy.linalg import solve_banded

from mach2d import problems as P

# flat-index offsets of the 9 slots (msi2d9-v03.f90, np = ni*(j-1)+i)
def offsets(nx):
    return [-nx - 1, -nx, -nx + 1, -1, 0, 1, nx - 1, nx, nx + 1]


def dense_A(a, nx, ny, two_d=True):
    """Dense matrix of a 9-diagonal array.  two_d: drop couplings whose
    neighbour lies outside the grid (the residual's convention)."""
    n = nx * ny
    A = np.zeros((n, n))
    mask = P.outside_mask(nx, ny)
    for k, off in enumerate(offsets(nx)):
        for p in range(n):
            q = p + off
            if 0 <= q < n and not (two_d and mask[k, p]):
                A[p, q] += a[k, p]
    return A


def dense_LU(dl, du, nx, ny):
    n = nx * ny
    L = np.zeros((n, n))
    U = np.eye(n)
    lo = [0, -1, -nx + 1, -nx, -nx - 1]     # dl slots: P, W, SE, S, SW
    up = [1, nx - 1, nx, nx + 1]            # du slots: E, NW, N, NE
    for k, off in enumerate(lo):
        for p in range(n):
            if 0 <= p + off < n:
                L[p, p + off] = dl[k, p]
    for k, off in enumerate(up):
        for p in range(n):
            if 0 <= p + off < n:
                U[p, p + off] = du[k, p]
    return L, U


SHAPES = [(5, 4), (6, 5), (9, 7), (12, 6)]


@pytest.mark.parametrize("nx,ny", SHAPES)
def test_alpha0_factor_is_ilu0(orc, nx, ny):
    """alpha = 0: (L U) equals A on A's 9-point flat-index pattern."""
    s = P.random_f3_system(nx, ny, seed=nx * ny)
    dl, du = orc.lu2d9(s.a, nx, ny, alpha0=0.0)
    L, U = dense_LU(dl, du, nx, ny)
    LU = L @ U
    A = dense_A(s.a, nx, ny, two_d=False)
    n = nx * ny
    for off in offsets(nx):
        for p in range(n):
            q = p + off
            if 0 <= q < n:
                assert abs(LU[p, q] - A[p, q]) <= 1e-12 * max(1.0, abs(A[p, p])), (p, off)


@pytest.mark.parametrize("nx,ny", SHAPES)
def test_alpha_changes_factor_only_on_real_volumes(orc, nx, ny):
    """alpha enters only at real volumes (msi2d9-v03.f90:347-351): on
    fictitious rows the factor is the same for alpha = 0 and 0.5."""
    s = P.random_f3_system(nx, ny, seed=3 + nx)
    dl0, du0 = orc.lu2d9(s.a, nx, ny, alpha0=0.0)
    dl5, du5 = orc.lu2d9(s.a, nx, ny, alpha0=0.5)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    fict = ((i == 0) | (i == nx - 1) | (j == 0) | (j == ny - 1)).reshape(-1)
    # the first grid row and the first point of row 2 precede any real volume
    first = np.arange(nx * ny) < nx + 1
    assert np.array_equal(dl0[:, first], dl5[:, first])
    assert np.array_equal(du0[:, first], du5[:, first])
    assert not np.array_equal(dl0[:, ~fict], dl5[:, ~fict])


@pytest.mark.parametrize("nx,ny", SHAPES)
def test_one_iteration_is_preconditioned_richardson(orc, nx, ny):
    """fb2d9 iteration: var+ = var - (LU)^-1 (A var - rhs), dense solve."""
    s = P.random_f3_system(nx, ny, seed=7 * nx + ny)
    dl, du = orc.lu2d9(s.a, nx, ny)
    L, U = dense_LU(dl, du, nx, ny)
    A = dense_A(s.a, nx, ny)
    x0 = np.linspace(-1, 1, nx * ny)
    x1, nit, r0, hist = orc.fb2d9(s.a, dl, du, nx, ny, x0, 0.0, 1, s.rhs)
    r = A @ x0 - s.rhs
    expect = x0 - np.linalg.solve(L @ U, r)
    assert nit == 1
    assert np.allclose(x1, expect, rtol=1e-12, atol=1e-12)
    assert abs(r0 - np.linalg.norm(r)) <= 1e-13 * np.linalg.norm(r)
    assert abs(hist[0] - np.linalg.norm(A @ x1 - s.rhs)) <= 1e-12 * r0


def test_fixed_point(orc):
    s = P.random_f3_system(9, 7, seed=5)
    A = dense_A(s.a, 9, 7)
    x = np.linalg.solve(A, s.rhs)
    dl, du = orc.lu2d9(s.a, 9, 7)
    x1, _, _, _ = orc.fb2d9(s.a, dl, du, 9, 7, x, 0.0, 1, s.rhs)
    assert np.max(np.abs(x1 - x)) <= 1e-13 * np.max(np.abs(x))


def test_diagonal_system_exact_in_one_iteration(orc):
    nx, ny = 8, 6
    a = np.zeros((9, nx * ny))
    a[4] = np.linspace(1.0, 3.0, nx * ny)
    rhs = np.cos(np.arange(nx * ny))
    dl, du = orc.lu2d9(a, nx, ny)
    x, nit, _, hist = orc.fb2d9(a, dl, du, nx, ny, np.zeros(nx * ny), 1e-14, 50, rhs)
    assert nit == 1 and np.allclose(x, rhs / a[4], rtol=1e-15, atol=0)


def test_row_tridiagonal_system_is_thomas(orc):
    """Only W/E couplings: the factor is the exact row-wise LU, so one
    iteration from zero solves every row (= scipy solve_banded)."""
    nx, ny = 11, 5
    rng = np.random.default_rng(9)
    a = np.zeros((9, nx * ny))
    a[3] = -rng.uniform(0.1, 1, nx * ny)
    a[5] = -rng.uniform(0.1, 1, nx * ny)
    a[4] = 2.5 + rng.uniform(0, 1, nx * ny)
    a[3].reshape(ny, nx)[:, 0] = 0.0
    a[5].reshape(ny, nx)[:, -1] = 0.0
    rhs = rng.uniform(-1, 1, nx * ny)
    dl, du = orc.lu2d9(a, nx, ny)
    x, _, _, _ = orc.fb2d9(a, dl, du, nx, ny, np.zeros(nx * ny), 0.0, 1, rhs)
    for j in range(ny):
        sl = slice(j * nx, (j + 1) * nx)
        ab = np.zeros((3, nx))
        ab[0, 1:] = a[5, sl][:-1]
        ab[1] = a[4, sl]
        ab[2, :-1] = a[3, sl][1:]
        assert np.allclose(x[sl], solve_banded((1, 1), ab, rhs[sl]), rtol=1e-13, atol=1e-14)


def test_stopping_rule_edges(orc):
    s = P.random_f3_system(10, 6, seed=2)
    dl, du = orc.lu2d9(s.a, 10, 6)
    z = np.zeros(s.N)
    # gamma = 0 never converges: exactly nitm iterations
    _, nit, _, _ = orc.fb2d9(s.a, dl, du, 10, 6, z, 0.0, 17, s.rhs)
    assert nit == 17
    # at least one iteration even when the guess is exact (r0 = 0 -> gammarp = 0)
    A = dense_A(s.a, 10, 6)
    xs = np.linalg.solve(A, s.rhs)
    _, nit, r0, _ = orc.fb2d9(s.a, dl, du, 10, 6, xs, 1e-3, 3, s.rhs)
    assert nit == 3
    # nitm <= 0: no count exit; converges on the residual like a large nitm
    _, nit_big, _, h1 = orc.fb2d9(s.a, dl, du, 10, 6, z, 1e-6, 10000, s.rhs)
    _, nit_0, _, h2 = orc.fb2d9(s.a, dl, du, 10, 6, z, 1e-6, 0, s.rhs, max_iter_cap=10000)
    assert nit_big == nit_0 < 10000
    # strict comparison: the last norm is below, the one before is not
    _, nit, r0, h = orc.fb2d9(s.a, dl, du, 10, 6, z, 1e-6, 10000, s.rhs)
    assert h[-1] < 1e-6 * r0 and (nit == 1 or h[-2] >= 1e-6 * r0)


def test_residual_is_2d_aware(orc):
    """r = A.x - rhs drops couplings across row ends even when they are
    non-zero (msi2d9-v03.f90:94-106)."""
    nx, ny = 7, 5
    rng = np.random.default_rng(4)
    a = rng.uniform(-1, 1, (9, nx * ny))
    x = rng.uniform(-1, 1, nx * ny)
    rhs = rng.uniform(-1, 1, nx * ny)
    r = orc.residual9(a, x, rhs, nx, ny)
    assert np.allclose(r, dense_A(a, nx, ny, two_d=True) @ x - rhs, rtol=1e-14, atol=1e-14)


def test_norms_closed_form(orc):
    nx, ny = 9, 6
    rng = np.random.default_rng(8)
    a = rng.uniform(-1, 1, (9, nx * ny))
    x = rng.uniform(-1, 1, nx * ny)
    b = rng.uniform(-1, 1, nx * ny)
    res = (dense_A(a, nx, ny, two_d=False) @ x - b).reshape(ny, nx)[1:-1, 1:-1]
    bb = b.reshape(ny, nx)[1:-1, 1:-1]
    assert np.isclose(orc.norm_l1_9d(a, b, x, nx, ny, 0.5), 0.5 + np.abs(res).sum(), rtol=1e-14)
    assert np.isclose(orc.norm_l1_9d_relative(a, b, x, nx, ny, 0.5),
                      0.5 + np.abs(res).sum() / np.abs(bb).sum(), rtol=1e-14)
    assert orc.norm_l1_b(b, nx, ny) == pytest.approx(np.abs(bb).sum(), rel=1e-14)
    z = np.zeros_like(b)
    assert np.isclose(orc.norm_l1_9d_relative(a, z, x, nx, ny, 0.0),
                      np.abs((dense_A(a, nx, ny, two_d=False) @ x).reshape(ny, nx)
                             [1:-1, 1:-1]).sum(), rtol=1e-14)


def test_tdma_rows_only_is_exact(orc):
    """W/E couplings only: the x-pass solves every row exactly and the y-pass
    (no S/N coupling) leaves it unchanged up to rounding."""
    nx, ny = 13, 4
    rng = np.random.default_rng(12)
    a = np.zeros((9, nx * ny))
    a[3] = -rng.uniform(0.1, 1, nx * ny)
    a[5] = -rng.uniform(0.1, 1, nx * ny)
    a[4] = 2.5
    b = rng.uniform(-1, 1, nx * ny)
    x = orc.tdma2d9(a, b, np.zeros(nx * ny), nx, ny, 1)
    for j in range(ny):
        sl = slice(j * nx, (j + 1) * nx)
        ab = np.zeros((3, nx))
        ab[0, 1:] = a[5, sl][:-1]
        ab[1] = a[4, sl]
        ab[2, :-1] = a[3, sl][1:]
        assert np.allclose(x[sl], solve_banded((1, 1), ab, b[sl]), rtol=1e-13, atol=1e-14)


def test_tdma_columns_only_is_exact(orc):
    nx, ny = 5, 14
    rng = np.random.default_rng(13)
    a = np.zeros((9, nx * ny))
    a[1] = -rng.uniform(0.1, 1, nx * ny)
    a[7] = -rng.uniform(0.1, 1, nx * ny)
    a[4] = 2.5
    b = rng.uniform(-1, 1, nx * ny)
    x = orc.tdma2d9(a, b, np.zeros(nx * ny), nx, ny, 1).reshape(ny, nx)
    A2, B2 = a.reshape(9, ny, nx), b.reshape(ny, nx)
    for i in range(nx):
        ab = np.zeros((3, ny))
        ab[0, 1:] = A2[7, :-1, i]
        ab[1] = A2[4, :, i]
        ab[2, :-1] = A2[1, 1:, i]
        assert np.allclose(x[:, i], solve_banded((1, 1), ab, B2[:, i]), rtol=1e-13, atol=1e-14)


def test_tdma_converges_to_solution(orc):
    """Block line-Jacobi on a diagonally dominant system converges to A^-1 b
    (explicit neighbour terms use the right signs and indices)."""
    nx, ny = 8, 7
    s = P.random_f3_system(nx, ny, seed=21, dominance=2.0)
    x = orc.tdma2d9(s.a, s.rhs, np.zeros(s.N), nx, ny, 60)
    assert np.allclose(x, np.linalg.solve(dense_A(s.a, nx, ny), s.rhs), rtol=1e-10, atol=1e-12)


# ---- coefficient assembly: properties the mathematics fixes ------------------
def _flow(nx=14, ny=9, coord=1, cart=False, seed=5):
    from mach2d import problems as P
    return P.flow_fields(nx, ny, seed=seed, coord=coord, cartesian=cart)


def _interior(nx, ny):
    m = np.zeros((ny, nx), bool)
    m[1:-1, 1:-1] = True
    return m.reshape(-1)


def test_face_density_pure_upwind_at_beta0(orc):
    nx, ny = 14, 9
    f = _flow(nx, ny)
    r = orc.coef("density_at_faces", nx, ny, dict(beta=0.0), f)
    ro, U, V = f["ro"], f["Uce"], f["Vcn"]
    for j in range(1, ny - 1):          # east faces: rows 2..ny-1, i = 1..nx-1
        for i in range(0, nx - 1):
            p = j * nx + i
            assert r["roe"][p] == (ro[p] if U[p] > 0 else ro[p + 1])
    for j in range(0, ny - 1):          # north faces
        for i in range(1, nx - 1):
            p = j * nx + i
            assert r["ron"][p] == (ro[p] if V[p] > 0 else ro[p + nx])


@pytest.mark.parametrize("sign", [1.0, -1.0])
@pytest.mark.parametrize("name", ["u_coefficients", "v_coefficients"])
def test_euler_coefficients_are_upwind(orc, name, sign):
    nx, ny = 14, 9
    f = _flow(nx, ny)
    f["Uce"] = sign * np.abs(f["Uce"]) + sign * 0.1
    f["Vcn"] = sign * np.abs(f["Vcn"]) + sign * 0.1
    a = orc.coef(name, nx, ny, dict(modvis=0, dt=1e-3, coord=1), f)["a"]
    m = _interior(nx, ny)
    idx = np.nonzero(m)[0]
    fmw = f["roe"][idx - 1] * f["re"][idx - 1] * f["Uce"][idx - 1]
    fms = f["ron"][idx - nx] * f["rn"][idx - nx] * f["Vcn"][idx - nx]
    fme = f["roe"][idx] * f["re"][idx] * f["Uce"][idx]
    fmn = f["ron"][idx] * f["rn"][idx] * f["Vcn"][idx]
    for k in (0, 2, 6, 8):
        assert np.all(a[k, idx] == 0.0)
    if sign > 0:   # flow towards +i, +j: the upstream (W, S) neighbours carry the flux
        assert np.all(a[5, idx] == 0.0) and np.all(a[7, idx] == 0.0)
        assert np.array_equal(a[3, idx], -fmw) and np.array_equal(a[1, idx], -fms)
    else:
        assert np.all(a[3, idx] == 0.0) and np.all(a[1, idx] == 0.0)
        assert np.array_equal(a[5, idx], fme) and np.array_equal(a[7, idx], fmn)


@pytest.mark.parametrize("name", ["u_coefficients", "v_coefficients",
                                  "T_coefficients_and_source"])
def test_orthogonal_grid_has_no_corner_couplings(orc, name):
    """betae = betan = 0 on an orthogonal grid: the cross-derivative
    (corner) couplings of the 9-point stencil vanish exactly."""
    nx, ny = 12, 8
    f = _flow(nx, ny, cart=True)
    m = _interior(nx, ny)
    # the boundary faces' one-sided metrics carry rounding: beta = 0 exactly
    f["betae"] = np.zeros_like(f["betae"])
    f["betan"] = np.zeros_like(f["betan"])
    a = orc.coef(name, nx, ny, dict(modvis=1, dt=1e-3, beta=0.2, coord=1), f)["a"]
    for k in (0, 2, 6, 8):
        assert np.all(a[k, m] == 0.0), k
    assert np.any(a[1, m] != 0.0) and np.any(a[3, m] != 0.0)


@pytest.mark.parametrize("name", ["u_coefficients", "v_coefficients"])
def test_diffusion_is_linear_in_viscosity(orc, name):
    """Without advection and transient terms every coefficient is linear in
    (vle, vln, vlp): doubling the viscosities doubles them exactly."""
    nx, ny = 14, 9
    f = _flow(nx, ny)
    f["Uce"] = np.zeros_like(f["Uce"])
    f["Vcn"] = np.zeros_like(f["Vcn"])
    f["roa"] = np.zeros_like(f["roa"])
    sc = dict(modvis=1, dt=1e-3, coord=1)
    a1 = orc.coef(name, nx, ny, sc, f)["a"]
    g = dict(f)
    for k in ("vle", "vln", "vlp"):
        g[k] = 2.0 * f[k]
    a2 = orc.coef(name, nx, ny, sc, g)["a"]
    m = _interior(nx, ny)
    assert np.array_equal(a2[:, m], 2.0 * a1[:, m])
    assert np.any(a1[:, m] != 0.0)


def test_p_source_vanishes_for_balanced_field(orc):
    nx, ny = 14, 9
    f = _flow(nx, ny)
    f["Uce"] = np.zeros_like(f["Uce"])
    f["Vcn"] = np.zeros_like(f["Vcn"])
    f["roa"] = f["ro"].copy()
    f["pl"] = np.zeros_like(f["pl"])
    b = orc.coef("p_source", nx, ny, dict(dt=1e-3), f)["b"]
    assert np.all(b == 0.0)
    f["pl"] = _flow(nx, ny, seed=9)["pl"]   # a pressure correction drives a mass source
    b = orc.coef("p_source", nx, ny, dict(dt=1e-3), f)["b"]
    assert np.any(b[_interior(nx, ny)] != 0.0)


def test_p_coefficients_balance(orc):
    """Without advection: a_P = g rp / (Jp dt) - (a_S + a_W + a_E + a_N)."""
    nx, ny = 14, 9
    f = _flow(nx, ny)
    f["Uce"] = np.zeros_like(f["Uce"])
    f["Vcn"] = np.zeros_like(f["Vcn"])
    dt = 1e-3
    a = orc.coef("p_coefficients", nx, ny, dict(dt=dt), f)["a"]
    m = _interior(nx, ny)
    lhs = a[2, m]
    rhs = f["g"][m] * f["rp"][m] / (f["Jp"][m] * dt) - (a[0, m] + a[1, m] + a[3, m] + a[4, m])
    assert np.allclose(lhs, rhs, rtol=1e-13, atol=0)
    assert np.all(a[[0, 1, 3, 4]][:, m] < 0.0)     # off-diagonals negative


def test_T_source_of_uniform_state_is_transient_only(orc):
    """Euler, fluid at rest, uniform T, no pressure change: b = mpa Ta cp / dt."""
    nx, ny = 14, 9
    f = _flow(nx, ny)
    for k in ("u", "v", "ua", "va", "ue", "ve", "un", "vn"):
        f[k] = np.zeros_like(f[k])
    f["T"] = np.full_like(f["T"], 300.0)
    f["pa"] = f["p"].copy()
    dt = 1e-3
    r = orc.coef("T_coefficients_and_source", nx, ny, dict(modvis=0, beta=0.3, dt=dt), f)
    m = _interior(nx, ny)
    mpa = f["roa"][m] * f["rp"][m] / f["Jp"][m]
    assert np.array_equal(r["b"][m], mpa * f["Ta"][m] * f["cp"][m] / dt)
