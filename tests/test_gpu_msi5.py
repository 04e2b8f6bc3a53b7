"""GPU parity of the 5-diagonal MSI path (lu2d5, fb2d5, class_solver_msi5d
and 5-diagonal handles) against the reference's own outputs and the oracle,
through the C ABI.

Bar, as for the 9-diagonal path: FP64 factor and iterates equal to the
oracle (and to the reference-written fixture) bit for bit, iteration counts
equal, per-iteration ||r||_2 within 1e-10 relative; FP32 inside FP64
refinement: solution within 1e-8 relative of the FP64 oracle solution."""
import numpy as np
import pytest

from conftest import load_golden
from mach2d import problems as P

pytestmark = pytest.mark.gpu

SHAPES = [(3, 3), (3, 70), (200, 3), (64, 65), (130, 66), (97, 141), (258, 130)]


def system5(nx, ny, seed):
    if min(nx, ny) < 4:   # grids without interior: random F3-valid 5-point rows
        s = P.random_f3_system(nx, ny, seed=seed)
        return s.a[[1, 3, 4, 5, 7]].copy(), s.rhs
    s = P.pcorr_system5(nx, ny, seed=seed)
    return s.a, s.rhs


def test_lu2d5_fb2d5_vs_reference_fixture(gpu):
    """embed5.npz holds the reference lu2d5 factor and fb2d5 iterate (7
    iterations from zero) of two 5-point systems."""
    g = load_golden("embed5.npz")
    for c in sorted({n.rsplit("_", 1)[0] for n in g if n.endswith("_shape")}):
        nx, ny = (int(v) for v in g[c + "_shape"])
        n = nx * ny
        dl, du = np.zeros((4, n)), np.zeros((3, n))
        st = gpu.lu2d5(g[c + "_a5"], dl, du, n, nx, ny)
        assert st.status == 0, c
        assert np.array_equal(dl, g[c + "_dl5"]) and np.array_equal(du, g[c + "_du5"]), c
        var = np.zeros(n)
        st = gpu.fb2d5(g[c + "_a5"], g[c + "_dl5"], g[c + "_du5"], ny, nx, n, var, 0.0, 7,
                       g[c + "_rhs"])
        assert st.status == 0 and st.nit == 7, c
        assert np.array_equal(var, g[c + "_x5"]), c


@pytest.mark.parametrize("nx,ny", SHAPES)
def test_lu2d5_bitwise_vs_oracle(gpu, orc, nx, ny):
    a, _ = system5(nx, ny, nx + 2 * ny)
    n = nx * ny
    dl, du = np.zeros((4, n)), np.zeros((3, n))
    st = gpu.lu2d5(a, dl, du, n, nx, ny)
    assert st.status == 0
    odl, odu = orc.lu2d5(a, nx, ny)
    assert np.array_equal(dl, odl) and np.array_equal(du, odu)


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("k", [1, 4])
def test_fb2d5_iterates_bitwise_vs_oracle(gpu, orc, nx, ny, k):
    a, rhs = system5(nx, ny, 3 * nx + ny)
    dl, du = orc.lu2d5(a, nx, ny)
    var = np.zeros(nx * ny)
    st = gpu.fb2d5(a, dl, du, ny, nx, nx * ny, var, 0.0, k, rhs)
    assert st.status == 0 and st.nit == k
    ovar, onit, _, _ = orc.fb2d5(a, dl, du, nx, ny, np.zeros(nx * ny), 0.0, k, rhs)
    assert onit == k and np.array_equal(var, ovar)


def test_solver_msi5d_full_solve_vs_oracle(gpu, orc):
    """class_solver_msi5d: 130x66 5-point p' system to a 1e-10 drop: same
    iteration count and iterate as the oracle."""
    nx, ny = 130, 66
    s = P.pcorr_system5(nx, ny, seed=11)
    sol = np.zeros(s.N)
    st = gpu.ClassSolverMSI5d().init(nx, ny, 2000, 1e-10).solve(s.a, s.rhs, sol)
    assert st.status == 0
    dl, du = orc.lu2d5(s.a, nx, ny)
    ovar, onit, _, _ = orc.fb2d5(s.a, dl, du, nx, ny, np.zeros(s.N), 1e-10, 2000, s.rhs)
    assert st.nit == onit
    assert np.array_equal(sol, ovar)


@pytest.mark.parametrize("staged", [False, True])
def test_handle_5d_matches_oracle_and_embedding(gpu, orc, staged):
    """A 5-diagonal handle: FP64 iterates bit-identical to the oracle's fb2d5
    and to a 9-diagonal handle on the zero-corner embedding; residual history
    within 1e-10 relative; register and LDS-staged sweeps alike."""
    import torch
    nx, ny = 101, 197          # 4 bands, the last one partial
    s = P.pcorr_system5(nx, ny, seed=5)
    dev5 = gpu.DeviceMSI(nx, ny, ndiag=5)
    dev9 = gpu.DeviceMSI(nx, ny)
    dev5.set_staging(staged)
    dev9.set_staging(staged)
    a5 = torch.tensor(s.a, device="cuda")
    a9 = torch.tensor(gpu.embed5(s.a, s.N), device="cuda")
    rhs = torch.tensor(s.rhs, device="cuda")
    v5 = torch.zeros(s.N, dtype=torch.float64, device="cuda")
    v9 = torch.zeros_like(v5)
    hist = torch.zeros(40, dtype=torch.float64, device="cuda")
    dev5.factor(a5)
    dev9.factor(a9)
    st5 = dev5.solve(a5, rhs, v5, 0.0, 40, hist)
    st9 = dev9.solve(a9, rhs, v9, 0.0, 40)
    assert st5.status == 0 and st5.nit == 40 and st9.nit == 40
    dl, du = orc.lu2d5(s.a, nx, ny)
    ovar, _, _, ohist = orc.fb2d5(s.a, dl, du, nx, ny, np.zeros(s.N), 0.0, 40, s.rhs)
    x5 = v5.cpu().numpy()
    assert np.array_equal(x5, ovar)
    assert np.array_equal(x5, v9.cpu().numpy())
    np.testing.assert_allclose(hist.cpu().numpy(), ohist, rtol=1e-10)
    gdl = torch.zeros(4 * s.N, dtype=torch.float64, device="cuda")
    gdu = torch.zeros(3 * s.N, dtype=torch.float64, device="cuda")
    dev5.get_factor(gdl, gdu)
    assert np.array_equal(gdl.cpu().numpy().reshape(4, -1), dl)
    assert np.array_equal(gdu.cpu().numpy().reshape(3, -1), du)
    r = dev5.residual_norm(a5, rhs, v5)
    ref = np.linalg.norm(orc.residual9(gpu.embed5(s.a, s.N), x5, s.rhs, nx, ny))
    assert abs(r - ref) <= 1e-12 * ref
    dev5.close()
    dev9.close()


def test_handle_5d_fp32_refinement(gpu, orc):
    """FP32 5-diagonal sweeps inside FP64 refinement: 1e-10 drop of the FP64
    residual, solution within 1e-8 of the FP64 oracle solution."""
    import torch
    nx, ny = 258, 130
    s = P.pcorr_system5(nx, ny, seed=8, rhs="gauss")
    dev = gpu.DeviceMSI(nx, ny, "fp32_ir", ndiag=5)
    a = torch.tensor(s.a, device="cuda")
    rhs = torch.tensor(s.rhs, device="cuda")
    v = torch.zeros(s.N, dtype=torch.float64, device="cuda")
    dev.factor(a)
    st = dev.solve(a, rhs, v, 1e-10, 20000)
    assert st.status == 0 and st.rn < 1e-10 * st.r0
    dl, du = orc.lu2d5(s.a, nx, ny)
    ox, _, _, _ = orc.fb2d5(s.a, dl, du, nx, ny, np.zeros(s.N), 1e-10, 20000, s.rhs)
    x = v.cpu().numpy()
    assert np.linalg.norm(x - ox) <= 1e-8 * np.linalg.norm(ox)
    dev.close()


def test_outside_coupling_rejected_5d(gpu):
    """A nonzero coupling across the grid edge (south slot of row 0) is
    refused like in the 9-diagonal path."""
    import torch
    from mach2d._lib import M2D_ERR_OUTSIDE_COUPLING
    nx, ny = 20, 12
    s = P.pcorr_system5(nx, ny, seed=1)
    a = s.a.copy()
    a[0, 5] = -0.5
    dev = gpu.DeviceMSI(nx, ny, ndiag=5)
    with pytest.raises(gpu.SolverError) as e:
        dev.factor(torch.tensor(a, device="cuda"))
    assert e.value.code == M2D_ERR_OUTSIDE_COUPLING
    dev.close()
