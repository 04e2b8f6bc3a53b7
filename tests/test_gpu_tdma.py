"""GPU parity of the line TDMA (tdma2d9) and the L1 norms through the C ABI.

TDMA: bit-exact against the reference's golden vectors and the oracle (the
Thomas recurrences run in the reference's order, one lane per line; P and
its denominators once per call, the right-hand sides D in parallel), on
short and long lines (lines longer than 64 KB take fewer lanes per LDS
tile), and for batches of systems in the same launches.  Norms: within
1e-13 relative (the GPU sums in a different order)."""
import numpy as np
import pytest

from conftest import load_golden
from mach2d import problems as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [1, 2, 5])
def test_tdma2d9_bitwise_vs_fixtures(gpu, k):
    g = load_golden("tdma.npz")
    for c in sorted({n.rsplit("_", 1)[0] for n in g if n.endswith("_shape")}):
        nx, ny = (int(v) for v in g[c + "_shape"])
        x = g[c + "_x0"].copy()
        st = gpu.tdma2d9(nx, ny, k, g[c + "_a"], g[c + "_b"], x)
        assert st.status == 0
        assert np.array_equal(x, g[c + "_x%d" % k]), c


@pytest.mark.parametrize("nx,ny", [(258, 130), (130, 66), (17, 300), (1026, 9), (3, 3),
                                   (1100, 37), (45, 2000), (1026, 514)])
def test_tdma2d9_bitwise_vs_oracle(gpu, orc, nx, ny):
    rng = np.random.default_rng(nx * ny)
    a = -rng.uniform(0.05, 1.0, size=(9, nx * ny))
    a[4] = 1.5 * np.sum(np.abs(a[[0, 1, 2, 3, 5, 6, 7, 8]]), axis=0) + 0.1
    b = rng.uniform(-1, 1, nx * ny)
    x0 = rng.uniform(-1, 1, nx * ny)
    x = x0.copy()
    gpu.tdma2d9(nx, ny, 3, a, b, x)
    assert np.array_equal(x, orc.tdma2d9(a, b, x0, nx, ny, 3))


@pytest.mark.parametrize("nx,ny", [(258, 130), (130, 66), (17, 300), (3, 3), (400, 37),
                                   (1026, 9)])
def test_tdma_pcr_mode_within_1e12_of_oracle(gpu, orc, nx, ny):
    """m2d_set_tdma_mode(1): the lines solved by parallel cyclic reduction.
    Same line-Jacobi iteration, other rounding: 3 iterations within 1e-12
    (max norm, relative) of the reference's recurrence; grids past 400
    points a line keep the recurrence (bit for bit)."""
    rng = np.random.default_rng(nx * ny + 7)
    a = -rng.uniform(0.05, 1.0, size=(9, nx * ny))
    a[4] = 1.5 * np.sum(np.abs(a[[0, 1, 2, 3, 5, 6, 7, 8]]), axis=0) + 0.1
    b = rng.uniform(-1, 1, nx * ny)
    x0 = rng.uniform(-1, 1, nx * ny)
    want = orc.tdma2d9(a, b, x0, nx, ny, 3)
    x = x0.copy()
    gpu.set_tdma_mode(1)
    try:
        st = gpu.tdma2d9(nx, ny, 3, a, b, x)
    finally:
        gpu.set_tdma_mode(0)
    assert st.status == 0
    err = np.max(np.abs(x - want)) / np.max(np.abs(want))
    if max(nx, ny) > 400:
        assert np.array_equal(x, want)
    else:
        assert err <= 1e-12, err
    # the default mode is the recurrence again
    x = x0.copy()
    gpu.tdma2d9(nx, ny, 3, a, b, x)
    assert np.array_equal(x, want)


def test_tdma_pcr_mode_rejects_other_modes(gpu):
    with pytest.raises(gpu.SolverError):
        gpu.set_tdma_mode(2)


def test_tdma_device_many_equals_single(gpu, orc):
    """Three systems of one grid size in the same launches: each result is
    its single-system result and the oracle's, bit for bit; the x-pass
    result of one system never leaks into another."""
    import torch
    nx, ny = 258, 130
    systems = [P.random_f3_system(nx, ny, seed=30 + q) for q in range(3)]
    x0 = [np.random.default_rng(q).uniform(-1, 1, nx * ny) for q in range(3)]
    ta = [torch.tensor(s.a.reshape(-1), device="cuda") for s in systems]
    tb = [torch.tensor(s.rhs, device="cuda") for s in systems]
    tx = [torch.tensor(v, device="cuda") for v in x0]
    gpu.tdma2d9_device_many(nx, ny, 5, ta, tb, tx)
    for q, s in enumerate(systems):
        want = orc.tdma2d9(s.a, s.rhs, x0[q], nx, ny, 5)
        assert np.array_equal(tx[q].cpu().numpy(), want), q
        one = torch.tensor(x0[q], device="cuda")
        gpu.ClassSolverTDMA9d().init(nx, ny, 5).solve(ta[q], tb[q], one)
        assert torch.equal(one, tx[q]), q
    # nitm = 0 leaves x untouched (the reference's loop runs zero times)
    keep = [t.clone() for t in tx]
    gpu.tdma2d9_device_many(nx, ny, 0, ta, tb, tx)
    assert all(torch.equal(u, v) for u, v in zip(keep, tx))


def test_tdma_pcr_device_many_equals_single(gpu, orc):
    """PCR mode: systems solved in the same launches (more than one group,
    so the scratch is reused between groups) give each its solo result bit
    for bit, each within 1e-12 of the oracle's recurrence; the scratch of
    one system never leaks into another."""
    import torch
    nx, ny, count = 258, 130, 10  # two groups of at most 8 systems
    systems = [P.random_f3_system(nx, ny, seed=50 + q) for q in range(count)]
    x0 = [np.random.default_rng(q + 9).uniform(-1, 1, nx * ny) for q in range(count)]
    ta = [torch.tensor(s.a.reshape(-1), device="cuda") for s in systems]
    tb = [torch.tensor(s.rhs, device="cuda") for s in systems]
    tx = [torch.tensor(v, device="cuda") for v in x0]
    gpu.set_tdma_mode(1)
    try:
        gpu.tdma2d9_device_many(nx, ny, 5, ta, tb, tx)
        for q, s in enumerate(systems):
            want = orc.tdma2d9(s.a, s.rhs, x0[q], nx, ny, 5)
            got = tx[q].cpu().numpy()
            assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), q
            one = torch.tensor(x0[q], device="cuda")
            gpu.tdma2d9_device_many(nx, ny, 5, [ta[q]], [tb[q]], [one])
            assert torch.equal(one, tx[q]), q
    finally:
        gpu.set_tdma_mode(0)


def test_tdma_solver_objects(gpu, orc):
    s = P.pcorr_system(66, 34, seed=2)
    x = np.zeros(s.N)
    gpu.ClassSolverTDMA9d().init(66, 34, 4).solve(s.a, s.rhs, x)
    assert np.array_equal(x, orc.tdma2d9(s.a, s.rhs, np.zeros(s.N), 66, 34, 4))
    a5 = s.a[[1, 3, 4, 5, 7]].copy()
    x5 = np.zeros(s.N)
    gpu.ClassSolverTDMA5d().init(66, 34, 4).solve(a5, s.rhs, x5)
    a9 = np.zeros_like(s.a)
    a9[[1, 3, 4, 5, 7]] = a5
    assert np.array_equal(x5, orc.tdma2d9(a9, s.rhs, np.zeros(s.N), 66, 34, 4))


def test_tdma_device_entry(gpu, orc):
    import torch
    s = P.pcorr_system(130, 66, seed=8)
    ta = torch.tensor(s.a, device="cuda")
    tb = torch.tensor(s.rhs, device="cuda")
    tx = torch.zeros(s.N, dtype=torch.float64, device="cuda")
    gpu.ClassSolverTDMA9d().init(130, 66, 2).solve(ta, tb, tx)
    assert np.array_equal(tx.cpu().numpy(), orc.tdma2d9(s.a, s.rhs, np.zeros(s.N), 130, 66, 2))


def test_norms_vs_fixtures(gpu):
    g = load_golden("norms.npz")
    for c in sorted({n.rsplit("_", 1)[0] for n in g if n.endswith("_shape")}):
        nx, ny = (int(v) for v in g[c + "_shape"])
        a, b, v = g[c + "_a"], g[c + "_b"], g[c + "_var"]
        assert gpu.norm_l1_9d(nx, ny, v, b, a, 0.25) == pytest.approx(float(g[c + "_l1"]),
                                                                      rel=1e-13)
        assert gpu.norm_l1_9d_relative(nx, ny, v, b, a, 0.25) == pytest.approx(
            float(g[c + "_rel"]), rel=1e-13)
        assert gpu.norm_l1_9d_relative(nx, ny, v, np.zeros_like(b), a, 0.25) == pytest.approx(
            float(g[c + "_rel0"]), rel=1e-13)
        assert gpu.norm_l1_b(nx, ny, b) == pytest.approx(float(g[c + "_nb"]), rel=1e-13)
