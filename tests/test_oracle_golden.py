"""The oracle (C restatement) against golden vectors written by the compiled
reference (tests/golden/make_golden.py).  Bit-exact: the oracle keeps the
reference's operation order and is built without FMA contraction."""
import numpy as np
import pytest

from conftest import load_golden


def _cases(g):
    return [str(c) for c in g["cases"]]


def test_lu2d9_matches_reference(orc):
    g = load_golden("msi_small.npz")
    for c in _cases(g):
        nx, ny = (int(v) for v in g[c + "_shape"])
        dl, du = orc.lu2d9(g[c + "_a"], nx, ny)
        assert np.array_equal(dl, g[c + "_dl"]), c
        assert np.array_equal(du, g[c + "_du"]), c


@pytest.mark.parametrize("k", [1, 2, 3, 5, 10])
def test_fb2d9_iterates_match_reference(orc, k):
    g = load_golden("msi_small.npz")
    for c in _cases(g):
        nx, ny = (int(v) for v in g[c + "_shape"])
        var, nit, _, _ = orc.fb2d9(g[c + "_a"], g[c + "_dl"], g[c + "_du"], nx, ny,
                                   np.zeros(nx * ny), 0.0, k, g[c + "_rhs"])
        assert nit == k
        assert np.array_equal(var, g[c + "_var_%d" % k]), c


def test_fb2d9_full_solve_matches_reference(orc):
    g = load_golden("msi_small.npz")
    for c in _cases(g):
        nx, ny = (int(v) for v in g[c + "_shape"])
        var, nit, r0, hist = orc.fb2d9(g[c + "_a"], g[c + "_dl"], g[c + "_du"], nx, ny,
                                       np.zeros(nx * ny), 1e-10, 500, g[c + "_rhs"])
        assert np.array_equal(var, g[c + "_full"]), c
        # the stopping rule fired on the strict test of msi2d9-v03.f90:258
        assert hist[-1] < 1e-10 * r0 or nit == 500


def test_cfg1_iterates_match_reference(orc):
    g = load_golden("cfg1.npz")
    nx, ny = (int(v) for v in g["shape"])
    dl, du = orc.lu2d9(g["a"], nx, ny)
    assert np.array_equal(dl[0], g["dl1"]) and np.array_equal(du[0], g["du1"])
    for k in (1, 10, 100):
        var, nit, _, _ = orc.fb2d9(g["a"], dl, du, nx, ny, np.zeros(nx * ny), 0.0, k, g["rhs"])
        assert np.array_equal(var, g["var_%d" % k]), k
    var, nit, r0, hist = orc.fb2d9(g["a"], dl, du, nx, ny, np.zeros(nx * ny), 1e-10, 500,
                                   g["rhs"])
    assert np.array_equal(var, g["var_full"])
    assert len(hist) == nit


@pytest.mark.parametrize("k", [1, 2, 5])
def test_tdma2d9_matches_reference(orc, k):
    g = load_golden("tdma.npz")
    names = sorted({n.rsplit("_", 1)[0] for n in g if n.endswith("_shape")})
    for c in names:
        nx, ny = (int(v) for v in g[c + "_shape"])
        x = orc.tdma2d9(g[c + "_a"], g[c + "_b"], g[c + "_x0"], nx, ny, k)
        assert np.array_equal(x, g[c + "_x%d" % k]), c


def test_norms_match_reference(orc):
    g = load_golden("norms.npz")
    names = sorted({n.rsplit("_", 1)[0] for n in g if n.endswith("_shape")})
    for c in names:
        nx, ny = (int(v) for v in g[c + "_shape"])
        a, b, v = g[c + "_a"], g[c + "_b"], g[c + "_var"]
        assert orc.norm_l1_9d(a, b, v, nx, ny, 0.25) == float(g[c + "_l1"])
        assert orc.norm_l1_9d_relative(a, b, v, nx, ny, 0.25) == float(g[c + "_rel"])
        assert orc.norm_l1_9d_relative(a, np.zeros_like(b), v, nx, ny, 0.25) == float(
            g[c + "_rel0"])
        assert orc.norm_l1_b(b, nx, ny) == float(g[c + "_nb"])


def test_msi5d_and_embedding_match_reference(orc):
    """F1 identity on the reference itself: lu2d9+fb2d9 on the zero-corner
    embedding reproduce lu2d5+fb2d5; the oracle reproduces both."""
    g = load_golden("embed5.npz")
    names = sorted({n.rsplit("_", 1)[0] for n in g if n.endswith("_shape")})
    for c in names:
        nx, ny = (int(v) for v in g[c + "_shape"])
        a5, rhs = g[c + "_a5"], g[c + "_rhs"]
        dl5, du5 = orc.lu2d5(a5, nx, ny)
        assert np.array_equal(dl5, g[c + "_dl5"]) and np.array_equal(du5, g[c + "_du5"])
        x5, _, _, _ = orc.fb2d5(a5, dl5, du5, nx, ny, np.zeros(nx * ny), 0.0, 7, rhs)
        assert np.array_equal(x5, g[c + "_x5"])
        a9 = np.zeros((9, nx * ny))
        a9[[1, 3, 4, 5, 7]] = a5
        dl9, du9 = orc.lu2d9(a9, nx, ny)
        x9, _, _, _ = orc.fb2d9(a9, dl9, du9, nx, ny, np.zeros(nx * ny), 0.0, 7, rhs)
        assert np.array_equal(x9, g[c + "_x9"])
        assert np.array_equal(x9, x5)


def test_solver_object_matches_reference(orc):
    g = load_golden("solver_msi9d.npz")
    nx, ny = (int(v) for v in g["shape"])
    var, _, _, _ = orc.solve_msi9d(g["a"], g["b"], g["sol0"], nx, ny, int(g["nitm"]),
                                   float(g["tol"]))
    assert np.array_equal(var, g["sol"])


def test_oracle_matches_live_reference_when_built(orc):
    """Where oracle/_ref exists (it travels with the tree), compare live on a
    system not in the fixtures."""
    if not orc.ref_available():
        pytest.skip("reference library not built")
    from mach2d import problems as P
    s = P.pcorr_system(41, 23, seed=77)
    dl, du = orc.lu2d9(s.a, s.nx, s.ny)
    rdl, rdu = orc.ref_lu2d9(s.a, s.nx, s.ny)
    assert np.array_equal(dl, rdl) and np.array_equal(du, rdu)
    v, nit, _, _ = orc.fb2d9(s.a, dl, du, s.nx, s.ny, np.zeros(s.N), 1e-8, 300, s.rhs)
    rv = orc.ref_fb2d9(s.a, dl, du, s.nx, s.ny, np.zeros(s.N), 1e-8, 300, s.rhs)
    assert np.array_equal(v, rv)


# ---- coefficient assembly (mach2d_coef.f90) -----------------------------------
COEF_TAGS = ["ns_axi", "euler", "ns_planar", "cart", "axis"]


def coef_case(g, tag):
    nx, ny, coord, mv, beta, dt = g["%s_scalars" % tag]
    nx, ny, coord, mv = int(nx), int(ny), int(coord), int(mv)
    pre = "%s_in_" % tag
    fields = {k[len(pre):]: v for k, v in g.items() if k.startswith(pre)}
    twall = float(g["%s_twall" % tag][0])
    return nx, ny, dict(coord=coord, modvis=mv, beta=beta, dt=dt, Rg=287.0, Twall=twall,
                        Cpref=1004.5, Href=837083.3, gamma=1.4, P0=1038350.4,
                        T0=833.33333333), fields


@pytest.mark.parametrize("tag", COEF_TAGS)
def test_coef_oracle_vs_reference_fixtures(orc, tag):
    """The C restatement of every assembly routine equals the reference's
    output bit for bit, including the points the routine leaves untouched."""
    g = load_golden("coef.npz")
    nx, ny, sc, f = coef_case(g, tag)
    for name, (_, _, outs) in orc.COEF.items():
        init = {k: g["%s_%s_init_%s" % (tag, name, k)] for k, _ in outs}
        got = orc.coef(name, nx, ny, sc, f, out=init)
        for k, m in outs:
            want = g["%s_%s_%s" % (tag, name, k)]
            if m in ("nx", "ny"):
                want = want.reshape(-1)
                got[k] = got[k].reshape(-1)
            if m == 0:
                want = float(np.asarray(want).reshape(-1)[0])
            assert np.array_equal(got[k], want), (tag, name, k)
            assert np.all(np.isfinite(want)), (tag, name, k)


@pytest.mark.parametrize("nx,ny", [(12, 9), (70, 33), (130, 66)])
def test_msi5d_embedding_identity_on_generated_systems(orc, nx, ny):
    """The identity the 5-diagonal GPU factorization relies on: on the
    generated 5-point p' systems, lu2d9 of the zero-corner embedding gives
    lu2d5's dl1..dl4 and du1..du3 and exact zeros in dl5 and du4, and
    fb2d9 on the embedding iterates exactly as fb2d5."""
    from mach2d import problems as P
    s = P.pcorr_system5(nx, ny, seed=nx)
    dl5, du5 = orc.lu2d5(s.a, nx, ny)
    a9 = np.zeros((9, s.N))
    a9[[1, 3, 4, 5, 7]] = s.a
    dl9, du9 = orc.lu2d9(a9, nx, ny)
    assert np.array_equal(dl9[:4], dl5) and np.array_equal(du9[:3], du5)
    assert not dl9[4].any() and not du9[3].any()
    x5, n5, _, h5 = orc.fb2d5(s.a, dl5, du5, nx, ny, np.zeros(s.N), 1e-8, 3000, s.rhs)
    x9, n9, _, h9 = orc.fb2d9(a9, dl9, du9, nx, ny, np.zeros(s.N), 1e-8, 3000, s.rhs)
    assert n5 == n9 and np.array_equal(x5, x9) and np.array_equal(h5, h9)


@pytest.mark.parametrize("modvis", [0, 1])
def test_pl_loop_oracle_vs_reference(orc, modvis):
    """The pressure-correction loop composed from the oracle's routines
    equals the same composition of the compiled reference's routines bit for
    bit (state after three itp iterations)."""
    import pl_loop_oracle as PL
    if not orc.ref_available():
        pytest.skip("reference library not built")
    nx, ny = 26, 14
    f = PL.loop_fields(nx, ny, seed=5)
    p_avg = float(np.mean(f["p"]))
    a, ia, na = PL.run(orc, f, nx, ny, modvis, 2e-3, 0.3, p_avg, 0.0, 3, 20, 1e-2)
    b, ib, nb = PL.run(orc, f, nx, ny, modvis, 2e-3, 0.3, p_avg, 0.0, 3, 20, 1e-2, ref=True)
    assert ia == ib == 3 and na == nb
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def _scarb_inputs(nx, ny, seed):
    from mach2d import problems as P
    rng = np.random.default_rng(seed)
    au = P.pcorr_system(nx, ny, seed=seed).a
    av = rng.uniform(-2.0, 2.0, size=(9, nx * ny))
    at = rng.uniform(-1.0, 1.0, size=(9, nx * ny))
    ap = P.pcorr_system5(nx, ny, seed=seed + 1).a
    return au, av, at, ap


@pytest.mark.parametrize("nx,ny", [(3, 3), (12, 9), (70, 33)])
def test_scarborough_oracle_vs_reference(orc, nx, ny, tmp_path):
    """get_mean_cc_9d / _5d (the reference's own module procedures, called
    through their symbols), tstepper_dt0_scarborough and
    tstepper_dt_scarborough on every branch of h (c < mincc, between,
    c > maxcc), bit for bit."""
    if not orc.ref_available():
        pytest.skip("reference library not built")
    au, av, at, ap = _scarb_inputs(nx, ny, nx + ny)
    for a in (au, av, at, ap):
        assert orc.mean_cc(a, nx, ny) == orc.ref_mean_cc(a, nx, ny)
    rng = np.random.default_rng(1)
    J, u, v = (rng.uniform(0.5, 2.0, nx * ny), rng.uniform(-5, 5, nx * ny),
               rng.uniform(-5, 5, nx * ny))
    assert orc.dt0_scarborough(J, u, v, nx, ny) == orc.ref_dt0_scarborough(J, u, v, nx, ny)
    c = max(orc.mean_cc(a, nx, ny) for a in (au, av, at, ap))
    for mincc, maxcc in ((2 * c, 4 * c), (0.5 * c, 2 * c), (0.1 * c, 0.5 * c)):
        pf = tmp_path / "scarb.txt"
        pf.write_text("   0.25 & h0 & amplitude\n %r & mincc & min\n %r & maxcc & max\n"
                      % (mincc, maxcc))
        want = orc.ref_dt_scarborough(str(pf), 1.5e-3, au, av, at, ap, nx, ny)
        assert orc.dt_scarborough(1.5e-3, au, av, at, ap, nx, ny, 0.25, mincc, maxcc) == want


# ---------------------------------------------------------------------------
# Grid centroids, metrics and radii (oracle/grid.c).  mod_grid.f90 does not
# build here (its module uses mod_extflow, whose geometry module needs the
# absent depp_interface), so these checks pin the restatement by the
# mathematics: affine maps give constant metrics at every point the
# reference writes (one-sided boundary differences included), tensor grids
# give the spacing formulas, 1/Jp is the area of each quadrilateral, the
# weighted centroid is the polygon centroid, and alpha, gamma, beta, J are
# invariant under rotation.
# ---------------------------------------------------------------------------

def _affine_grid(nx, ny, A, b):
    """corner (i, j), 1-based, at A (i, j) + b; flat index nx*(j-1) + i - 1"""
    i, j = np.meshgrid(np.arange(1, nx + 1, dtype=float), np.arange(1, ny + 1, dtype=float))
    x = A[0][0] * i + A[0][1] * j + b[0]
    y = A[1][0] * i + A[1][1] * j + b[1]
    return x.reshape(-1), y.reshape(-1)


def _masks(nx, ny):
    I, J = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1))
    east = (J >= 2) & (J <= ny - 1) & (I <= nx - 1)
    north = (J <= ny - 1) & (I >= 2) & (I <= nx - 1)
    inner = (I >= 2) & (I <= nx - 1) & (J >= 2) & (J <= ny - 1)
    return east.reshape(-1), north.reshape(-1), inner.reshape(-1)


@pytest.mark.parametrize("nx,ny", [(3, 3), (4, 7), (19, 11)])
def test_grid_metrics_affine_closed_forms(orc, nx, ny):
    A = [[0.7, -0.4], [0.3, 1.1]]
    x, y = _affine_grid(nx, ny, A, (2.0, -1.0))
    for opt in (1, 2):
        xp, yp = orc.centroids(x, y, nx, ny, opt)
        east, north, inner = _masks(nx, ny)
        m = orc.grid_metrics(x, y, xp, yp, nx, ny)
        det = A[0][0] * A[1][1] - A[0][1] * A[1][0]
        want_e = {"xe": A[0][1], "ye": A[1][1], "xke": A[0][0], "yke": A[1][0],
                  "alphae": A[0][1] ** 2 + A[1][1] ** 2,
                  "betae": A[0][0] * A[0][1] + A[1][0] * A[1][1], "Je": 1 / abs(det)}
        want_n = {"xk": A[0][0], "yk": A[1][0], "xen": A[0][1], "yen": A[1][1],
                  "gamman": A[0][0] ** 2 + A[1][0] ** 2,
                  "betan": A[0][0] * A[0][1] + A[1][0] * A[1][1], "Jn": 1 / abs(det)}
        for k, v in want_e.items():
            np.testing.assert_allclose(m[k][east], v, rtol=1e-12, atol=1e-12, err_msg=k)
            assert not m[k][~east].any(), k
        for k, v in want_n.items():
            np.testing.assert_allclose(m[k][north], v, rtol=1e-12, atol=1e-12, err_msg=k)
            assert not m[k][~north].any(), k
        np.testing.assert_allclose(m["Jp"][inner], 1 / abs(det), rtol=1e-12)
        assert not m["Jp"][~inner].any()
        # centroids: the parallelogram's centre, zero outside real volumes
        cx, cy = _affine_grid(nx, ny, A, (2.0 - 0.5 * (A[0][0] + A[0][1]),
                                          -1.0 - 0.5 * (A[1][0] + A[1][1])))
        np.testing.assert_allclose(xp[inner], cx[inner], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(yp[inner], cy[inner], rtol=1e-12, atol=1e-12)
        assert not xp[~inner].any() and not yp[~inner].any()


def test_grid_metrics_stretched_tensor_grid(orc):
    """Cartesian grid with geometric spacing: xke is the centroid spacing
    inside and the adjacent cell width at i = 1 and nx-1 (likewise yen);
    Jp = 1 / (cell area); beta = 0."""
    nx, ny = 13, 9
    X = np.cumsum(np.r_[0.0, 0.3 * 1.2 ** np.arange(nx - 1)])    # corners i = 1..nx
    Y = np.cumsum(np.r_[0.0, 0.1 * 1.15 ** np.arange(ny - 1)])
    x = np.tile(X, ny)
    y = np.repeat(Y, nx)
    xp, yp = orc.centroids(x, y, nx, ny)
    m = orc.grid_metrics(x, y, xp, yp, nx, ny)
    g = lambda a, i, j: a[nx * (j - 1) + i - 1]
    for j in range(2, ny):
        assert g(m["xke"], 1, j) == pytest.approx(X[1] - X[0], rel=1e-12)
        assert g(m["xke"], nx - 1, j) == pytest.approx(X[nx - 2] - X[nx - 3], rel=1e-12)
        for i in range(2, nx - 1):
            assert g(m["xke"], i, j) == pytest.approx((X[i] - X[i - 2]) / 2, rel=1e-12)
    for i in range(2, nx):
        assert g(m["yen"], i, 1) == pytest.approx(Y[1] - Y[0], rel=1e-12)
        assert g(m["yen"], i, ny - 1) == pytest.approx(Y[ny - 2] - Y[ny - 3], rel=1e-12)
        for j in range(2, ny - 1):
            assert g(m["yen"], i, j) == pytest.approx((Y[j] - Y[j - 2]) / 2, rel=1e-12)
        for j in range(2, ny):
            area = (X[i - 1] - X[i - 2]) * (Y[j - 1] - Y[j - 2])
            assert g(m["Jp"], i, j) == pytest.approx(1 / area, rel=1e-12)
    for k in ("betae", "betan", "yke", "xen"):    # zero up to rounding
        assert np.abs(m[k]).max() < 1e-14, k


def _perturbed_grid(nx, ny, seed, amp=0.2):
    rng = np.random.default_rng(seed)
    x, y = _affine_grid(nx, ny, [[1.0, 0.15], [-0.1, 0.8]], (0.0, 0.0))
    return x + amp * rng.uniform(-1, 1, x.size), y + amp * rng.uniform(-1, 1, y.size)


def test_grid_jp_area_and_weighted_centroid(orc):
    """1/Jp is the shoelace area of each quadrilateral; opt 2 gives its
    polygon centroid."""
    nx, ny = 12, 10
    x, y = _perturbed_grid(nx, ny, 3)
    xp, yp = orc.centroids(x, y, nx, ny, opt=2)
    m = orc.grid_metrics(x, y, xp, yp, nx, ny)
    for j in range(2, ny):
        for i in range(2, nx):
            q = [nx * (j - 2) + i - 2, nx * (j - 2) + i - 1, nx * (j - 1) + i - 1,
                 nx * (j - 1) + i - 2]                          # SW, SE, NE, NW
            px, py = x[q], y[q]
            cr = px * np.roll(py, -1) - np.roll(px, -1) * py
            area = 0.5 * cr.sum()
            cx = ((px + np.roll(px, -1)) * cr).sum() / (6 * area)
            cy = ((py + np.roll(py, -1)) * cr).sum() / (6 * area)
            p = nx * (j - 1) + i - 1
            assert m["Jp"][p] == pytest.approx(1 / abs(area), rel=1e-12)
            assert xp[p] == pytest.approx(cx, rel=1e-12, abs=1e-12)
            assert yp[p] == pytest.approx(cy, rel=1e-12, abs=1e-12)
            assert xp[p] != pytest.approx(px.mean(), rel=1e-9, abs=1e-9) or \
                yp[p] != pytest.approx(py.mean(), rel=1e-9, abs=1e-9)


def test_grid_metrics_rotation_invariance(orc):
    nx, ny = 15, 8
    x, y = _perturbed_grid(nx, ny, 5)
    th = 0.7
    xr, yr = np.cos(th) * x - np.sin(th) * y, np.sin(th) * x + np.cos(th) * y
    m = orc.grid_metrics(x, y, *orc.centroids(x, y, nx, ny), nx, ny)
    mr = orc.grid_metrics(xr, yr, *orc.centroids(xr, yr, nx, ny), nx, ny)
    for k in ("alphae", "gamman", "betae", "betan", "Je", "Jn", "Jp"):
        np.testing.assert_allclose(mr[k], m[k], rtol=1e-10, atol=1e-12, err_msg=k)
    # face tangents rotate as vectors
    np.testing.assert_allclose(mr["xke"], np.cos(th) * m["xke"] - np.sin(th) * m["yke"],
                               atol=1e-12)


@pytest.mark.parametrize("nx,ny", [(3, 3), (6, 5), (17, 12)])
def test_grid_radius_affine_reflection(orc, nx, ny):
    """coord 1 on an affine grid: every rp (fictitious volumes and corners
    included) is y at the volume's centre, re / rn the face centres; coord 0
    gives ones."""
    A = [[0.4, 0.2], [0.1, 0.9]]
    x, y = _affine_grid(nx, ny, A, (0.5, 3.0))
    xp, yp = orc.centroids(x, y, nx, ny)
    rad, re, rn, rp = orc.radius(1, y, yp, nx, ny)
    _, cy = _affine_grid(nx, ny, A, (0.5 - 0.5 * (A[0][0] + A[0][1]),
                                     3.0 - 0.5 * (A[1][0] + A[1][1])))
    np.testing.assert_allclose(rp, cy, rtol=1e-12, atol=1e-12)
    assert np.array_equal(rad, y)
    east, north, _ = _masks(nx, ny)
    _, ey = _affine_grid(nx, ny, A, (0.5, 3.0 - 0.5 * A[1][1]))
    _, ny_ = _affine_grid(nx, ny, A, (0.5, 3.0 - 0.5 * A[1][0]))
    np.testing.assert_allclose(re[east], ey[east], rtol=1e-12)
    np.testing.assert_allclose(rn[north], ny_[north], rtol=1e-12)
    assert (re[~east] == 1).all() and (rn[~north] == 1).all()
    for a in orc.radius(0, y, yp, nx, ny):
        assert (a == 1).all()


# ---------------------------------------------------------------------------
# The rest of one internal-flow time step (oracle/intflow.c) against the
# compiled reference: bit for bit.
# ---------------------------------------------------------------------------

def _step_inputs(nx, ny, seed):
    from mach2d import problems as P
    g = P.ogive_grid(nx, ny, 1.05)
    m = P.metrics(g)
    rng = np.random.default_rng(seed)
    f = {k: np.ascontiguousarray(m[k].reshape(-1)) for k in ("xp", "xk", "yk")}
    f["x"] = np.ascontiguousarray(g.x.reshape(-1))
    n = nx * ny
    for k in ("u", "ue", "un", "v", "ve", "vn"):
        f[k] = rng.uniform(20.0, 120.0, n) * rng.choice([-1.0, 1.0], n)
    f["u"] = rng.uniform(30.0, 90.0, n)
    f["Tin"] = rng.uniform(780.0, 830.0, ny)
    return f, rng


@pytest.mark.parametrize("nx,ny", [(4, 3), (5, 6), (23, 14), (130, 66)])
def test_intflow_step_routines_vs_reference(orc, nx, ny):
    if not orc.ref_available():
        pytest.skip("reference library not built")
    f, rng = _step_inputs(nx, ny, nx * ny)
    n = nx * ny
    for modvis in (0, 1):
        init = {k: rng.uniform(-1, 1, n) for k in ("due", "dve", "dun", "dvn", "de", "dn")}
        a = orc.coef("intflow_boundary_simplec", nx, ny, {"modvis": modvis}, {}, init)
        b = orc.ref_coef("intflow_boundary_simplec", nx, ny, {"modvis": modvis}, {}, init)
        for k in a:
            assert np.array_equal(a[k], b[k]), (modvis, k)
        init = {"u": rng.uniform(10, 90, n), "v": rng.uniform(-9, 9, n)}
        a = orc.coef("intflow_extrapolate_u_v_to_fictitious", nx, ny, {"modvis": modvis}, f, init)
        b = orc.ref_coef("intflow_extrapolate_u_v_to_fictitious", nx, ny, {"modvis": modvis}, f,
                         init)
        for k in a:
            assert np.array_equal(a[k], b[k]), (modvis, k)
            assert not np.array_equal(a[k], init[k])
    sc = {"Cpref": 1004.5, "Href": 837000.0}
    init = {"T": rng.uniform(700, 830, n), "Tbe": np.full(ny, -1.0), "Tbw": np.full(ny, -1.0),
            "Tbn": np.full(nx, -1.0), "Tbs": np.full(nx, -1.0)}
    a = orc.coef("intflow_boundary_T_from_H_conservation", nx, ny, sc, f, init)
    b = orc.ref_coef("intflow_boundary_T_from_H_conservation", nx, ny, sc, f, init)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    sc = {"gamma": 1.4, "Rg": 287.0, "P0": 1038350.4, "T0": 833.33333333}
    init = {k: rng.uniform(1, 2, ny) for k in ("uin", "vin", "Tin", "Mw", "plin", "pina")}
    init["pin"] = rng.uniform(9.9e5, 1.03e6, ny)
    init["p"] = rng.uniform(9.0e5, 1.0e6, n)
    a = orc.coef("intflow_time_cycle_update", nx, ny, sc, f, init)
    b = orc.ref_time_cycle_update(nx, ny, sc, f, init)
    for k in a:
        assert np.array_equal(a[k].reshape(-1), b[k].reshape(-1)), k
    assert np.array_equal(a["pina"].reshape(-1)[1:-1], init["pin"][1:-1])
    # norms and the average
    a5 = rng.uniform(-1, 1, (5, n))
    var, rhs = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    assert orc.norm_l1_5d(nx, ny, var, rhs, a5, 0.25) == \
        orc.norm_l1_5d(nx, ny, var, rhs, a5, 0.25, ref_=True)
    F = rng.uniform(9e5, 1e6, n)
    assert orc.get_average(nx, ny, F) == orc.get_average(nx, ny, F, ref_=True)


def test_intflow_step_routines_invariants(orc):
    """Closed forms the reference's formulas fix: at rest the inflow keeps
    T0 and P0 (Mw = 0), plin is pin - pina, and the west fictitious pressure
    averages to pin; the T writer puts the face temperature midway between
    the fictitious and real values; the extrapolation is exact for u linear
    in xp along rows."""
    nx, ny = 9, 7
    f, rng = _step_inputs(nx, ny, 3)
    n = nx * ny
    f0 = dict(f, u=np.zeros(n))
    sc = {"gamma": 1.4, "Rg": 287.0, "P0": 1.0e6, "T0": 800.0}
    init = {k: np.zeros(ny) for k in ("uin", "vin", "Tin", "Mw", "plin", "pina")}
    init["pin"] = np.full(ny, 0.9e6)
    init["p"] = rng.uniform(9.0e5, 1.0e6, n)
    a = orc.coef("intflow_time_cycle_update", nx, ny, sc, f0, init)
    J = slice(1, ny - 1)
    assert (a["Tin"].reshape(-1)[J] == 800.0).all() and (a["pin"].reshape(-1)[J] == 1.0e6).all()
    assert np.allclose(a["plin"].reshape(-1)[J], 0.1e6)
    p = a["p"].reshape(ny, nx)
    assert np.allclose((p[J, 0] + p[J, 1]) / 2, 1.0e6)
    assert np.array_equal(p[0, 1:-1], p[1, 1:-1]) and np.array_equal(p[-1, 1:-1], p[-2, 1:-1])
    sc = {"Cpref": 1000.0, "Href": 9.0e5}
    init = {"T": rng.uniform(700, 830, n), "Tbe": np.zeros(ny), "Tbw": np.zeros(ny),
            "Tbn": np.zeros(nx), "Tbs": np.zeros(nx)}
    a = orc.coef("intflow_boundary_T_from_H_conservation", nx, ny, sc, f, init)
    T = a["T"].reshape(ny, nx)
    np.testing.assert_allclose((T[J, 0] + T[J, 1]) / 2, f["Tin"][J], rtol=1e-14)
    ue, ve = f["ue"].reshape(ny, nx), f["ve"].reshape(ny, nx)
    np.testing.assert_allclose((T[J, -1] + T[J, -2]) / 2,
                               (9.0e5 - (ue[J, -2] ** 2 + ve[J, -2] ** 2) / 2) / 1000.0,
                               rtol=1e-13)
    xp = f["xp"].reshape(ny, nx)
    u = 3.0 + 0.5 * xp
    a = orc.coef("intflow_extrapolate_u_v_to_fictitious", nx, ny, {"modvis": 1}, f,
                 {"u": u.reshape(-1), "v": np.zeros(n)})
    uw = a["u"].reshape(ny, nx)[J, 0]
    np.testing.assert_allclose(uw, 3.0 + 0.5 * (-xp[J, 1]), rtol=1e-12)


@pytest.mark.parametrize("modvis,solver9d,nx,ny,mach,dt0",
                         [(0, 0, 22, 12, None, 2e-5), (0, 1, 22, 12, None, 2e-5),
                          (1, 0, 18, 10, None, 2e-5), (0, 0, 66, 34, 0.2, 5e-6),
                          (0, 1, 66, 34, 0.2, 5e-6)])
def test_time_step_oracle_vs_reference_composition(orc, modvis, solver9d, nx, ny, mach, dt0):
    """One internal-flow time step composed from the oracle's routines equals
    the same step composed from the compiled reference's own routines, bit
    for bit: every field, the next dt, the norms, rmass and itp (mach: the
    nozzle state at that inlet Mach number)."""
    import os

    import time_step_oracle as TS
    if not orc.ref_available():
        pytest.skip("reference library not built")
    f = TS.step_fields(nx, ny, seed=nx + modvis, mach_in=mach)
    prm = dict(TS.PARAMS, modvis=modvis, solver9d=solver9d)
    pf = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "parametersNozzleEulerPTC.txt")
    a, dta, ra = TS.step(orc, f, nx, ny, prm, dt0)
    b, dtb, rb = TS.step(orc, f, nx, ny, prm, dt0, ref=True, paramfile=pf)
    assert dta == dtb and ra == rb, (dta, dtb, ra, rb)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert ra["itp"] >= 1 and np.isfinite(ra["norm"]) and dta > 0
    for k in ("u", "v", "p", "T", "ro", "pl", "pin"):
        assert not np.array_equal(a[k], f[k]), k


def test_problem_radii_match_oracle_radius(orc):
    """The synthetic-problem radii (mach2d.problems.radii) equal get_radius's
    restatement bit for bit."""
    import time_step_oracle as TS
    from mach2d import problems as P
    nx, ny = 21, 9
    g = TS.nozzle_grid(nx, ny)
    m = P.metrics(g)
    r = P.radii(g, m["yp"])
    _, re, rn, rp = orc.radius(1, g.y.reshape(-1), m["yp"].reshape(-1), nx, ny)
    assert np.array_equal(r["re"].reshape(-1), re)
    assert np.array_equal(r["rn"].reshape(-1), rn)
    assert np.array_equal(r["rp"].reshape(-1), rp)


def _grid_cases():
    g = load_golden("grid.npz")
    for tag in [str(c) for c in g["cases"]]:
        nx, ny = (int(v) for v in tag.rsplit("_", 1)[1].split("x"))
        yield g, tag, nx, ny


def test_grid_oracle_vs_reference_fixtures(orc):
    """get_real_centroids_xy, get_metrics and get_radius restated in
    oracle/grid.c equal the reference's own module procedures bit for bit
    (fixture written by the compiled mod_grid.f90) on ogive, nozzle,
    perturbed skewed and strongly stretched grids."""
    for g, tag, nx, ny in _grid_cases():
        x, y = g[tag + "_x"], g[tag + "_y"]
        for opt in (1, 2):
            xp, yp = orc.centroids(x, y, nx, ny, opt)
            assert np.array_equal(xp, g["%s_xp%d" % (tag, opt)]), (tag, opt)
            assert np.array_equal(yp, g["%s_yp%d" % (tag, opt)]), (tag, opt)
        m = orc.grid_metrics(x, y, xp, yp, nx, ny)
        for k in orc.METRICS:
            assert np.array_equal(m[k], g["%s_m_%s" % (tag, k)]), (tag, k)
        for coord in (0, 1):
            for k, v in zip(("radius", "re", "rn", "rp"), orc.radius(coord, y, yp, nx, ny)):
                assert np.array_equal(v, g["%s_r%d_%s" % (tag, coord, k)]), (tag, coord, k)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_grid_oracle_vs_reference_live(orc, seed):
    """The same three routines on seeded random grids (sizes and
    perturbations drawn per seed), against the compiled reference."""
    if not orc.ref_available():
        pytest.skip("reference library not built")
    rng = np.random.default_rng(seed)
    nx, ny = int(rng.integers(3, 60)), int(rng.integers(3, 40))
    x, y = _perturbed_grid(nx, ny, seed, amp=float(rng.uniform(0.0, 0.3)))
    for opt in (1, 2):
        a, b = orc.centroids(x, y, nx, ny, opt), orc.ref_centroids(x, y, nx, ny, opt)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), opt
    xp, yp = b
    a, b = orc.grid_metrics(x, y, xp, yp, nx, ny), orc.ref_grid_metrics(x, y, xp, yp, nx, ny)
    for k in orc.METRICS:
        assert np.array_equal(a[k], b[k]), k
    for coord in (0, 1):
        for u, v in zip(orc.radius(coord, y, yp, nx, ny), orc.ref_radius(coord, y, yp, nx, ny)):
            assert np.array_equal(u, v), coord
