"""GPU parity of the grid centroids, metrics and radii
(src/grid/mod_grid.f90:237-723), bit for bit, with host and device arrays:
against the fixture the reference's own module procedures wrote
(tests/golden/grid.npz) and against the oracle, which equals the compiled
reference bit for bit (tests/test_oracle_golden.py)."""
import numpy as np
import pytest

from conftest import load_golden
from mach2d import problems as P
from test_oracle_golden import _perturbed_grid

pytestmark = pytest.mark.gpu

CASES = [("ogive", 3, 3), ("ogive", 4, 9), ("perturbed", 17, 12), ("ogive", 130, 66),
         ("perturbed", 300, 257), ("ogive", 1026, 514)]


def _grid(kind, nx, ny):
    if kind == "ogive":
        g = P.ogive_grid(nx, ny, 1.05)
        return g.x.reshape(-1).copy(), g.y.reshape(-1).copy()
    return _perturbed_grid(nx, ny, nx + ny)


@pytest.mark.parametrize("kind,nx,ny", CASES)
def test_grid_create_tail_matches_oracle(gpu, orc, kind, nx, ny):
    import torch

    from mach2d import grid
    x, y = _grid(kind, nx, ny)
    for opt in (1, 2):
        oxp, oyp = orc.centroids(x, y, nx, ny, opt)
        xp, yp = grid.centroids(x, y, nx, ny, opt)
        assert np.array_equal(xp, oxp) and np.array_equal(yp, oyp), opt
    om = orc.grid_metrics(x, y, oxp, oyp, nx, ny)
    m = grid.metrics(x, y, oxp, oyp, nx, ny)
    for k in grid.METRICS:
        assert np.array_equal(m[k], om[k]), k
    for coord in (0, 1):
        want = orc.radius(coord, y, oyp, nx, ny)
        got = grid.radius(coord, y, oyp, nx, ny)
        for a, b in zip(got, want):
            assert np.array_equal(a, b), coord
    # device arrays, every step on the GPU
    dx, dy = (torch.tensor(v, device="cuda") for v in (x, y))
    d = grid.grid_create(dx, dy, nx, ny, kcm=2, coord=1)
    torch.cuda.synchronize()
    for k in grid.METRICS:
        assert np.array_equal(d[k].cpu().numpy(), om[k]), k
    for k, b in zip(("radius", "re", "rn", "rp"), orc.radius(1, y, oyp, nx, ny)):
        assert np.array_equal(d[k].cpu().numpy(), b), k


def test_grid_rejects_small_and_mixed(gpu):
    import torch

    from mach2d import grid
    from mach2d._lib import M2D_ERR_ARG
    x, y = _grid("ogive", 4, 4)
    with pytest.raises(gpu.SolverError) as e:
        grid.centroids(x[:4], y[:4], 2, 2)
    assert e.value.code == M2D_ERR_ARG
    # one device and one host input: refused before any launch
    import ctypes as C

    from mach2d import _lib
    t = torch.tensor(y, device="cuda")
    outs = [torch.zeros(16, dtype=torch.float64, device="cuda") for _ in range(4)]
    i4 = C.byref(C.c_int(4))
    _lib.load().m2d_get_radius(C.byref(C.c_int(1)), i4, i4, C.c_void_p(t.data_ptr()),
                               C.c_void_p(y.ctypes.data), *[C.c_void_p(o.data_ptr()) for o in outs])
    assert _lib.last_status().status == M2D_ERR_ARG


def test_grid_kernels_vs_reference_fixtures(gpu):
    """k_centroids, k_metrics and k_radius against the outputs of the
    reference's own get_real_centroids_xy, get_metrics and get_radius."""
    import torch

    from mach2d import grid
    g = load_golden("grid.npz")
    for tag in [str(c) for c in g["cases"]]:
        nx, ny = (int(v) for v in tag.rsplit("_", 1)[1].split("x"))
        x, y = g[tag + "_x"], g[tag + "_y"]
        for opt in (1, 2):
            xp, yp = grid.centroids(x, y, nx, ny, opt)
            assert np.array_equal(xp, g["%s_xp%d" % (tag, opt)]), (tag, opt)
            assert np.array_equal(yp, g["%s_yp%d" % (tag, opt)]), (tag, opt)
        m = grid.metrics(x, y, xp, yp, nx, ny)
        for k in grid.METRICS:
            assert np.array_equal(m[k], g["%s_m_%s" % (tag, k)]), (tag, k)
        for coord in (0, 1):
            for k, v in zip(("radius", "re", "rn", "rp"), grid.radius(coord, y, yp, nx, ny)):
                assert np.array_equal(v, g["%s_r%d_%s" % (tag, coord, k)]), (tag, coord, k)
        dx, dy = (torch.tensor(v, device="cuda") for v in (x, y))
        d = grid.grid_create(dx, dy, nx, ny, kcm=2, coord=1)
        torch.cuda.synchronize()
        for k in grid.METRICS:
            assert np.array_equal(d[k].cpu().numpy(), g["%s_m_%s" % (tag, k)]), (tag, k)
        for k in ("radius", "re", "rn", "rp"):
            assert np.array_equal(d[k].cpu().numpy(), g["%s_r1_%s" % (tag, k)]), (tag, k)
