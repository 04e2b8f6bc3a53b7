"""GPU parity of the coefficient assembly (m2d_get_*, mach2d_coef.f90)
against the oracle, through the C ABI, with host and with device arrays.

Bar: bit for bit (the kernels keep the reference's operation order without
contraction); points the reference does not write keep their values."""
import numpy as np
import pytest

from mach2d import problems as P

pytestmark = pytest.mark.gpu

CASES = [  # nx, ny, coord, modvis, cartesian
    (12, 8, 1, 1, False), (12, 8, 0, 0, False), (34, 18, 1, 1, False), (10, 7, 1, 1, True),
    (300, 70, 1, 1, False), (3, 3, 1, 1, False), (258, 3, 0, 1, False)]


def _fields(nx, ny, coord, cart, seed):
    f = P.flow_fields(nx, ny, seed=seed, coord=coord, cartesian=cart)
    return f


@pytest.mark.parametrize("case", CASES)
def test_assembly_host_arrays_bitwise(gpu, orc, case):
    from mach2d import assembly as A
    nx, ny, coord, mv, cart = case
    f = _fields(nx, ny, coord, cart, seed=nx * ny + coord)
    sc = dict(beta=0.3, modvis=mv, dt=1e-3, coord=coord)
    rng = np.random.default_rng(7)
    for name, (_, _, outs) in A.ROUTINES.items():
        init = {k: rng.uniform(-1, 1, size=m * nx * ny) for k, m in outs}
        want = orc.coef(name, nx, ny, sc, f, out={k: v.copy() for k, v in init.items()})
        got = A.call(name, nx, ny, sc, f, out={k: v.copy() for k, v in init.items()})
        for k in want:
            assert np.array_equal(np.asarray(got[k]), want[k]), (name, k, case)


@pytest.mark.parametrize("case", CASES[:4])
def test_assembly_device_arrays_bitwise(gpu, orc, case):
    import torch

    from mach2d import assembly as A
    nx, ny, coord, mv, cart = case
    f = _fields(nx, ny, coord, cart, seed=11 + nx)
    tf = {k: torch.tensor(v, device="cuda") for k, v in f.items()}
    sc = dict(beta=0.2, modvis=mv, dt=2e-3, coord=coord)
    for name in A.ROUTINES:
        want = orc.coef(name, nx, ny, sc, f)
        got = A.call(name, nx, ny, sc, tf)
        torch.cuda.synchronize()
        for k in want:
            assert np.array_equal(got[k].cpu().numpy(), want[k]), (name, k, case)


def test_assembly_rejects_mixed_memory(gpu):
    import torch

    from mach2d import assembly as A
    from mach2d._lib import SolverError
    f = _fields(12, 8, 1, False, seed=3)
    f["rp"] = torch.tensor(f["rp"], device="cuda")   # one device array among host ones
    with pytest.raises(SolverError):
        A.call("p_coefficients", 12, 8, dict(dt=1e-3), f)


@pytest.mark.parametrize("tag", ["ns_axi", "euler", "ns_planar", "cart", "axis"])
def test_assembly_vs_reference_fixtures(gpu, tag):
    """Every routine on the GPU against the reference's own outputs
    (tests/golden/coef.npz), untouched points included."""
    from conftest import load_golden
    from mach2d import assembly as A
    g = load_golden("coef.npz")
    nx, ny, coord, mv, beta, dt = g["%s_scalars" % tag]
    nx, ny = int(nx), int(ny)
    sc = dict(coord=int(coord), modvis=int(mv), beta=beta, dt=dt)
    pre = "%s_in_" % tag
    f = {k[len(pre):]: v for k, v in g.items() if k.startswith(pre)}
    for name, (_, _, outs) in A.ROUTINES.items():
        init = {k: g["%s_%s_init_%s" % (tag, name, k)].reshape(-1).copy() for k, _ in outs}
        got = A.call(name, nx, ny, sc, f, out=init)
        for k, _ in outs:
            assert np.array_equal(np.asarray(got[k]), g["%s_%s_%s" % (tag, name, k)]), (name, k)
