"""GPU parity of the coefficient assembly (m2d_get_*, mach2d_coef.f90)
against the oracle, through the C ABI, with host and with device arrays.

Bar: bit for bit (the kernels keep the reference's operation order without
contraction); points the reference does not write keep their values.  The
one exception is the inflow pressure of intflow_time_cycle_update, a real
power x**y: the device pow and the host libm pow may differ in the last bit,
so pin, plin (= pin - pina) and the pressures set from pin are held to 4 ulp
of the inflow pressure."""
import os
import sys

import numpy as np
import pytest

from mach2d import problems as P

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))

pytestmark = pytest.mark.gpu

CASES = [  # nx, ny, coord, modvis, cartesian
    (12, 8, 1, 1, False), (12, 8, 0, 0, False), (34, 18, 1, 1, False), (10, 7, 1, 1, True),
    (300, 70, 1, 1, False), (3, 3, 1, 1, False), (258, 3, 0, 1, False)]


def _fields(nx, ny, coord, cart, seed):
    """Assembly inputs plus those of the SIMPLEC glue (u, v systems from the
    assembly, unit diagonal on fictitious rows; seeded sources and face
    values)."""
    import oracle as O
    f = P.flow_fields(nx, ny, seed=seed, coord=coord, cartesian=cart)
    rng = np.random.default_rng(seed + 1)
    sc = dict(modvis=1, dt=2e-3, coord=coord)
    for key, name in (("au", "u_coefficients"), ("av", "v_coefficients")):
        a = O.coef(name, nx, ny, sc, f)["a"]
        a[4][a[4] == 0.0] = 1.0
        f[key] = a
    for k in ("cup", "cvp", "sup", "svp", "uea", "vea", "una", "vna", "due", "dve", "dun",
              "dvn"):
        f[k] = rng.uniform(-1.0, 1.0, size=nx * ny)
    return f


STEP = dict(Cpref=1004.5, Href=837083.3, gamma=1.4, P0=1038350.4, T0=833.33333333)
POW_OUTPUTS = {("intflow_time_cycle_update", k) for k in ("pin", "plin", "p")}


def _size(m, nx, ny):
    return ny if m == "ny" else nx if m == "nx" else m * nx * ny if m else 1


def _same(got, want, name=None, k=None, ref_scale=None):
    if isinstance(want, float):
        return float(got) == want
    got, want = np.asarray(got).reshape(-1), np.asarray(want).reshape(-1)
    if (name, k) in POW_OUTPUTS:
        return np.abs(got - want).max() <= 4 * np.finfo(float).eps * ref_scale
    return np.array_equal(got, want)


def _skip(name, nx):
    # the north slip extrapolation reads the third column (nx >= 4)
    return name == "intflow_extrapolate_u_v_to_fictitious" and nx < 4


@pytest.mark.parametrize("case", CASES)
def test_assembly_host_arrays_bitwise(gpu, orc, case):
    from mach2d import assembly as A
    nx, ny, coord, mv, cart = case
    f = _fields(nx, ny, coord, cart, seed=nx * ny + coord)
    sc = dict(beta=0.3, modvis=mv, dt=1e-3, coord=coord, Rg=287.0,
              Twall=300.0 if coord else -1.0, **STEP)
    rng = np.random.default_rng(7)
    for name, (_, _, outs) in A.ROUTINES.items():
        if _skip(name, nx):
            continue
        init = {k: rng.uniform(-1, 1, size=_size(m, nx, ny)) for k, m in outs}
        want = orc.coef(name, nx, ny, sc, f, out={k: v.copy() for k, v in init.items()})
        got = A.call(name, nx, ny, sc, f, out={k: v.copy() for k, v in init.items()})
        for k in want:
            assert _same(got[k], want[k], name, k, STEP["P0"]), (name, k, case)


@pytest.mark.parametrize("case", CASES[:4])
def test_assembly_device_arrays_bitwise(gpu, orc, case):
    import torch

    from mach2d import assembly as A
    nx, ny, coord, mv, cart = case
    f = _fields(nx, ny, coord, cart, seed=11 + nx)
    tf = {k: torch.tensor(v.reshape(-1), device="cuda") for k, v in f.items()}
    sc = dict(beta=0.2, modvis=mv, dt=2e-3, coord=coord, Rg=287.0,
              Twall=-1.0 if coord else 310.0, **STEP)
    for name, (_, _, outs) in A.ROUTINES.items():
        if _skip(name, nx):
            continue
        want = orc.coef(name, nx, ny, sc, f)
        if name in ("boundary_faces_flags", "intflow_boundary_simplec"):
            # no input fields: the outputs choose the memory
            out = {k: torch.zeros(nx * ny, dtype=torch.float64, device="cuda") for k, _ in outs}
            got = A.call(name, nx, ny, sc, tf, out=out)
        else:
            got = A.call(name, nx, ny, sc, tf)
        torch.cuda.synchronize()
        for k in want:
            g = got[k] if isinstance(got[k], float) else got[k].cpu().numpy()
            assert _same(g, want[k], name, k, STEP["P0"]), (name, k, case)


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
    sc = dict(coord=int(coord), modvis=int(mv), beta=beta, dt=dt, Rg=287.0,
              Twall=float(g["%s_twall" % tag][0]), **STEP)
    pre = "%s_in_" % tag
    f = {k[len(pre):]: v for k, v in g.items() if k.startswith(pre)}
    for name, (_, _, outs) in A.ROUTINES.items():
        init = {k: g["%s_%s_init_%s" % (tag, name, k)].reshape(-1).copy() for k, _ in outs}
        got = A.call(name, nx, ny, sc, f, out=init)
        for k, m in outs:
            want = g["%s_%s_%s" % (tag, name, k)]
            if m == 0:
                want = float(np.asarray(want).reshape(-1)[0])
            assert _same(got[k], want if m == 0 else want.reshape(np.asarray(got[k]).shape),
                         name, k, STEP["P0"]), (name, k)
