"""GPU parity of one internal-flow time step (m2d_intflow_time_step,
src/mach2d_main.f90:94-404) and of the routines it adds (get_average,
norm_l1_5d) against the oracle composition, which the CPU tests pin bit for
bit to the same step composed from the compiled reference's routines.

Bar: get_average bit for bit; norm_l1_5d (a fixed-tree sum) within 1e-13
relative.  The time step: every field within 1e-10 relative of the oracle
after three steps, the same pressure-correction iteration counts; the only
source of difference is the inflow pressure's real power (device pow
against host libm, at most an ulp or two), which the solves carry forward."""
import numpy as np
import pytest

import time_step_oracle as TS

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny", [(3, 3), (22, 12), (258, 130), (1026, 514)])
def test_get_average_and_norm5(gpu, orc, nx, ny):
    import torch

    from mach2d import assembly as A
    rng = np.random.default_rng(nx)
    F = rng.uniform(9e4, 1.1e5, nx * ny)
    want = orc.get_average(nx, ny, F)
    assert A.get_average(nx, ny, F) == want
    assert A.get_average(nx, ny, torch.tensor(F, device="cuda")) == want
    a5 = rng.uniform(-1, 1, (5, nx * ny))
    var, b = rng.uniform(-1, 1, nx * ny), rng.uniform(-1, 1, nx * ny)
    w = orc.norm_l1_5d(nx, ny, var, b, a5, 0.5)
    assert gpu.norm_l1_5d(nx, ny, var, b, a5, 0.5) == pytest.approx(w, rel=1e-13)


def _rel(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    return d / max(np.abs(np.asarray(b)).max(), 1e-300)


@pytest.mark.parametrize("modvis,solver9d,nx,ny", [(0, 0, 22, 12), (0, 1, 34, 18),
                                                   (1, 0, 30, 14), (0, 0, 258, 130)])
def test_time_step_matches_oracle(gpu, orc, modvis, solver9d, nx, ny):
    import torch

    from mach2d import simplec
    f = TS.step_fields(nx, ny, seed=nx + modvis)
    prm = dict(TS.PARAMS, modvis=modvis, solver9d=solver9d)
    d = {k: torch.tensor(np.ascontiguousarray(v).reshape(-1), device="cuda")
         for k, v in f.items()}
    dev9 = gpu.DeviceMSI(nx, ny) if solver9d == 0 else None
    dev5 = gpu.DeviceMSI(nx, ny, ndiag=5)
    torch.cuda.synchronize()   # the copies above run on torch's stream
    s, dt_o, dt_g = f, 2e-5, 2e-5
    for it in range(3):
        s, dt_o, ro = TS.step(orc, s, nx, ny, prm, dt_o)
        dt_g, rg = simplec.intflow_time_step(dev9, dev5, d, prm, dt_g)
        assert rg["itp"] == ro["itp"], (it, rg, ro)
        assert dt_g == pytest.approx(dt_o, rel=1e-12)
        for k in ("norm", "rmass", "normpl"):
            assert rg[k] == pytest.approx(ro[k], rel=1e-9, abs=1e-300), (it, k)
    for k in simplec.TS_STATE + simplec.TS_LEVEL + ("pin", "plin", "Tin", "au", "av", "ap",
                                                     "bp"):
        got = d[k].cpu().numpy().reshape(np.asarray(s[k]).shape)
        assert _rel(got, s[k]) <= 1e-10, (k, _rel(got, s[k]))
    for h in (dev9, dev5):
        if h is not None:
            h.close()
