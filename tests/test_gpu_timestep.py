"""GPU parity of one internal-flow time step (m2d_intflow_time_step,
src/mach2d_main.f90:94-404) and of the routines it adds (get_average,
norm_l1_5d) against the oracle composition, which the CPU tests pin bit for
bit to the same step composed from the compiled reference's routines.

Bar: get_average bit for bit; norm_l1_5d (a fixed-tree sum) within 1e-13
relative.  The time step: every field, dt, max|pl|/p_avg, the mass residual
and the pressure-correction iteration counts bit for bit after each of
three steps (the inflow pressure's real power is taken from the C library's
pow, as in the reference); the relative L1 norms (fixed-tree sums) within
1e-12."""
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


@pytest.mark.parametrize("modvis,solver9d,nx,ny,mach,dt0",
                         [(0, 0, 22, 12, None, 2e-5), (0, 1, 34, 18, None, 2e-5),
                          (1, 0, 30, 14, None, 2e-5), (0, 0, 258, 130, None, 2e-5),
                          (0, 0, 258, 130, 0.2, 5e-6), (0, 1, 258, 130, 0.2, 5e-6)])
def test_time_step_matches_oracle(gpu, orc, modvis, solver9d, nx, ny, mach, dt0):
    """Three time steps: every field, the time step and the p' iteration
    counts bit for bit; the relative L1 norms within 1e-12 (fixed-tree sums
    on the GPU, left to right in the reference).  mach = 0.2: the 258x130
    nozzle at inlet Mach 0.2 (supersonic past the throat), u and v by MSI or
    by TDMA."""
    import torch

    from mach2d import simplec
    f = TS.step_fields(nx, ny, seed=nx + modvis, mach_in=mach)
    prm = dict(TS.PARAMS, modvis=modvis, solver9d=solver9d)
    d = {k: torch.tensor(np.ascontiguousarray(v).reshape(-1), device="cuda")
         for k, v in f.items()}
    dev9 = gpu.DeviceMSI(nx, ny) if solver9d == 0 else None
    dev5 = gpu.DeviceMSI(nx, ny, ndiag=5)
    torch.cuda.synchronize()   # the copies above run on torch's stream
    s, dt_o, dt_g = f, dt0, dt0
    for it in range(3):
        s, dt_o, ro = TS.step(orc, s, nx, ny, prm, dt_o)
        dt_g, rg = simplec.intflow_time_step(dev9, dev5, d, prm, dt_g)
        assert rg["itp"] == ro["itp"], (it, rg, ro)
        assert dt_g == dt_o, (it, dt_g, dt_o)
        assert rg["normpl"] == ro["normpl"] and rg["rmass"] == ro["rmass"], (it, rg, ro)
        assert rg["norm"] == pytest.approx(ro["norm"], rel=1e-12, abs=1e-300), it
        for k in simplec.TS_STATE + simplec.TS_LEVEL + ("pin", "plin", "Tin", "au", "av",
                                                         "ap", "bp"):
            got = d[k].cpu().numpy().reshape(np.asarray(s[k]).shape)
            assert np.array_equal(got, s[k]), (it, k, np.abs(got - s[k]).max())
    for h in (dev9, dev5):
        if h is not None:
            h.close()


def test_time_step_pcr_tdma_within_1e9_of_oracle(gpu, orc):
    """u and v by TDMA with the lines solved by parallel cyclic reduction
    (m2d_set_tdma_mode(1)) on the 258x130 nozzle at inlet Mach 0.2: three
    steps, the same p' iteration counts, every field, dt and the norms
    within 1e-9 of the reference's trajectory (largest difference relative
    to the field's largest magnitude); the line solutions differ from the
    recurrence's by rounding only."""
    import torch

    from mach2d import simplec
    nx, ny, dt0 = 258, 130, 5e-6
    f = TS.step_fields(nx, ny, seed=nx, mach_in=0.2)
    prm = dict(TS.PARAMS, modvis=0, solver9d=1)
    d = {k: torch.tensor(np.ascontiguousarray(v).reshape(-1), device="cuda")
         for k, v in f.items()}
    dev5 = gpu.DeviceMSI(nx, ny, ndiag=5)
    torch.cuda.synchronize()
    s, dt_o, dt_g = f, dt0, dt0
    gpu.set_tdma_mode(1)
    try:
        for it in range(3):
            s, dt_o, ro = TS.step(orc, s, nx, ny, prm, dt_o)
            dt_g, rg = simplec.intflow_time_step(None, dev5, d, prm, dt_g)
            assert rg["itp"] == ro["itp"], (it, rg, ro)
            assert dt_g == pytest.approx(dt_o, rel=1e-9), (it, dt_g, dt_o)
            assert rg["normpl"] == pytest.approx(ro["normpl"], rel=1e-9), it
            for k in simplec.TS_STATE + ("pin", "plin", "Tin"):
                got = d[k].cpu().numpy().reshape(np.asarray(s[k]).shape)
                err = np.abs(got - s[k]).max() / max(np.abs(s[k]).max(), 1e-300)
                assert err <= 1e-9, (it, k, err)
    finally:
        gpu.set_tdma_mode(0)
        dev5.close()


@pytest.mark.parametrize("cap", [0, 1])
def test_time_step_p_solve_capture_same_bits(gpu, orc, cap):
    """The 258x130 nozzle at inlet Mach 0.2, u and v by TDMA, three steps
    with the p' solves' enqueued iterations trimmed (cap 0: from the
    previous call's longest solve, fewer than nitm_p = 20 after the first
    step, no rerun needed) or cut to one (cap 1: a solve that needs more
    makes its loop run again with all 20).  Every field, dt and the p'
    iteration counts bit for bit the reference's."""
    import torch

    from mach2d import simplec
    nx, ny, dt0 = 258, 130, 5e-6
    f = TS.step_fields(nx, ny, seed=nx, mach_in=0.2)
    prm = dict(TS.PARAMS, modvis=0, solver9d=1)
    d = {k: torch.tensor(np.ascontiguousarray(v).reshape(-1), device="cuda")
         for k, v in f.items()}
    dev5 = gpu.DeviceMSI(nx, ny, ndiag=5)
    simplec.set_pl_solve_cap(dev5, cap)
    torch.cuda.synchronize()
    s, dt_o, dt_g = f, dt0, dt0
    for it in range(3):
        s, dt_o, ro = TS.step(orc, s, nx, ny, prm, dt_o)
        dt_g, rg = simplec.intflow_time_step(None, dev5, d, prm, dt_g)
        used, reruns = simplec.pl_loop_cap_info(dev5)
        if cap == 0:
            assert (used < 20 and reruns == 0) if it else used == 20, (it, used, reruns)
        else:
            assert used == 1, (it, used)
        assert rg["itp"] == ro["itp"] and dt_g == dt_o, (it, rg, ro)
        assert rg["normpl"] == ro["normpl"] and rg["rmass"] == ro["rmass"], it
        for k in simplec.TS_STATE + simplec.TS_LEVEL + ("pin", "plin", "Tin"):
            got = d[k].cpu().numpy().reshape(np.asarray(s[k]).shape)
            assert np.array_equal(got, s[k]), (it, k, np.abs(got - s[k]).max())
    dev5.close()


def test_time_step_follows_replaced_tensors(gpu, orc):
    """The wrapper reuses its checked field structure while the same tensors
    come back: a field replaced by a copy between two steps is the one the
    second step updates (the same two steps on one set of tensors give the
    same fields)."""
    import torch

    from mach2d import simplec
    nx, ny = 34, 18
    f = TS.step_fields(nx, ny, seed=nx, mach_in=None)
    prm = dict(TS.PARAMS, modvis=0, solver9d=1)
    dev5 = gpu.DeviceMSI(nx, ny, ndiag=5)
    out = []
    for replace in (False, True):
        d = {k: torch.tensor(np.ascontiguousarray(v).reshape(-1), device="cuda")
             for k, v in f.items()}
        torch.cuda.synchronize()
        dt = simplec.intflow_time_step(None, dev5, d, prm, 2e-5)[0]
        if replace:
            d["u"] = d["u"].clone()
            d["p"] = d["p"].clone()
            torch.cuda.synchronize()
        simplec.intflow_time_step(None, dev5, d, prm, dt)
        out.append({k: d[k].cpu().numpy() for k in ("u", "v", "p", "T")})
    dev5.close()
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


@pytest.mark.parametrize("bad", [dict(itpmax=0), dict(nitm_p=0), dict(modvis=2),
                                 dict(itmmax=0), dict(solver9d=3)])
def test_time_step_rejects_bad_parameters_before_changing_fields(gpu, bad):
    """m2d_intflow_time_step checks every parameter at entry: a bad one is
    M2D_ERR_ARG with every field left as it was."""
    import torch

    from mach2d import _lib, simplec
    from mach2d._lib import SolverError
    nx, ny = 22, 12
    f = TS.step_fields(nx, ny, seed=3)
    d = {k: torch.tensor(np.ascontiguousarray(v).reshape(-1), device="cuda")
         for k, v in f.items()}
    before = {k: v.clone() for k, v in d.items()}
    dev9, dev5 = gpu.DeviceMSI(nx, ny), gpu.DeviceMSI(nx, ny, ndiag=5)
    torch.cuda.synchronize()
    with pytest.raises(SolverError) as e:
        simplec.intflow_time_step(dev9, dev5, d, dict(TS.PARAMS, **bad), 2e-5)
    assert e.value.code == _lib.M2D_ERR_ARG
    for k in d:
        assert torch.equal(d[k], before[k]), k
    dev9.close()
    dev5.close()
