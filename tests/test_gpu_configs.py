"""GPU parity at the full sizes of the three large workloads, through the C
ABI, against the oracle (a C restatement of msi2d9-v03.f90 pinned bit for
bit to the compiled reference):

* the 1026x514 wall-resolved p' system (near-wall aspect 1e4, seeded
  Gaussian right-hand side, zero guess, gamma = 1e-10): plain FP64 equal to
  the oracle bit for bit with the same iteration count; FP32 sweeps inside
  FP64 refinement within 1e-8 relative of the FP64 oracle, with the FP64
  residual below 1e-10 ||b|| (zero guess: r0 = ||b||, msi2d9-v03.f90:133,
  258-259);
* the 256-system uneven batch (real sizes 64s x 32s, s log-uniform on
  [1, 8]): every system of the FP64 batch equal to the oracle bit for bit
  with the oracle's iteration count, and the FP32-refined batch within 1e-8;
* the shipped U-Net weights on the 1026x514 system: the guessed solve ends
  within 1e-8 of the oracle, and the guess is kept only when its residual
  norm is below the zero guess's (otherwise the solve is the zero-guess one,
  bit for bit).

Tolerances: bitwise for FP64 iterates; 1e-8 relative (2-norm) for results
that leave the reference's arithmetic (FP32 sweeps, a nonzero guess)."""
import concurrent.futures as cf
import os

import numpy as np
import pytest

from conftest import load_golden
from mach2d import problems as P

pytestmark = pytest.mark.gpu

GAMMA = 1e-10
NITM = 20000


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def cfg3(orc):
    c = P.CFG3
    s = P.pcorr_system(c["nx"], c["ny"], seed=c["seed"], aspect=c["aspect"], rhs=c["rhs"])
    print("cfg3: %dx%d, oracle FP64 solve to gamma %g" % (s.nx, s.ny, GAMMA), flush=True)
    ov, onit, or0, ohist = orc.solve_msi9d(s.a, s.rhs, np.zeros(s.N), s.nx, s.ny, NITM, GAMMA)
    return s, ov, onit, or0, ohist


@pytest.fixture(scope="module")
def cfg4(orc):
    systems = P.cfg4_systems(256)
    print("cfg4: %d systems, N = %d, oracle FP64 solves on %d threads" % (
        len(systems), sum(s.N for s in systems), _threads()), flush=True)

    def one(s):
        v, nit, r0, _ = orc.solve_msi9d(s.a, s.rhs, np.zeros(s.N), s.nx, s.ny, NITM, GAMMA)
        return v, nit, r0

    with cf.ThreadPoolExecutor(max_workers=_threads()) as ex:
        ref = list(ex.map(one, systems))
    return systems, ref


def _t(x):
    import torch
    return torch.tensor(np.ascontiguousarray(x).reshape(-1), device="cuda")


def test_cfg3_fp64_bitwise_vs_oracle(gpu, cfg3):
    import torch
    s, ov, onit, or0, ohist = cfg3
    dev = gpu.DeviceMSI(s.nx, s.ny)
    ta, tr = _t(s.a), _t(s.rhs)
    tv = torch.zeros(s.N, dtype=torch.float64, device="cuda")
    hist = torch.zeros(onit + 8, dtype=torch.float64, device="cuda")
    dev.factor(ta)
    st = dev.solve(ta, tr, tv, GAMMA, NITM, hist=hist)
    assert st.status == 0 and st.nit == onit, (st.as_dict(), onit)
    assert np.array_equal(tv.cpu().numpy(), ov)
    assert abs(st.r0 - or0) <= 1e-12 * or0
    # residual history: the device sums r.r in a fixed tree, the oracle left
    # to right, so the norms agree to the last bits only
    assert np.allclose(hist.cpu().numpy()[:onit], ohist, rtol=1e-10, atol=0)
    dev.close()


def test_cfg3_fp32_refinement_vs_oracle(gpu, orc, cfg3):
    import torch
    s, ov, onit, or0, _ = cfg3
    dev = gpu.DeviceMSI(s.nx, s.ny, precision="fp32_ir")
    ta, tr = _t(s.a), _t(s.rhs)
    tv = torch.zeros(s.N, dtype=torch.float64, device="cuda")
    dev.factor(ta)
    st = dev.solve(ta, tr, tv, GAMMA, NITM)
    assert st.status == 0 and st.outer >= 1, st.as_dict()
    x = tv.cpu().numpy()
    rb = np.linalg.norm(s.rhs)
    assert np.linalg.norm(orc.residual9(s.a, x, s.rhs, s.nx, s.ny)) < GAMMA * rb * (1 + 1e-9)
    err = np.linalg.norm(x - ov) / np.linalg.norm(ov)
    print("cfg3 fp32_ir: %d inner / %d outer iterations, error vs FP64 oracle %.3g" % (
        st.nit, st.outer, err), flush=True)
    assert err < 1e-8, err
    dev.close()


def test_cfg4_batch_fp64_bitwise_vs_oracle(gpu, cfg4):
    import torch
    systems, ref = cfg4
    b = gpu.BatchMSI([(s.nx, s.ny) for s in systems])
    a = [_t(s.a) for s in systems]
    rhs = [_t(s.rhs) for s in systems]
    var = [torch.zeros(s.N, dtype=torch.float64, device="cuda") for s in systems]
    b.factor(a)
    rc, st = b.solve(a, rhs, var, GAMMA, NITM)
    assert rc == 0
    bad = []
    for q, (s, (ov, onit, _)) in enumerate(zip(systems, ref)):
        if st[q].status != 0 or st[q].nit != onit or not np.array_equal(var[q].cpu().numpy(), ov):
            bad.append((q, s.nx, s.ny, st[q].status, st[q].nit, onit))
    assert not bad, bad[:8]
    b.close()


def test_cfg4_batch_fp32_refinement_vs_oracle(gpu, orc, cfg4):
    import torch
    systems, ref = cfg4
    b = gpu.BatchMSI([(s.nx, s.ny) for s in systems], precision="fp32_ir")
    a = [_t(s.a) for s in systems]
    rhs = [_t(s.rhs) for s in systems]
    var = [torch.zeros(s.N, dtype=torch.float64, device="cuda") for s in systems]
    b.factor(a)
    rc, st = b.solve(a, rhs, var, GAMMA, NITM)
    assert rc == 0
    worst = 0.0
    for q, (s, (ov, _, _)) in enumerate(zip(systems, ref)):
        x = var[q].cpu().numpy()
        rb = np.linalg.norm(s.rhs)
        assert st[q].status == 0 and st[q].rn < st[q].gammarp, (q, st[q].as_dict())
        assert np.linalg.norm(orc.residual9(s.a, x, s.rhs, s.nx, s.ny)) < GAMMA * rb * (1 + 1e-6)
        worst = max(worst, np.linalg.norm(x - ov) / np.linalg.norm(ov))
    print("cfg4 fp32_ir: worst error vs FP64 oracle %.3g" % worst, flush=True)
    assert worst < 1e-8, worst
    b.close()


def test_cfg5_shipped_weights_on_cfg3(gpu, cfg3):
    """The trained network's guess on the 1026x514 system, judged by the
    residual norm against the zero guess, then FP64 MSI to gamma ||b||."""
    import torch

    from mach2d import learned as Lr
    from mach2d import unet as U
    s, ov, onit, _, _ = cfg3
    net = U.UNet()
    net.load(Lr.default_weights())
    dev = gpu.DeviceMSI(s.nx, s.ny)
    ta, tr = _t(s.a), _t(s.rhs)
    dev.factor(ta)
    tv = torch.zeros(s.N, dtype=torch.float64, device="cuda")
    st, used, rg, rb = Lr.guessed_solve(dev, net, ta, tr, tv, s.nx, s.ny, GAMMA, NITM)
    assert st.status == 0
    x = tv.cpu().numpy()
    print("cfg5: guess used %s, ||r(guess)|| / ||b|| = %.3g, %d iterations (zero guess %d)" % (
        used, rg / rb, st.nit, onit), flush=True)
    if used:
        assert rg < rb
        assert st.rn < GAMMA * rb * (1 + 1e-9)
        err = np.linalg.norm(x - ov) / np.linalg.norm(ov)
        assert err < 1e-8, err
    else:   # the fallback: plain MSI from zero, the oracle's iterate
        assert rg == rb and st.nit == onit
        assert np.array_equal(x, ov)
    dev.close()


def test_worse_than_zero_guess_falls_back_to_zero(gpu):
    """A guess whose residual exceeds ||b|| is replaced by the zero guess,
    so the solve is the reference's zero-guess solve bit for bit (130x66
    system, iterate written by the compiled reference)."""
    import torch

    from mach2d import learned as Lr
    g = load_golden("cfg1.npz")
    nx, ny = (int(v) for v in g["shape"])
    dev = gpu.DeviceMSI(nx, ny)
    ta, tr = _t(g["a"]), _t(g["rhs"])
    dev.factor(ta)
    bad = torch.tensor(1e3 * np.random.default_rng(0).standard_normal(nx * ny), device="cuda")
    used, rg, rb = Lr.judge_guess(dev, ta, tr, bad)
    assert not used and rg == rb and float(bad.abs().max()) == 0.0
    st = dev.solve(ta, tr, bad, 1e-10 * rb / rg, 500)
    assert st.status == 0
    assert np.array_equal(bad.cpu().numpy(), g["var_full"])
    # a guess better than zero is kept: half the converged solution
    good = torch.tensor(0.5 * g["var_full"], device="cuda")
    used, rg, rb = Lr.judge_guess(dev, ta, tr, good)
    assert used and rg < rb and float(good.abs().max()) > 0.0
    dev.close()


def test_cfg5_shipped_weights_on_heldout_simplec_systems(gpu):
    """The trained network on p' systems of a SIMPLEC run it was not trained
    on (258x130 nozzle, variant seed 99): the guessed solve stops at
    gamma ||b||, agrees with the zero-guess solution to 1e-8 relative, and
    keeps the guess only where its residual is below ||b||."""
    from mach2d import learned as Lr
    from mach2d import unet as U
    net = U.UNet()
    net.load(Lr.default_weights())
    systems = Lr.simplec_systems(258, 130, 99, 60, 4)
    for z, g, used, rr, rel in Lr.compare_guess(net, systems, 258, 130):
        assert rel < 1e-8, rel
        assert (rr < 1.0) if used else (rr == 1.0 and g == z)
