"""GPU parity of the batched MSI (m2d_batch_*) through the C ABI.

Bar: every system of an FP64 batch ends bit for bit on the oracle's fb2d9
iterate with the oracle's iteration count, whatever the launch placement
and across repeated graph replays; FP32 sweeps inside FP64 refinement reach
the FP64 stopping rule with the solution within 1e-8 relative of the FP64
oracle."""
import numpy as np
import pytest

from mach2d import problems as P

pytestmark = pytest.mark.gpu

# uneven sizes: partial bands, one-band systems, several bands, odd widths
SIZES = [(10, 6), (66, 34), (3, 3), (130, 66), (37, 129), (200, 3), (64, 65), (98, 50),
         (5, 140), (258, 70)]


def _systems(sizes, seed0=10):
    return [P.random_f3_system(nx, ny, seed=seed0 + q) for q, (nx, ny) in enumerate(sizes)]


def _dev(systems):
    import torch
    a = [torch.tensor(s.a.reshape(-1), device="cuda") for s in systems]
    rhs = [torch.tensor(s.rhs, device="cuda") for s in systems]
    var = [torch.zeros(s.N, dtype=torch.float64, device="cuda") for s in systems]
    return a, rhs, var


def _oracle(orc, systems, nitm, gamma):
    out = []
    for s in systems:
        v, nit, _, _ = orc.solve_msi9d(s.a, s.rhs, np.zeros(s.N), s.nx, s.ny, nitm, gamma,
                                       max_iter_cap=0 if nitm > 0 else 100000)
        out.append((v, nit))
    return out


@pytest.mark.parametrize("layout", [0, 1, 2, 3, 4, 5])   # 3..5: LDS-staged sweeps
def test_batch_fp64_bitwise_vs_oracle(gpu, orc, layout):
    systems = _systems(SIZES)
    ref = _oracle(orc, systems, 40, 1e-6)
    b = gpu.BatchMSI([(s.nx, s.ny) for s in systems])
    b.set_tuning(layout, 4)
    a, rhs, var = _dev(systems)
    b.factor(a)
    rc, st = b.solve(a, rhs, var, 1e-6, 40)
    assert rc == 0
    for q, (s, (ov, onit)) in enumerate(zip(systems, ref)):
        assert st[q].nit == onit, (q, s.nx, s.ny, st[q].nit, onit)
        assert np.array_equal(var[q].cpu().numpy(), ov), (q, s.nx, s.ny)
    b.close()


def test_batch_replays_repeat_exactly(gpu):
    """Three solves through the same captured graph: identical results."""
    import torch
    systems = _systems(SIZES[:6], seed0=40)
    b = gpu.BatchMSI([(s.nx, s.ny) for s in systems])
    b.set_tuning(4, 3)   # LDS-staged sweeps, one XCD per system
    a, rhs, var = _dev(systems)
    b.factor(a)
    first = None
    for rep in range(3):
        for v in var:
            v.zero_()
        torch.cuda.synchronize()
        rc, st = b.solve(a, rhs, var, 1e-8, 25)
        assert rc == 0
        got = [v.cpu().numpy().copy() for v in var] + [[x.nit for x in st]]
        if first is None:
            first = got
        else:
            for x, y in zip(first[:-1], got[:-1]):
                assert np.array_equal(x, y), rep
            assert first[-1] == got[-1]
    b.close()


def test_batch_matches_single_handle(gpu):
    import torch
    s = P.pcorr_system(130, 66, seed=3)
    b = gpu.BatchMSI([(130, 66)])
    a, rhs, var = _dev([s])
    b.factor(a)
    rc, st = b.solve(a, rhs, var, 1e-10, 500)
    d = gpu.DeviceMSI(130, 66)
    v1 = torch.zeros(s.N, dtype=torch.float64, device="cuda")
    d.factor(a[0])
    st1 = d.solve(a[0], rhs[0], v1, 1e-10, 500)
    assert rc == 0 and st1.status == 0 and st[0].nit == st1.nit
    assert np.array_equal(var[0].cpu().numpy(), v1.cpu().numpy())
    b.close()
    d.close()


def test_batch_fp32_refinement_accuracy(gpu, orc):
    sizes = [(66, 34), (130, 66), (98, 50), (37, 129)]
    systems = [P.pcorr_system(nx, ny, seed=20 + q) for q, (nx, ny) in enumerate(sizes)]
    ref = _oracle(orc, systems, 5000, 1e-12)  # the FP64 solution, tightly converged
    b = gpu.BatchMSI(sizes, precision="fp32_ir")
    b.set_inner(1e-4, 200)
    a, rhs, var = _dev(systems)
    b.factor(a)
    rc, st = b.solve(a, rhs, var, 1e-10, 50)
    assert rc == 0
    for q, (s, (ov, _)) in enumerate(zip(systems, ref)):
        x = var[q].cpu().numpy()
        r0 = np.linalg.norm(s.rhs)                       # zero guess: r0 = ||rhs||
        assert np.linalg.norm(orc.residual9(s.a, x, s.rhs, s.nx, s.ny)) < 1e-10 * r0 * (1 + 1e-6)
        err = np.linalg.norm(x - ov) / np.linalg.norm(ov)
        assert err < 1e-8, (q, err)                      # tolerance: 1e-8 relative
        assert st[q].rn < st[q].gammarp and st[q].outer >= 1, q
    b.close()


def test_batch_autotune(gpu):
    systems = _systems(SIZES, seed0=60)
    b = gpu.BatchMSI([(s.nx, s.ny) for s in systems])
    a, rhs, var = _dev(systems)
    b.factor(a)
    layout, ms = b.tune(a, rhs, iters=3)
    assert layout in range(6) and len(ms) == 6 and all(t > 0 for t in ms)
    assert b.get_tuning()[0] == layout
    # the caller's var is untouched, and the tuned batch still solves exactly
    assert all(float(v.abs().max()) == 0.0 for v in var)
    b.close()
