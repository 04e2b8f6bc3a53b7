"""GPU parity of the device-resident pressure-correction loop of the
internal-flow case (m2d_intflow_pl_loop, src/mach2d_main.f90:257-309)
against the same loop composed from the oracle's routines
(tests/pl_loop_oracle.py): every updated field bit for bit, the same number
of itp iterations and the same max|pl| / p_avg.  The loop's batches are
replayed from captured HIP graphs: replays with eager work between them
equal the directly enqueued kernels bit for bit, an exit inside a batch
turns the rest of it into no-ops, and a steady run reads the host once per
call."""
import numpy as np
import pytest

import pl_loop_oracle as PL

pytestmark = pytest.mark.gpu


def gpu_loop(gpu, f, nx, ny, **kw):
    import torch

    from mach2d import simplec
    dev = gpu.DeviceMSI(nx, ny, "fp64", ndiag=5)
    t = {k: torch.tensor(v.reshape(-1), device="cuda") for k, v in f.items()}
    itp, normpl = simplec.intflow_pl_loop(dev, t, **kw)
    out = {k: t[k].cpu().numpy() for k in PL.INOUT + ("ap", "bp")}
    out["ap"] = out["ap"].reshape(5, -1)
    dev.close()
    return out, itp, normpl


@pytest.mark.parametrize("modvis,coord", [(0, 1), (1, 1), (0, 0)])
def test_pl_loop_matches_oracle(gpu, orc, modvis, coord):
    nx, ny = 70, 38
    f = PL.loop_fields(nx, ny, seed=11 + modvis + 2 * coord, coord=coord)
    p_avg = float(np.mean(f["p"]))
    kw = dict(modvis=modvis, dt=2e-3, beta=0.3, p_avg=p_avg, tolm=0.0, itpmax=4, nitm_p=20,
              tol_p=1e-2)
    want, itp0, n0 = PL.run(orc, f, nx, ny, kw["modvis"], kw["dt"], kw["beta"], p_avg,
                            kw["tolm"], kw["itpmax"], kw["nitm_p"], kw["tol_p"])
    got, itp, normpl = gpu_loop(gpu, f, nx, ny, **kw)
    assert itp == itp0 == 4 and normpl == n0
    for k in want:
        assert np.array_equal(got[k], want[k]), k


def test_pl_loop_exit_test(gpu, orc):
    """A tolm below every earlier itp's max|pl| / p_avg and above the k-th
    stops the loop after iteration k, on the GPU as in the oracle."""
    nx, ny = 40, 22
    f = PL.loop_fields(nx, ny, seed=3)
    p_avg = float(np.mean(f["p"]))
    args = (0, 2e-3, 0.3, p_avg)
    seq = [PL.run(orc, f, nx, ny, *args, 0.0, k, 20, 1e-2)[2] for k in range(1, 6)]
    k = next(k for k in range(2, 6) if seq[k - 1] < min(seq[:k - 1]))
    tolm = 0.5 * (seq[k - 1] + min(seq[:k - 1]))
    want, itp0, _ = PL.run(orc, f, nx, ny, *args, tolm, 10, 20, 1e-2)
    got, itp, _ = gpu_loop(gpu, f, nx, ny, modvis=0, dt=2e-3, beta=0.3, p_avg=p_avg, tolm=tolm,
                           itpmax=10, nitm_p=20, tol_p=1e-2)
    assert itp == itp0 == k
    for key in want:
        assert np.array_equal(got[key], want[key]), key


def _tensors(f):
    import torch
    return {k: torch.tensor(v.reshape(-1), device="cuda") for k, v in f.items()}


def test_pl_loop_graph_replays_equal_direct_launches(gpu):
    """Three calls in a row on one handle replaying captured batches, and on
    another enqueueing the same kernels directly, with eager torch work on
    the fields and a new dt between the calls: every field equal bit for bit
    after every call.  The first call (no history) reads the host after each
    of its itpmax one-iteration batches; the later ones replay one batch of
    itpmax iterations and read once."""
    import torch

    from mach2d import simplec
    nx, ny = 70, 38
    f = PL.loop_fields(nx, ny, seed=5)
    p_avg = float(np.mean(f["p"]))
    kw = dict(modvis=0, beta=0.3, p_avg=p_avg, tolm=0.0, itpmax=3, nitm_p=20, tol_p=1e-2)
    devs = [gpu.DeviceMSI(nx, ny, "fp64", ndiag=5) for _ in range(2)]
    simplec.set_pl_graph(devs[1], 0)
    ts = [_tensors(f), _tensors(f)]
    for call, dt in enumerate((2e-3, 1.7e-3, 2.3e-3)):
        res = []
        for dev, t in zip(devs, ts):
            res.append(simplec.intflow_pl_loop(dev, t, dt=dt, **kw))
            torch.cuda.synchronize()
        assert res[0] == res[1] and res[0][0] == 3, res
        for k in PL.INOUT + ("ap", "bp"):
            assert torch.equal(ts[0][k], ts[1][k]), (call, k)
        reads, batches = simplec.pl_loop_info(devs[0])
        assert (reads, batches) == ((3, 3) if call == 0 else (1, 1)), (call, reads, batches)
        for t in ts:  # eager work between the replays
            t["p"].mul_(1.0 + 1e-3)
            t["ro"].mul_(1.0 - 1e-3)
        torch.cuda.synchronize()
    for d in devs:
        d.close()


def test_pl_loop_exit_inside_a_batch(gpu, orc):
    """After a call that ran 6 iterations, the next call replays a batch of 6;
    its exit test fires at iteration k < 6 and the rest of the batch must
    leave every field as the oracle's loop leaves it at k."""
    from mach2d import simplec
    nx, ny = 40, 22
    f = PL.loop_fields(nx, ny, seed=3)
    p_avg = float(np.mean(f["p"]))
    args = (0, 2e-3, 0.3, p_avg)
    seq = [PL.run(orc, f, nx, ny, *args, 0.0, k, 20, 1e-2)[2] for k in range(1, 6)]
    k = next(k for k in range(2, 6) if seq[k - 1] < min(seq[:k - 1]))
    tolm = 0.5 * (seq[k - 1] + min(seq[:k - 1]))
    want, itp0, n0 = PL.run(orc, f, nx, ny, *args, tolm, 10, 20, 1e-2)
    dev = gpu.DeviceMSI(nx, ny, "fp64", ndiag=5)
    kw = dict(modvis=0, dt=2e-3, beta=0.3, p_avg=p_avg, itpmax=10, nitm_p=20, tol_p=1e-2)
    assert simplec.intflow_pl_loop(dev, _tensors(f), tolm=0.0, **dict(kw, itpmax=6))[0] == 6
    t = _tensors(f)
    itp, normpl = simplec.intflow_pl_loop(dev, t, tolm=tolm, **kw)
    assert simplec.pl_loop_info(dev) == (1, 1)
    assert itp == itp0 == k and normpl == n0
    for key in want:
        got = t[key].cpu().numpy()
        assert np.array_equal(got.reshape(want[key].shape), want[key]), key
    dev.close()


@pytest.mark.parametrize("cap,graph", [(1, 1), (19, 1), (1, 0)])
def test_pl_loop_short_solve_capture_reruns_same_bits(gpu, orc, cap, graph):
    """Each p' solve enqueued with `cap` of its nitm_p = 20 iterations
    (m2d_set_pl_solve_cap): these solves take 18-20 iterations, so the
    first one runs out, the loop ends early and the call runs again from
    the saved fields with all 20 enqueued.  Every field, itp and
    max|pl| / p_avg equal the oracle's loop bit for bit, on the call that
    captures the batches and on the one that replays them (graph 1), and
    with the kernels enqueued directly (graph 0)."""
    from mach2d import simplec
    nx, ny = 70, 38
    f = PL.loop_fields(nx, ny, seed=11)
    p_avg = float(np.mean(f["p"]))
    kw = dict(modvis=0, dt=2e-3, beta=0.3, p_avg=p_avg, tolm=0.0, itpmax=4, nitm_p=20,
              tol_p=1e-2)
    want, itp0, n0 = PL.run(orc, f, nx, ny, kw["modvis"], kw["dt"], kw["beta"], p_avg,
                            kw["tolm"], kw["itpmax"], kw["nitm_p"], kw["tol_p"])
    dev = gpu.DeviceMSI(nx, ny, "fp64", ndiag=5)
    simplec.set_pl_solve_cap(dev, cap)
    simplec.set_pl_graph(dev, graph)
    for call in range(2):
        t = _tensors(f)
        itp, normpl = simplec.intflow_pl_loop(dev, t, **kw)
        assert simplec.pl_loop_cap_info(dev) == (cap, 1), call
        assert itp == itp0 and normpl == n0, call
        for key in want:
            got = t[key].cpu().numpy()
            assert np.array_equal(got.reshape(want[key].shape), want[key]), (call, key)
    dev.close()
