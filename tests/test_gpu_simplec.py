"""GPU parity of the device-resident pressure-correction loop of the
internal-flow case (m2d_intflow_pl_loop, src/mach2d_main.f90:257-309)
against the same loop composed from the oracle's routines
(tests/pl_loop_oracle.py): every updated field bit for bit, the same number
of itp iterations and the same max|pl| / p_avg."""
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
