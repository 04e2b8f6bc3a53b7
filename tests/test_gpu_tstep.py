"""GPU parity of the Scarborough time-step statistics
(src/tstepper/mod_tstepper_scarborough.f90) against the oracle, which the
CPU tests pin to the compiled reference: bit for bit, host and device
arrays."""
import numpy as np
import pytest

from test_oracle_golden import _scarb_inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny", [(3, 3), (70, 33), (258, 130), (1026, 514)])
def test_mean_cc_and_dt_match_oracle(gpu, orc, nx, ny):
    import torch

    from mach2d import tstep
    au, av, at, ap = _scarb_inputs(nx, ny, nx + ny)
    for a in (au, av, at, ap):
        want = orc.mean_cc(a, nx, ny)
        assert tstep.mean_cc(a, nx, ny) == want
        assert tstep.mean_cc(torch.tensor(a.reshape(-1), device="cuda"), nx, ny) == want
    rng = np.random.default_rng(2)
    J, u, v = (rng.uniform(0.5, 2.0, nx * ny), rng.uniform(-5, 5, nx * ny),
               rng.uniform(-5, 5, nx * ny))
    assert tstep.dt0(J, u, v, nx, ny) == orc.dt0_scarborough(J, u, v, nx, ny)
    c = max(orc.mean_cc(a, nx, ny) for a in (au, av, at, ap))
    dev = [torch.tensor(a.reshape(-1), device="cuda") for a in (au, av, at, ap)]
    for mincc, maxcc in ((2 * c, 4 * c), (0.5 * c, 2 * c), (0.1 * c, 0.5 * c)):
        want = orc.dt_scarborough(1.5e-3, au, av, at, ap, nx, ny, 0.25, mincc, maxcc)
        assert tstep.dt(1.5e-3, au, av, at, ap, nx, ny, 0.25, mincc, maxcc) == want
        assert tstep.dt(1.5e-3, *dev, nx, ny, 0.25, mincc, maxcc) == want
