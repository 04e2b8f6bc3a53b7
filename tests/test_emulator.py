"""The gfx950 wavefront kernels' schedule, emulated on the CPU
(tests/native/wavefront_emu.cpp), against the oracle, bit for bit.

The emulator runs the kernels' band/lane/step loops with their shuffles,
history registers, boundary-row reads and progress publications under
random band interleavings; it fails on any read of a value its producer had
not published, on a deadlock and on an out-of-range index.  Equality with
the oracle shows that the general 2-D stencil formulas reproduce the
reference's flat-index special rows when outside couplings are zero."""
import numpy as np
import pytest

from mach2d import problems as P

SHAPES = [(3, 3), (4, 3), (3, 9), (10, 6), (20, 70), (7, 200), (130, 66), (66, 129),
          (200, 3), (65, 64), (64, 65)]


def _sys(nx, ny, kind, seed):
    if kind == "p" and nx >= 4 and ny >= 4:
        return P.pcorr_system(nx, ny, seed=seed)
    return P.random_f3_system(nx, ny, seed=seed)


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("kind", ["f3", "p"])
def test_factor_schedule_matches_oracle(orc, emu, nx, ny, kind):
    s = _sys(nx, ny, kind, nx + 3 * ny)
    dl, du = orc.lu2d9(s.a, nx, ny)
    edl, edu = np.zeros_like(dl), np.zeros_like(du)
    for K, seed in [(1, 1), (5, 2), (32, 3)]:
        rc = emu.emu_lu2d9(s.a, edl, edu, nx, ny, 0.5, K, seed)
        assert rc == 0, emu.emu_error()
        assert np.array_equal(edl, dl) and np.array_equal(edu, du), (K, seed)


@pytest.mark.parametrize("nx,ny", SHAPES)
def test_iteration_schedule_matches_oracle(orc, emu, nx, ny):
    s = _sys(nx, ny, "p", 5 * nx + ny)
    dl, du = orc.lu2d9(s.a, nx, ny)
    ref, _, _, _ = orc.fb2d9(s.a, dl, du, nx, ny, np.zeros(s.N), 0.0, 3, s.rhs)
    x = np.zeros(s.N)
    r = np.zeros(s.N)
    w = np.zeros(s.N)
    for k in range(3):
        emu.emu_residual(s.a, x, s.rhs, r, nx, ny)
        assert np.array_equal(r, orc.residual9(s.a, x, s.rhs, nx, ny))
        rc = emu.emu_msi_iteration(r, dl, du, x, w, nx, ny, 1 + 7 * k, 100 + k)
        assert rc == 0, emu.emu_error()
    assert np.array_equal(x, ref)
