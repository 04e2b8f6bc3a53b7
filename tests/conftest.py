import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "mach-2d_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP path)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(oracle.ORACLE_SO):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"])
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def emu():
    import ctypes as C

    import numpy as np
    d = os.path.join(ROOT, "tests", "native")
    so = os.path.join(d, "libemu.so")
    if not os.path.exists(so):
        subprocess.check_call(["make", "-s", "-C", d])
    E = C.CDLL(so)
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    E.emu_lu2d9.argtypes = [dp, dp, dp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_uint]
    E.emu_msi_iteration.argtypes = [dp, dp, dp, dp, dp, C.c_int, C.c_int, C.c_int, C.c_uint]
    E.emu_residual.argtypes = [dp, dp, dp, dp, C.c_int, C.c_int]
    E.emu_error.restype = C.c_char_p
    return E


def load_golden(name):
    import numpy as np
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip("golden fixture %s missing" % name)
    return dict(np.load(path, allow_pickle=False))


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a GPU; tests marked gpu only."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from mach2d import _lib, solvers
    _lib.load()
    return solvers
