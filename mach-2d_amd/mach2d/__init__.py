"""mach2d -- MI355X (gfx950) SIMPLEC linear-algebra hot path of MACH-2D.

Host-side mirror of the reference solver interface (solvers.py) and of the
coefficient assembly (assembly.py) over the C ABI of libmach2d_hip.so
(_lib.py), and the synthetic workload generators (problems.py).  Importing touches no GPU; the library loads on first use.
"""
from . import assembly, problems, solvers  # noqa: F401

__all__ = ["assembly", "problems", "solvers"]
