"""Drop-in for the reference's compiled module: ``from build_ip import Solver``
followed by ``Solver.CudaSolver(n, m)`` (CUDAmanager.py:1, 29)."""
from . import Solver  # noqa: F401
