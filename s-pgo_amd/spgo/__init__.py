"""MI355X (gfx950) Levenberg-Marquardt stereo bundle-adjustment step.

``spgo.Solver`` drives ``lib/libspgo.so`` (hand-written HIP kernels behind a C
ABI) through ctypes; ``build_ip.Solver.CudaSolver`` is the drop-in for the
reference's Cython class.
"""
from .solver import Batch, Solver, make_options, solver_from_problem  # noqa: F401

__all__ = ["Batch", "Solver", "make_options", "solver_from_problem"]
