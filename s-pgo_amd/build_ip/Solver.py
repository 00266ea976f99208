"""``CudaSolver``: drop-in for the reference's Cython class
(src/CythonAPI/Solver.pyx:16-107), backed by the gfx950 library.

Same constructor, method names, argument meaning, dtypes and return shapes.
Array arguments are typed ``float[:]`` in the reference, so a buffer that is not
float32 raises ``ValueError`` here as Cython's buffer check does there.  The
default semantics are the reference step (no damping, every update accepted,
Cauchy weights with the near-frame boost, pose-only step size, inverse-depth
clamp), evaluated in double precision; ``set_options`` switches to damped LM,
Huber or the mixed-precision path.
"""
from __future__ import annotations

import numpy as np

from spgo import _lib
from spgo.solver import Solver as _Solver
from spgo.solver import make_options


def _float_view(a, name: str) -> np.ndarray:
    arr = np.asarray(a)
    if arr.dtype != np.float32:
        raise ValueError(f"Buffer dtype mismatch for {name}: expected 'float' but got '{arr.dtype}'")
    if arr.ndim != 1:
        raise ValueError(f"Buffer has wrong number of dimensions for {name} (expected 1, got {arr.ndim})")
    return np.ascontiguousarray(arr)


class CudaSolver:
    def __init__(self, N: int, M: int):
        self._s = _Solver(int(N), int(M), make_options(mode="reference"))
        self._N = int(N)
        self._M = int(M)

    # Solver.pyx:25-26
    def step(self, iterations: int) -> None:
        self._s.step(int(iterations))

    # Solver.pyx:28-32
    def loadCalibration(self, intrinsics_py, T_r_to_l_py) -> None:
        i = _float_view(intrinsics_py, "intrinsics_py")
        t = _float_view(T_r_to_l_py, "T_r_to_l_py")
        self._s.load_calibration(i, t)

    # Solver.pyx:34-38
    def writeObservations(self, anchor_frame_ID: int, target_frame_ID: int, global_feat_ID: int,
                          left_obs_py, right_obs_py) -> None:
        left = _float_view(left_obs_py, "left_obs_py")
        right = _float_view(right_obs_py, "right_obs_py")
        self._s.write_observation(int(anchor_frame_ID), int(target_frame_ID), int(global_feat_ID), left, right)

    # Solver.pyx:41-52
    def getIncrementalPose(self, keyFrameID: int) -> np.ndarray:
        return self._s.get_incremental_pose(int(keyFrameID))

    # Solver.pyx:54-57
    def writeIncrementalPose(self, keyFrameID: int, T_curr_to_next_py) -> None:
        self._s.write_incremental_pose(int(keyFrameID), _float_view(T_curr_to_next_py, "T_curr_to_next_py"))

    # Solver.pyx:59-72
    def getObservation(self, keyFrameID: int, global_feat_ID: int):
        left, right = self._s.get_observation(int(keyFrameID), int(global_feat_ID))
        return left.reshape((3, 1)), right.reshape((3, 1))

    # Solver.pyx:75-88 (the 8 intrinsics come back reshaped to (4, 2))
    def getCalibration(self):
        intr, T = self._s.get_calibration()
        return intr.reshape((4, 2)), T.reshape((4, 4))

    # Solver.pyx:90-93
    def loadInverseDepths(self, inverse_depths_py) -> None:
        self._s.load_inverse_depths(_float_view(inverse_depths_py, "inverse_depths_py"))

    # Solver.pyx:95-104
    def getInverseDepths(self) -> np.ndarray:
        return self._s.get_inverse_depths()

    # Solver.pyx:106-107
    def setStepSize(self, step_size) -> None:
        self._s.set_step_size(float(step_size))

    # -- additions (no reference counterpart) --------------------------------
    def set_options(self, **kw) -> None:
        """Switch semantics, e.g. ``set_options(mode='lm', robust='huber')``."""
        self._s.set_options(make_options(**kw))

    @property
    def solver(self) -> _Solver:
        return self._s


__all__ = ["CudaSolver"]
_ = _lib
