"""C-ABI library checks that need no GPU: the library loads, exports every
entry point include/spgo.h declares, and the host-side API (ingestion,
getters, layout construction, argument errors) behaves like the reference's
Solver facade (src/Optimization/Solver.cpp, src/CythonAPI/Solver.pyx)."""
import ctypes
import os
import re

import numpy as np
import pytest

from spgo import _lib
from spgo.problems import make_problem
from spgo.solver import Solver, make_options, solver_from_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    src = open(os.path.join(ROOT, "include", "spgo.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(spgo_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes prototypes cover exactly the declared functions
    assert sorted(_lib.SIGNATURES) == names


def test_default_options_and_struct_layout():
    o = _lib.default_options()
    assert o.mode == _lib.SPGO_MODE_LM and o.robust == _lib.SPGO_ROBUST_CAUCHY
    assert o.robust_param == 9.0 and o.eps == 1e-4 and o.lambda0 == 1e-4
    assert o.alpha_min == pytest.approx(1 / 30) and o.alpha_max == pytest.approx(5.0)
    assert ctypes.sizeof(_lib.Options) == 4 * 4 + 9 * 8 + 5 * 4 + 7 * 4 + 8
    assert _lib.Options.pcg_tol.offset == 136
    assert o.factorization == _lib.SPGO_FACTOR_AUTO
    assert o.pcg_max_iter == 200 and o.pcg_tol == 1e-8
    assert ctypes.sizeof(_lib.IterStats) == 3 * 8 + 4 * 4


def test_reference_facade_semantics_on_host():
    s = Solver(3, 2)
    # identity poses, inverse depths 1.0 (include/Solver.hpp:36-44)
    assert np.array_equal(s.get_incremental_pose(1), np.eye(4, dtype=np.float32))
    assert np.array_equal(s.get_inverse_depths(), np.ones(6, np.float32))
    intr = np.array([500, 501, 320, 240, 502, 503, 321, 241], np.float32)
    T = np.eye(4, dtype=np.float32)
    T[0, 3] = 0.12
    s.load_calibration(intr, T.reshape(16))
    i2, t2 = s.get_calibration()
    assert np.array_equal(i2, intr) and np.array_equal(t2, T.reshape(16))
    # observation cells: later writes win, unwritten cells read zero
    s.write_observation(0, 1, 4, [1, 2, 1], [3, 4, 1])
    s.write_observation(0, 1, 4, [5, 6, 1], [7, 8, 1])
    l3, r3 = s.get_observation(1, 4)
    assert np.array_equal(l3, [5, 6, 1]) and np.array_equal(r3, [7, 8, 1])
    l3, r3 = s.get_observation(2, 5)
    assert not l3.any() and not r3.any()
    # writeIncrementalPose writes rows 0-2 only (Solver.cpp:69-84)
    P = np.arange(16, dtype=np.float32)
    s.write_incremental_pose(2, P)
    got = s.get_incremental_pose(2)
    assert np.array_equal(got[:3], P.reshape(4, 4)[:3]) and np.array_equal(got[3], [0, 0, 0, 1])
    a = np.linspace(0.1, 0.6, 6).astype(np.float32)
    s.load_inverse_depths(a)
    assert np.array_equal(s.get_inverse_depths(), a)
    X, al = s.get_state()
    assert X.dtype == np.float64 and np.array_equal(al, a.astype(np.float64))


def test_argument_errors():
    s = Solver(3, 2)
    with pytest.raises(RuntimeError, match="out of range"):
        s.write_observation(0, 4, 0, [0, 0, 1], [0, 0, 1])      # frames are 0..N
    with pytest.raises(RuntimeError, match="out of range"):
        s.write_observation(0, 1, 6, [0, 0, 1], [0, 0, 1])      # slots are 0..N*M-1
    with pytest.raises(RuntimeError, match="out of range"):
        s.get_incremental_pose(3)
    with pytest.raises(RuntimeError):
        Solver(0, 5)
    # capacity N(N+1)M (include/Solver.hpp:20)
    s1 = Solver(1, 1)
    s1.write_observation(0, 0, 0, [0, 0, 1], [0, 0, 1])
    s1.write_observation(0, 1, 0, [0, 0, 1], [0, 0, 1])
    with pytest.raises(RuntimeError, match="capacity"):
        s1.write_observation(0, 1, 0, [0, 0, 1], [0, 0, 1])


def test_layout_of_generated_problem():
    p = make_problem(10, 500, 4, seed=1001)
    s = Solver(p.n_poses, p.m_per_frame, make_options(chunk_landmarks=64, group_chunks=1))
    s.write_observations(p.meas_a, p.meas_t, p.meas_l, p.uv)
    info = s.layout_info()
    assert info["n_measurements"] == 2000 and info["n_landmarks"] == 500
    assert info["max_span"] == 3                      # host + 3 following frames
    per_anchor = np.bincount(p.meas_a[p.meas_t == p.meas_a])
    assert info["n_groups"] == int(np.sum((per_anchor + 63) // 64))
    assert info["n_anchors_used"] == int(np.count_nonzero(per_anchor))
    # 6 x 10 pose unknowns padded to whole 32-row blocks; the profile of a
    # landmark span of 3 frames fits the envelope factorisation
    assert info["reduced_size"] == 64
    assert info["factorization"] == _lib.SPGO_FACTOR_ENVELOPE
    s2 = Solver(p.n_poses, p.m_per_frame, make_options(factorization=_lib.SPGO_FACTOR_DENSE))
    s2.write_observations(p.meas_a, p.meas_t, p.meas_l, p.uv)
    assert s2.layout_info()["factorization"] == _lib.SPGO_FACTOR_DENSE


def test_envelope_profile_depth_of_the_64_keyframe_problem():
    # 8 observations per landmark over consecutive frames: pose rows couple 6
    # poses back (36 scalars), at most 3 blocks of 16 below the diagonal
    from spgo.problems import make_config
    p = make_config("c2_fp64")
    s = Solver(p.n_poses, p.m_per_frame)
    s.write_observations(p.meas_a, p.meas_t, p.meas_l, p.uv)
    info = s.layout_info()
    a, t, l = p.meas_a, p.meas_t, p.meas_l
    lo = np.full(p.n_slots, 1 << 30)
    hi = np.full(p.n_slots, -1)
    np.minimum.at(lo, l, np.minimum(a, t))
    np.maximum.at(hi, l, np.maximum(a, t))
    first_pose = np.arange(p.n_poses)
    for f0, f1 in zip(lo[hi >= 0], hi[hi >= 0]):
        first_pose[f0:f1] = np.minimum(first_pose[f0:f1], f0)
    rows = np.arange(6 * p.n_poses)
    first_row = 6 * first_pose[rows // 6]
    depth = max(b - first_row[16 * b:16 * b + 16].min() // 16 for b in range(6 * p.n_poses // 16))
    assert info["profile_depth"] == depth == 3
    # 24 block rows of a profile 3 deep: 8 super-rows of 48, long enough for
    # cyclic reduction
    assert info["factorization"] == _lib.SPGO_FACTOR_CR and info["reduced_size"] == 384


def test_layout_coordinates_for_mixed_anchors_and_wide_spans():
    s = Solver(20, 2)
    s.write_observation(0, 0, 1, [1, 1, 1], [1, 1, 1])
    s.write_observation(1, 2, 1, [1, 1, 1], [1, 1, 1])
    with pytest.raises(RuntimeError, match="different anchor"):
        s.layout_info()
    # a landmark spanning 19 poses: frame coordinates (AUTO), refused by the
    # incremental layout when that is asked for
    s2 = Solver(20, 2)
    s2.write_observation(0, 0, 1, [1, 1, 1], [1, 1, 1])
    s2.write_observation(0, 19, 1, [1, 1, 1], [1, 1, 1])
    info = s2.layout_info()
    assert info["coordinates"] == _lib.SPGO_COORD_FRAMES and info["max_span"] == 19
    s2.set_options(make_options(coordinates=_lib.SPGO_COORD_INCREMENTAL))
    with pytest.raises(RuntimeError, match="spans"):
        s2.layout_info()
    s2.set_options(make_options(coordinates=_lib.SPGO_COORD_FRAMES))
    assert s2.layout_info()["coordinates"] == _lib.SPGO_COORD_FRAMES
    # short spans stay incremental under AUTO
    p = make_problem(10, 200, 4, seed=3)
    assert solver_from_problem(p).layout_info()["coordinates"] == _lib.SPGO_COORD_INCREMENTAL
    # one landmark in more frames than the frame-coordinate kernel holds
    s3 = Solver(70, 1)
    for t in range(0, 67):
        s3.write_observation(0, t, 0, [1, 1, 1], [1, 1, 1])
    with pytest.raises(RuntimeError, match="at most 64"):
        s3.layout_info()


def test_rejected_bulk_write_leaves_the_table_unchanged():
    s = Solver(4, 2)
    s.write_observations([0, 0], [0, 1], [0, 0], np.ones((2, 4), np.float32))
    before = s.layout_info()["n_measurements"]
    with pytest.raises(RuntimeError, match="nothing was written"):
        s.write_observations([0, 0, 0], [2, 3, 9], [1, 1, 1], np.full((3, 4), 7, np.float32))
    assert s.layout_info()["n_measurements"] == before
    l3, _ = s.get_observation(2, 1)
    assert np.array_equal(l3, np.zeros(3, np.float32))


def test_step_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    p = make_problem(4, 20, 3, seed=3)
    s = Solver(p.n_poses, p.m_per_frame)
    s.load_calibration(p.intrinsics, p.T_r_to_l)
    s.write_observations(p.meas_a, p.meas_t, p.meas_l, p.uv)
    with pytest.raises(RuntimeError, match="device"):
        s.step(1)


def test_dropin_module_surface():
    from build_ip import Solver as S
    cs = S.CudaSolver(3, 4)
    for name in ["step", "loadCalibration", "writeObservations", "getIncrementalPose", "writeIncrementalPose",
                 "getObservation", "getCalibration", "loadInverseDepths", "getInverseDepths", "setStepSize"]:
        assert callable(getattr(cs, name))
    intr = np.arange(8, dtype=np.float32)
    cs.loadCalibration(intr, np.eye(4, dtype=np.float32).reshape(16))
    ci, cT = cs.getCalibration()
    assert ci.shape == (4, 2) and np.array_equal(ci.reshape(8), intr) and cT.shape == (4, 4)
    cs.writeObservations(0, 1, 2, np.array([1, 2, 1], np.float32), np.array([3, 4, 1], np.float32))
    l, r = cs.getObservation(1, 2)
    assert l.shape == (3, 1) and r.shape == (3, 1) and l[1, 0] == 2
    assert cs.getIncrementalPose(0).shape == (4, 4) and cs.getIncrementalPose(0).dtype == np.float32
    assert cs.getInverseDepths().shape == (12,)
    with pytest.raises(ValueError, match="dtype"):
        cs.loadInverseDepths(np.ones(12))                       # float64 buffer, as Cython's float[:] check


def test_problem_npz_round_trip_and_ingestion(tmp_path):
    from spgo.problems import load_problem, make_problem, save_problem
    p = make_problem(8, 300, 4, seed=17, outlier_frac=0.05, name="npz")
    f = str(tmp_path / "p.npz")
    save_problem(f, p, seed=17)
    q = load_problem(f)
    assert (q.n_poses, q.m_per_frame, q.n_landmarks, q.name) == (p.n_poses, p.m_per_frame, p.n_landmarks, "npz")
    for k in ("meas_a", "meas_t", "meas_l", "uv", "intrinsics", "T_r_to_l", "X0", "alpha0", "X_true",
              "alpha_true", "outlier"):
        a, b = getattr(p, k), getattr(q, k)
        assert a.dtype == b.dtype and np.array_equal(a, b), k
    # the loaded problem ingests into the same host layout
    a = solver_from_problem(p).layout_info()
    b = solver_from_problem(q).layout_info()
    assert a == b
