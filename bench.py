#!/usr/bin/env python3
"""LM iterations/s of the MI355X stereo bundle-adjustment step.

Workload (default): one seeded problem with 256 keyframes (1536 x 1536
reduced pose system), 100,000 landmarks and 10 stereo observations per
landmark (1,000,000 stereo measurements), 5 % outliers of up to 50 px, Huber
2 px, damped LM (lambda0 = 1e-4).  It is solved twice from the same state:
by the plain FP64 GPU path (the baseline) and by the mixed-precision path
(FP32 residuals, Jacobians, assembly and landmark Schur on MFMA, FP64 reduced
system).  A step is one LM iteration: reduced-system formation, Cholesky,
solves, pose update and the linearisation of the candidate with
accept/reject -- all on the GPU, inputs resident in HBM.  ``value`` is the
fastest form whose final cost matches the FP64 path's to 1e-6 relative;
``vs_baseline`` is its ratio to the FP64 path.  Prints one JSON line.

The roofline object describes the kernel that takes the most device time,
timed with HIP events on the solver's stream over the timed iterations (a
second pass of the same iterations launched one phase at a time); the
assembly kernel's HBM roofline is reported beside it.  HBM traffic per launch
comes from the rocprofv3 counter summaries under profiles/ (FETCH_SIZE and
WRITE_SIZE, scaled by the reading of a plain copy of known size; see
--calibrate-copy).  The CPU baseline is the FP64 oracle timed on this host,
one thread, on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "s-pgo_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    """Progress on stderr (the JSON line is the only stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)

FP64_PEAK_TFLOPS = 78.6


def assembly_bytes(p, real_bytes: int = 8) -> int:
    """Algorithmic HBM bytes of one linearisation launch (candidate mode):
    per measurement the target id, packed landmark/position index and four
    float32 pixels (24 B); per landmark its slot id, ray pixel, position range,
    inverse depth in and out, C and g_alpha in and out, and its B column v_l
    (6 k_l values, read for the back-substitution and written new); per anchor
    frame the reduced (6 k_a)^2 block and 6 k_a gradient written."""
    a, t, l = p.meas_a.astype(np.int64), p.meas_t.astype(np.int64), p.meas_l.astype(np.int64)
    O = a.size
    span = np.zeros(p.n_slots, np.int64)
    np.maximum.at(span, l, np.maximum(t - a, 0))
    used = np.unique(l)
    k_l = span[used]
    anchors = np.zeros(p.n_slots, np.int64)
    anchors[l] = a
    k_a = np.zeros(p.n_poses + 1, np.int64)
    np.maximum.at(k_a, anchors[used], k_l)
    b = O * 24
    b += used.size * (4 + 8 + 8 + 2 * real_bytes + 4 * real_bytes)
    b += int(np.sum(6 * k_l)) * real_bytes * 2
    b += int(np.sum((6 * k_a) ** 2 + 6 * k_a)) * real_bytes
    return int(b)


def minimal_bytes_fields(nbytes: int, ms: float) -> dict:
    """The roofline fields of the minimal byte count beside the builder's own."""
    ach = nbytes / (ms * 1e-3) / 1e9
    return {"algorithmic_bytes_minimal": nbytes, "achieved_minimal": ach, "frac_minimal": ach / HBM_PEAK_GBS}


def assembly_bytes_minimal(p, real_bytes: int = 4) -> int:
    """The minimal algorithmic bytes of one linearisation, counted per item
    of the problem with nothing re-read: per measurement 3 int32 ids and 4
    pixels (28 B), per landmark the anchor pixel and alpha (12 B), per pose
    its 3x4 matrix (48 B) read; per landmark C and g_alpha (8 B), its B
    column (6 k_l values) and per coupled pose pair one 6x6 block written.
    Floating-point terms scale with real_bytes / 4.  Unlike assembly_bytes it
    leaves out the B-column read of the back-substitution and the per-anchor
    blocks' full squares."""
    f = real_bytes / 4
    a, t, l = p.meas_a.astype(np.int64), p.meas_t.astype(np.int64), p.meas_l.astype(np.int64)
    O = a.size
    span = np.zeros(p.n_slots, np.int64)
    np.maximum.at(span, l, np.maximum(t - a, 0))
    used = np.unique(l)
    k_l = span[used]
    anchors = np.zeros(p.n_slots, np.int64)
    anchors[l] = a
    # coupled pose pairs (i >= j): poses a .. a + k - 1 of each landmark couple pairwise
    first = np.arange(p.n_poses)
    for a0, k in zip(anchors[used], k_l):
        if k > 0:
            first[a0:a0 + k] = np.minimum(first[a0:a0 + k], a0)
    covered = np.zeros(p.n_poses, bool)
    for a0, k in zip(anchors[used], k_l):
        covered[a0:a0 + k] = True
    n_pairs = int(np.sum((np.arange(p.n_poses) - first + 1)[covered]))
    rd = O * (12 + 16 * f) + used.size * 12 * f + p.n_poses * 48 * f
    wr = used.size * 8 * f + int(np.sum(6 * k_l)) * 4 * f + n_pairs * 36 * 4 * f
    return int(rd + wr)


def scalar_profile(p) -> np.ndarray:
    """First coupled column of every row of the reduced pose system: pose j
    couples back to the first frame of every landmark whose observations span
    it (the frames of one landmark are consecutive)."""
    a, t, l = p.meas_a, p.meas_t, p.meas_l
    lo = np.full(p.n_slots, 1 << 30)
    hi = np.full(p.n_slots, -1)
    np.minimum.at(lo, l, np.minimum(a, t))
    np.maximum.at(hi, l, np.maximum(a, t))
    first_pose = np.arange(p.n_poses)
    for f0, f1 in zip(lo[hi >= 0], hi[hi >= 0]):
        first_pose[f0:f1] = np.minimum(first_pose[f0:f1], f0)
    return 6 * first_pose[np.arange(6 * p.n_poses) // 6]


def envelope_flops(p) -> float:
    """Algorithmic flops of the envelope Cholesky and both triangular solves:
    column j with n_j entries below the diagonal inside the profile costs one
    square root, n_j scalings and n_j (n_j + 1) / 2 multiply-adds of the
    update; each solve costs 2 n_j + 1 per column."""
    first = scalar_profile(p)
    P = first.size
    n = np.array([np.count_nonzero(first[j + 1:] <= j) for j in range(P)], dtype=np.float64)
    return float(np.sum(1 + n + n * (n + 1)) + 2 * np.sum(2 * n + 1))


def dense_flops(n_poses: int) -> float:
    P = 6 * n_poses
    return P ** 3 / 3.0 + 2 * P ** 2


# factorisation paths (layout_info()["factorization"]) that run the whole solve
# in the "cholesky" phase, and the kernels each one launches
BAND_PATHS = (2, 3, 5)
CHOL_KERNELS = {1: "k_chol_panel + k_chol_update + k_solve_fwd + k_solve_bwd",
                2: "k_chol_env", 3: "k_chol_split1 + k_chol_split2",
                5: "k_cr_factor + k_cr_update + k_cr_tail3 + k_cr_back (cyclic reduction; the candidate's pose "
                   "update runs in k_cr_back's last workgroup, so the pose_update phase holds no kernel)"}
CHOL_TRAFFIC = {2: "k_chol_env", 3: "k_chol_split", 5: "k_cr"}
CHOL_NAMES = {1: "dense", 2: "envelope", 3: "envelope, split", 5: "cyclic reduction"}


def profile_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed counter summary
    profiles/traffic_<kernel>.txt (a line 'traffic_bytes_per_launch: N')."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}.txt")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(p, n_iter: int, gpu_recs):
    """FP64 oracle on the host: seconds per LM iteration on n_iter iterations,
    and the relative cost difference of the GPU's first iterations."""
    from oracle import ba_oracle as O

    try:
        from threadpoolctl import threadpool_info
        cores = max([i.get("num_threads", 1) for i in threadpool_info()] + [1])
    except Exception:
        cores = 1
    pa = O.from_problem(p)
    X0 = p.X0.astype(np.float64)
    a0 = p.alpha0.astype(np.float64)
    t0 = time.perf_counter()
    O.lm_solve(pa, X0, a0, 0, O.Options())
    t1 = time.perf_counter()
    _, _, _, c0, recs = O.lm_solve(pa, X0, a0, n_iter, O.Options())
    t2 = time.perf_counter()
    per_iter = max((t2 - t1) - (t1 - t0), 1e-9) / n_iter
    errs = [abs(q["cost"] - r.cost) / abs(r.cost) for q, r in zip(gpu_recs, recs)]
    return per_iter, cores, max(errs) if errs else None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-cg", action="store_true", help="c3: skip the CG forms (and the network's training)")
    ap.add_argument("--chunk-landmarks", type=int, default=0)
    ap.add_argument("--group-chunks", type=int, default=0)
    ap.add_argument("--config", default="c3", choices=["c2", "c3", "c4", "c5"],
                    help="c3 (the headline): 256 keyframes, 100k landmarks, 5%% outliers, Huber 2 px, run in FP64 "
                         "and in mixed precision; c2: 64 keyframes, 20k landmarks, FP64; c4: 512 independent windows batched, "
                         "FP32 assembly, 10 LM iterations per solve (a step is one batched solve); c5: 1024 "
                         "keyframes, 400k landmarks, 12 observations each, LM with CG preconditioned by the "
                         "trained network against the Cholesky path")
    ap.add_argument("--no-plain", action="store_true",
                    help="c4: skip the one-window-at-a-time FP64 solver and the FP64 batch (profiling runs)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="c4: use the tile shape given by --chunk-landmarks/--group-chunks instead of the autotuner")
    ap.add_argument("--c5-systems", type=int, default=2000, help="c5: systems the network is trained on")
    ap.add_argument("--pcg-max-iter", type=int, default=200, help="c5: CG iterations before the Cholesky fallback")
    ap.add_argument("--c5-forms", default="", help="c5: comma-separated subset of the forms to run (default: all)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of 256 MiB five times (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8}))
        return 0
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the single-problem LM step is measured on one MI355X)", file=sys.stderr)
        return 2

    if args.config == "c2":
        return bench_c2(args)
    if args.config == "c4":
        return bench_c4(args)
    if args.config == "c5":
        return bench_c5(args)
    return bench_c3(args)


def bench_c2(args) -> int:
    """64 keyframes, 20,000 landmarks, 8 observations each, FP64 LM."""
    from spgo.problems import make_config
    from spgo.solver import make_options, solver_from_problem
    from spgo import accuracy

    p = make_config("c2_fp64")
    kw = {}
    if args.chunk_landmarks:
        kw["chunk_landmarks"] = args.chunk_landmarks
    if args.group_chunks:
        kw["group_chunks"] = args.group_chunks
    s = solver_from_problem(p, make_options(mode="lm", **kw))

    # warm-up iterations, then restart from the initial state
    if args.warmup > 0:
        s.step(args.warmup)
    X0 = p.X0.astype(np.float64)
    s.set_state(X0, p.alpha0.astype(np.float64))
    s.step(0)                                   # linearise the initial state (synchronises)
    t0 = time.perf_counter()
    s.step(args.steps)                          # synchronises before returning
    t1 = time.perf_counter()
    wall = t1 - t0
    dev_ms = s.last_step_ms()
    c0, recs = s.stats()

    # second pass: the same iterations, one launch per phase, HIP events per phase
    s.set_state(X0, p.alpha0.astype(np.float64))
    phases = s.profile_phases(args.steps)
    info = s.layout_info()
    fac = info["factorization"]
    envelope = fac in BAND_PATHS   # envelope (one workgroup, or split) or cyclic reduction
    if envelope:
        # one kernel factors and solves; the "solve" phase holds no kernel
        phases["cholesky"] += phases.pop("solve")
    nbytes = assembly_bytes(p)
    ach_a = nbytes / (phases["assemble"] * 1e-3) / 1e9
    roof_asm = {"kernel": "k_assemble<double>", "bound": "hbm", "achieved": ach_a, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": ach_a / HBM_PEAK_GBS, "traffic": profile_traffic("k_assemble"),
                "algorithmic_bytes": nbytes}
    roof_asm.update(minimal_bytes_fields(assembly_bytes_minimal(p, 8), phases["assemble"]))
    dom = max(phases, key=phases.get)
    if dom == "cholesky":
        flops = envelope_flops(p) if envelope else dense_flops(p.n_poses)
        ach = flops / (phases[dom] * 1e-3) / 1e12
        roof = {"kernel": CHOL_KERNELS[fac], "bound": "mfma", "achieved": ach, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
                "traffic": profile_traffic(CHOL_TRAFFIC[fac]) if envelope else None,
                "algorithmic_flops": flops,
                "note": "envelope flops (profile Cholesky and both solves) over the phase's event time; the "
                        "factorisation is a chain of small dense factorisations (latency-bound); FP64 peak is "
                        "the whole chip's"}
    else:
        roof = dict(roof_asm)
        if dom != "assemble":
            roof["note"] = f"dominant phase is {dom}; roofline given for the assembly kernel"

    X, al = s.get_state()
    used = np.unique(p.meas_l)
    acc = accuracy.summary(X, p.X_true, al, p.alpha_true, used)

    cpu = None
    err = None
    if not args.no_cpu:
        per_iter, cores, err = cpu_baseline(p, 2, recs)
        cpu = {"value": 1.0 / per_iter, "unit": "LM iterations/s", "cores": cores, "kind": "port",
               "sample": "FP64 oracle (numpy/scipy restatement of the reference step), 2 LM iterations "
                         "of the same 64-keyframe problem"}

    line = {
        "metric": "LM iters/s & batched solves/s vs plain FP64 GPU solver at matched final cost",
        "value": args.steps / wall,
        "unit": "LM iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "device_ms_per_step": dev_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator: circle trajectory r=2 m, landmarks in a 4 m box, 1 px noise)",
        "config": {"workload": "64 keyframes, 20000 landmarks, 8 stereo observations each (160000 stereo "
                               "measurements), FP64 plain LM, lambda0 1e-4, Cauchy c^2=9",
                   "keyframes": p.n_poses, "landmarks": p.n_landmarks, "measurements": p.n_obs,
                   "reduced_system": 6 * p.n_poses},
        "accuracy": acc,
        "cost_initial": c0,
        "cost_final": recs[-1]["cost"] if recs else c0,
        "oracle_rel_err": err,
        "phases_ms": phases,
        "roofline": roof,
        "roofline_assemble": roof_asm,
        "factorization": CHOL_NAMES.get(fac, str(fac)),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


def oracle_sample_rate(p, opts, n_frames: int):
    """CPU baseline on a bounded sample of problem ``p``: the FP64 oracle, one
    BLAS thread, on the sub-problem of the first ``n_frames`` keyframes (every
    measurement whose anchor and target frames lie in it).  Returns (LM
    iterations/s scaled to the whole problem by its measurement count, the
    sample's per-iteration seconds, its measurement count)."""
    from oracle import ba_oracle as O

    sel = (p.meas_a <= n_frames) & (p.meas_t <= n_frames)
    pa = O.problem_arrays(n_frames, p.n_slots, p.intrinsics, p.T_r_to_l, p.meas_a[sel], p.meas_t[sel],
                          p.meas_l[sel], p.uv[sel])
    X0 = p.X0[:n_frames].astype(np.float64)
    a0 = p.alpha0.astype(np.float64)
    import contextlib
    try:
        from threadpoolctl import threadpool_limits
        limit = threadpool_limits(1)
    except Exception:
        limit = contextlib.nullcontext()
    with limit:
        t0 = time.perf_counter()
        O.lm_solve(pa, X0, a0, 0, opts)
        t1 = time.perf_counter()
        O.lm_solve(pa, X0, a0, 1, opts)
        t2 = time.perf_counter()
    per_iter = max((t2 - t1) - (t1 - t0), 1e-9)
    n_sub = int(sel.sum())
    return n_sub / (per_iter * p.n_obs), per_iter, n_sub


def reference_cpu_path_rate(n_timed: int = 5):
    """The reference's CPU path (Optimizer.step restated, oracle/cpu_path.py)
    on the 10-keyframe, 500-landmark problem: seconds per call, median of
    ``n_timed`` after one warm-up with the default BLAS threads, and one call
    with one BLAS thread."""
    import contextlib
    import statistics
    from oracle import cpu_path as CP
    from spgo.problems import make_config

    p1 = make_config("c1_small")
    c = CP.from_problem(p1)
    c.step()
    ts = []
    for _ in range(n_timed):
        t0 = time.perf_counter()
        c.step()
        ts.append(time.perf_counter() - t0)
    try:
        from threadpoolctl import threadpool_info, threadpool_limits
        blas = max([i.get("num_threads", 1) for i in threadpool_info()] + [1])
        limit = threadpool_limits(1)
    except Exception:
        blas, limit = None, contextlib.nullcontext()
    with limit:
        t0 = time.perf_counter()
        c.step()
        t1 = time.perf_counter() - t0
    med = statistics.median(ts)
    return {"value": 1.0 / med, "unit": "LM iterations/s", "cores": blas, "host_cpu_count": os.cpu_count(),
            "one_thread_value": 1.0 / t1, "kind": "port",
            "sample": f"the reference CPU path (Optimizer.step, PythonUtils/Optimizer.py:135-288, restated in "
                      f"float64 numpy with its Python loops and dense Jacobians) on 10 keyframes / 500 landmarks / "
                      f"{p1.n_obs} stereo measurements: median of {n_timed} calls after one warm-up "
                      f"({med:.2f} s), BLAS at {blas} threads; one call with one BLAS thread {t1:.2f} s"}


def chol_roofline(p, phases, info):
    """Roofline object of the reduced-system factorisation phase."""
    fac = info["factorization"]
    envelope = fac in BAND_PATHS
    flops = envelope_flops(p) if envelope else dense_flops(p.n_poses)
    ms = phases["cholesky"] + (phases.get("solve", 0.0) if not envelope else 0.0)
    ach = flops / (ms * 1e-3) / 1e12
    traffic = {3: "k_chol_split_c3", 2: "k_chol_env", 5: "k_cr_c3"}.get(fac)
    return {"kernel": CHOL_KERNELS[fac], "bound": "mfma", "achieved": ach, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
            "traffic": profile_traffic(traffic) if traffic else None,
            "algorithmic_flops": flops,
            "note": "envelope flops (profile Cholesky and both solves) over the phase's event time; the "
                    "factorisation is a chain of small dense factorisations (latency-bound); FP64 peak is "
                    "the whole chip's"}


PCG_PATH = 4   # SPGO_FACTOR_PCG


def c3_network():
    """The preconditioner network for the 256-keyframe workload: trained on
    systems the solver generates for problems of the same generator (128-256
    keyframes, 10 observations per landmark, 5 % outliers, Huber), none of them
    the benchmark problem; its CG iterations on held-out systems of a further
    problem beside block Jacobi's."""
    from spgo.lpcg import LearnedPCG, solver_systems
    from spgo.problems import make_problem
    from spgo.solver import make_options

    K = 10
    W = K - 1
    opts = make_options(mode="lm", robust="huber", robust_param=2.0)
    t0 = time.perf_counter()
    train = [make_problem(n, 390 * n, K, seed=3300 + n, outlier_frac=0.05) for n in (128, 192, 256)]
    held = [make_problem(256, 390 * 256, K, seed=3700, outlier_frac=0.05)]
    systems = solver_systems(train, W, states_per_problem=3, iters_between=1, restarts=20, options=opts, seed=31)
    held_sys = solver_systems(held, W, states_per_problem=3, iters_between=1, restarts=2, options=opts, seed=37)
    t_data = time.perf_counter() - t0
    net = LearnedPCG(W, seed=13)
    net.load(systems)
    t0 = time.perf_counter()
    steps = 300
    losses = net.train(steps, lr=1e-3)
    t_train = time.perf_counter() - t0
    net.load(held_sys)
    it_l = [int(v) for v in net.solve(learned=True, tol=1e-8, max_iter=2000)[0]]
    it_b = [int(v) for v in net.solve(learned=False, tol=1e-8, max_iter=2000)[0]]
    log(f"network: {len(systems)} systems in {t_data:.1f} s, {steps} steps in {t_train:.1f} s, loss "
        f"{losses[0]:.4f} -> {losses[-1]:.4f}; held-out CG iterations learned {it_l}, block Jacobi {it_b}")
    info = {"systems": len(systems), "problem_keyframes": [q.n_poses for q in train], "steps": steps,
            "loss_first": float(losses[0]), "loss_last": float(losses[-1]), "data_s": t_data, "train_s": t_train,
            "held_out_systems": len(held_sys), "held_out_cg_iterations_learned": it_l,
            "held_out_cg_iterations_block_jacobi": it_b}
    return net, info


def bench_c3(args) -> int:
    """256 keyframes, 100,000 landmarks, 10 observations each, 5% outliers, Huber
    2 px: LM iterations/s of the plain FP64 path (the baseline) and of the
    mixed-precision path (FP32 linearisation and landmark Schur, FP64 reduced
    system), their final costs compared."""
    from oracle import ba_oracle as O
    from spgo import accuracy
    from spgo.problems import make_config
    from spgo.solver import make_options, solver_from_problem

    t0 = time.perf_counter()
    p = make_config("c3_mixed")
    log(f"problem generated in {time.perf_counter() - t0:.1f} s: {p.n_poses} keyframes, {p.n_obs} measurements")
    out = {}
    # tile shape: 64 landmarks per chunk; chunks per group left to the library
    # (0), which sizes the groups to the assembly's resident workgroups
    tile = {"chunk_landmarks": args.chunk_landmarks or 64, "group_chunks": args.group_chunks or 0}
    X0 = p.X0.astype(np.float64)
    a0 = p.alpha0.astype(np.float64)
    net, cg_training = (None, None) if args.no_cg else c3_network()
    forms = [("fp64", {"precision": "fp64"}, False), ("mixed", {"precision": "mixed"}, False)]
    if net is not None:
        # CG on the reduced system (tolerance 1e-8, Cholesky fallback), block
        # Jacobi and the trained network
        forms += [("mixed_cg_block_jacobi", {"precision": "mixed", "factorization": PCG_PATH}, False),
                  ("mixed_cg_learned", {"precision": "mixed", "factorization": PCG_PATH}, True)]
    for prec, kw, use_net in forms:
        s = solver_from_problem(p, make_options(mode="lm", robust="huber", robust_param=2.0, **kw, **tile))
        if use_net:
            s.set_preconditioner(net)
        if args.warmup > 0:
            s.step(args.warmup)
        s.set_state(X0, a0)
        s.step(0)                                   # linearise the initial state (synchronises)
        t1 = time.perf_counter()
        s.step(args.steps)                          # synchronises before returning
        wall = time.perf_counter() - t1
        c0, recs = s.stats()
        X, al = s.get_state()
        out[prec] = {"value": args.steps / wall, "ms_per_step": wall * 1e3 / args.steps,
                     "device_ms_per_step": s.last_step_ms() / args.steps, "cost_initial": c0,
                     "cost_final": recs[-1]["cost"] if recs else c0,
                     "accepted": int(sum(r["accepted"] for r in recs)),
                     "accuracy": accuracy.summary(X, p.X_true, al, p.alpha_true, np.unique(p.meas_l))}
        if kw.get("factorization") == PCG_PATH:
            out[prec]["cg_iterations"] = [r["pcg_iters"] for r in recs]
            out[prec]["fallbacks"] = int(sum(r["pcg_fallback"] for r in recs))
        s.set_state(X0, a0)
        out[prec]["phases_ms"] = s.profile_phases(min(args.steps, 10))
        out[prec]["layout"] = s.layout_info()
        log(f"{prec}: {out[prec]['value']:.1f} LM it/s, final cost {out[prec]['cost_final']:.10g}"
            + (f", CG iterations {out[prec]['cg_iterations']}, fallbacks {out[prec]['fallbacks']}"
               if "cg_iterations" in out[prec] else ""))
        s.close()
    if net is not None:
        net.close()
    base = out["fp64"]
    for r in out.values():
        r["final_cost_rel_diff_vs_fp64"] = abs(r["cost_final"] - base["cost_final"]) / base["cost_final"]
    # the headline form: the fastest one that reaches the FP64 path's final cost to 1e-6
    form = max((f for f in out if out[f]["final_cost_rel_diff_vs_fp64"] <= 1e-6), key=lambda f: out[f]["value"])
    best = out[form]
    real = 4 if form == "mixed" else 8
    phases = dict(best["phases_ms"])
    envelope = best["layout"]["factorization"] in BAND_PATHS
    if envelope:
        phases["cholesky"] += phases.pop("solve")
    nbytes = assembly_bytes(p, real)
    ach_a = nbytes / (phases["assemble"] * 1e-3) / 1e9
    roof_asm = {"kernel": f"k_assemble<{'float' if real == 4 else 'double'}>", "bound": "hbm", "achieved": ach_a,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach_a / HBM_PEAK_GBS,
                "traffic": profile_traffic(f"k_assemble_c3_{form}"), "algorithmic_bytes": nbytes}
    roof_asm.update(minimal_bytes_fields(assembly_bytes_minimal(p, real), phases["assemble"]))
    dom = max(phases, key=phases.get)
    if dom == "cholesky" and best["layout"]["factorization"] != PCG_PATH:
        roof = chol_roofline(p, phases, best["layout"])
    else:
        roof = dict(roof_asm)
        if dom != "assemble":
            roof["note"] = f"dominant phase is {dom}; roofline given for the assembly kernel"

    cpu = ref_cpu = None
    if not args.no_cpu:
        oo = O.Options(robust="huber", robust_param=2.0)
        rate, per_iter, n_sub = oracle_sample_rate(p, oo, 64)
        cpu = {"value": rate, "unit": "LM iterations/s", "cores": 1, "kind": "port",
               "host_cpu_count": os.cpu_count(),
               "sample": f"FP64 oracle (numpy/scipy restatement of the reference step), one BLAS thread, one LM "
                         f"iteration of the first 64 keyframes of the same problem ({n_sub} measurements, "
                         f"{per_iter:.2f} s), scaled to the full {p.n_obs} measurements"}
        log(f"cpu baseline: {rate:.4g} LM it/s (sample {per_iter:.2f} s per iteration)")
        ref_cpu = reference_cpu_path_rate()
        log(f"reference CPU path at 10 keyframes: {ref_cpu['value']:.3g} LM it/s")

    line = {
        "metric": "LM iters/s & batched solves/s vs plain FP64 GPU solver at matched final cost",
        "value": best["value"], "unit": "LM iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": best["ms_per_step"], "device_ms_per_step": best["device_ms_per_step"],
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": best["value"] / base["value"],
        "vs_baseline_note": "ratio to the plain FP64 GPU path on the same problem and steps "
                            "(BASELINE.md publishes no number for this metric)",
        "dtype": "f32" if form == "mixed" else "f64",
        "form": form,
        "data": "synthetic (seeded generator: circle trajectory r=2 m, landmarks in a 4 m box, 1 px noise, "
                "5% outliers of up to 50 px)",
        "config": {"workload": "256 keyframes, 100000 landmarks, 10 stereo observations each (1000000 stereo "
                               "measurements), 5% outliers, Huber 2 px, LM lambda0 1e-4",
                   "tile": tile,
                   "keyframes": p.n_poses, "landmarks": p.n_landmarks, "measurements": p.n_obs,
                   "reduced_system": 6 * p.n_poses},
        "final_cost_rel_diff_vs_fp64": best["final_cost_rel_diff_vs_fp64"],
        "cg_training": cg_training,
        "forms": out,
        "phases_ms": phases,
        "roofline": roof,
        "roofline_assemble": roof_asm,
        "cpu_baseline": cpu,
        "reference_cpu_path": ref_cpu,
    }
    print(json.dumps(line))
    return 0


def bench_c4(args) -> int:
    """512 independent sliding windows (keyframes U{8..32}, landmarks U{200..2000},
    observations per landmark U{3..8}), 10 LM iterations each.  A step is one
    batched solve of all windows (linearisation of the initial states and 10
    LM iterations, one launch per phase for all windows, graph-replayed, FP32
    assembly, tile shape from the autotuner).  The plain FP64 GPU solver solves
    the same windows one after another; final costs are compared per window."""
    from spgo.problems import WINDOW_BATCH, make_window_batch
    from spgo.solver import Batch, make_options, solver_from_problem

    iters = 10
    t_gen = time.perf_counter()
    ws = make_window_batch(**WINDOW_BATCH)
    t_gen = time.perf_counter() - t_gen
    nw = len(ws)
    init = [(w.X0.astype(np.float64), w.alpha0.astype(np.float64)) for w in ws]

    # plain FP64 GPU solver, one window at a time (buffers built by a warm-up step)
    t_lone = None
    if not args.no_plain:
        fo = make_options(mode="lm")
        lone = [solver_from_problem(w, fo) for w in ws]
        for s in lone:
            s.step(1)
        for s, (X0, a0) in zip(lone, init):
            s.set_state(X0, a0)
        t0 = time.perf_counter()
        for s in lone:
            s.step(iters)
        t_lone = time.perf_counter() - t0
        lone_final = np.array([s.stats()[1][-1]["cost"] for s in lone])
        lone_c0 = np.array([s.stats()[0] for s in lone])
        del lone

    def run_batch(precision):
        kw = {}
        if args.no_autotune:
            kw = {"chunk_landmarks": args.chunk_landmarks or 64, "group_chunks": args.group_chunks or 1}
        o = make_options(mode="lm", precision=precision, use_graph=1, **kw)
        bt = Batch([solver_from_problem(w, o) for w in ws], o)
        tune = dict(kw, ms=None) if args.no_autotune else bt.autotune(3)
        for _ in range(max(args.warmup, 1)):
            bt.reset_states()
            bt.step(iters)
        walls, dev = [], []
        for _ in range(args.steps):
            bt.reset_states()
            t0 = time.perf_counter()
            bt.step(iters)
            walls.append(time.perf_counter() - t0)
            dev.append(bt.last_step_ms())
        final = np.array([bt.stats(p)[1][-1]["cost"] for p in range(nw)])
        c0 = np.array([bt.stats(p)[0] for p in range(nw)])
        bt.reset_states()
        phases = bt.profile_phases(iters)
        return bt, tune, float(np.sum(walls)), float(np.mean(dev)), final, c0, phases

    bt, tune, wall, dev_ms, final, c0, phases = run_batch("mixed")
    info = bt.info()
    value = args.steps * nw / wall
    plain = {}
    if t_lone is not None:
        rel = np.abs(final - lone_final) / lone_final
        _, tune64, wall64, dev64, final64, _, phases64 = run_batch("fp64")
        rel64 = np.abs(final64 - lone_final) / lone_final
        plain = {
            "plain_fp64_gpu": {"value": nw / t_lone, "unit": "solves/s (one window at a time)",
                               "ms_per_batch": t_lone * 1e3},
            "speedup_vs_plain_fp64_gpu": value / (nw / t_lone),
            "final_cost_rel_diff_vs_plain_fp64": {"max": float(rel.max()), "median": float(np.median(rel))},
            "initial_cost_rel_diff": float(np.max(np.abs(c0 - lone_c0) / lone_c0)),
            "batched_fp64": {"value": args.steps * nw / wall64, "device_ms_per_step": dev64, "autotune": tune64,
                             "final_cost_rel_diff_max": float(rel64.max()), "phases_ms": phases64},
        }

    nbytes = int(sum(assembly_bytes(w, 4) for w in ws))
    ach = nbytes / (phases["assemble"] * 1e-3) / 1e9
    roof = {"kernel": "k_assemble<float>", "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": profile_traffic("k_assemble_c4"), "algorithmic_bytes": nbytes}
    roof.update(minimal_bytes_fields(int(sum(assembly_bytes_minimal(w, 4) for w in ws)), phases["assemble"]))
    dom = max(phases, key=phases.get)
    if dom != "assemble":
        roof["note"] = f"dominant phase is {dom}; roofline given for the assembly kernel"

    cpu = None
    if not args.no_cpu:
        from oracle import ba_oracle as O
        try:
            from threadpoolctl import threadpool_info
            cores = max([i.get("num_threads", 1) for i in threadpool_info()] + [1])
        except Exception:
            cores = 1
        sample = ws[:4]
        t0 = time.perf_counter()
        for w in sample:
            O.lm_solve(O.from_problem(w), w.X0.astype(np.float64), w.alpha0.astype(np.float64), iters, O.Options())
        tc = time.perf_counter() - t0
        cpu = {"value": len(sample) / tc, "unit": "solves/s (10 LM iterations each)", "cores": cores,
               "kind": "port", "sample": "FP64 oracle on the first 4 windows of the batch, 10 LM iterations each"}

    line = {
        "metric": "LM iters/s & batched solves/s vs plain FP64 GPU solver at matched final cost",
        "value": value,
        "unit": "solves/s (10 LM iterations each)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "device_ms_per_step": dev_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 linearisation / f64 reduced system",
        "data": "synthetic (seeded generator per window: circle trajectory r=2 m, landmarks in a 4 m box, 1 px noise)",
        "config": {"workload": f"{nw} independent windows, keyframes U{{8..32}}, landmarks U{{200..2000}}, "
                               "observations per landmark U{3..8}, 10 LM iterations per solve, Cauchy c^2=9",
                   "windows": nw, "keyframes": info["n_poses"], "landmarks": info["n_landmarks"],
                   "measurements": info["n_measurements"], "groups": info["n_groups"],
                   "max_profile_depth": info["max_depth"]},
        "autotune": tune,
        **plain,
        "phases_ms": phases,
        "roofline": roof,
        "generator_s": t_gen,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


C5_TRAIN_SIZES = (128, 192, 256, 384, 512, 768, 1024)


def bench_c5(args) -> int:
    """1024 keyframes (6144 x 6144 reduced system), 400,000 landmarks, 12
    observations each, the last 20 % of the frames revisiting the first (a
    revisited host's landmarks are observed again some 820 frames later).  The
    reduced system is solved in frame coordinates.  The preconditioner network
    is trained on about 2,000 systems the solver generates for revisit problems
    of 128-1024 keyframes; then LM runs with the Cholesky (FP64 and FP32
    linearisation), with block-Jacobi CG and with CG preconditioned by the
    trained network (tolerance 1e-8, Cholesky fallback).  A step is one LM
    iteration."""
    from spgo.lpcg import LearnedPCG, solver_systems
    from spgo.problems import make_revisit_problem
    from spgo.solver import make_options, solver_from_problem

    K = 12
    W = K - 1
    t0 = time.perf_counter()
    train_probs = [make_revisit_problem(n, 390 * n, K, seed=5100 + n) for n in C5_TRAIN_SIZES]
    log(f"training problems generated in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    per = max(1, args.c5_systems // (3 * len(train_probs)))
    systems = solver_systems(train_probs, W, states_per_problem=3, iters_between=1, restarts=per, seed=17)
    held = [make_revisit_problem(n, 390 * n, K, seed=7100 + n) for n in (160, 640)]
    held_sys = solver_systems(held, W, states_per_problem=2, iters_between=1, restarts=2, seed=23)
    t_data = time.perf_counter() - t0
    log(f"{len(systems)} training systems ({sum(h.shape[0] for h, _ in systems)} nodes) in {t_data:.1f} s")
    net = LearnedPCG(W, seed=11)
    net.load(systems)
    t0 = time.perf_counter()
    train_steps = 300
    losses = net.train(train_steps, lr=1e-3)
    t_train = time.perf_counter() - t0
    log(f"trained {train_steps} steps in {t_train:.1f} s: loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    # held-out band systems: CG iterations with the network and with block Jacobi
    net.load(held_sys)
    it_l, _ = net.solve(learned=True, tol=1e-8, max_iter=2000)[:2]
    it_b, _ = net.solve(learned=False, tol=1e-8, max_iter=2000)[:2]
    held_out = {"systems": len(held_sys), "cg_iterations_learned": [int(v) for v in it_l],
                "cg_iterations_block_jacobi": [int(v) for v in it_b]}
    log(f"held-out CG iterations: learned {held_out['cg_iterations_learned']}, "
        f"block Jacobi {held_out['cg_iterations_block_jacobi']}")
    del systems, held_sys

    p = make_revisit_problem(1024, 400000, K, seed=5005)
    X0 = p.X0.astype(np.float64)
    a0 = p.alpha0.astype(np.float64)
    runs = {}
    forms = (("cholesky", {}, False), ("cholesky_mixed", {"precision": "mixed"}, False),
             ("cg_block_jacobi", {"factorization": 4}, False),
             ("cg_learned", {"factorization": 4}, True),
             ("cg_learned_mixed", {"factorization": 4, "precision": "mixed"}, True))
    wanted = (set(args.c5_forms.split(",")) | {"cholesky"}) if args.c5_forms else None
    for tag, kw, use_net in forms:
        if wanted is not None and tag not in wanted:
            continue
        s = solver_from_problem(p, make_options(mode="lm", pcg_max_iter=args.pcg_max_iter, **kw))
        if use_net:
            s.set_preconditioner(net)
        s.step(max(args.warmup, 1))
        s.set_state(X0, a0)
        s.step(0)
        t1 = time.perf_counter()
        s.step(args.steps)
        wall = time.perf_counter() - t1
        dev = s.last_step_ms() / args.steps
        c0, recs = s.stats()
        info = s.layout_info()
        s.set_state(X0, a0)
        phases = s.profile_phases(min(args.steps, 5))
        runs[tag] = {"value": args.steps / wall, "ms_per_step": wall * 1e3 / args.steps, "device_ms_per_step": dev,
                     "cost_initial": c0, "cost_final": recs[-1]["cost"],
                     "cg_iterations": [r["pcg_iters"] for r in recs],
                     "fallbacks": int(sum(r["pcg_fallback"] for r in recs)),
                     "accepted": int(sum(r["accepted"] for r in recs)), "phases_ms": phases,
                     "factorization": info["factorization"], "coordinates": info["coordinates"]}
        log(f"{tag}: {runs[tag]['value']:.1f} LM it/s, final cost {runs[tag]['cost_final']:.10g}, "
            f"fallbacks {runs[tag]['fallbacks']}")
        s.close()
    net.close()
    base = runs["cholesky"]
    for r in runs.values():
        r["final_cost_rel_diff_vs_cholesky"] = abs(r["cost_final"] - base["cost_final"]) / base["cost_final"]
    # the headline form: the fastest run that reaches the FP64 Cholesky path's
    # final cost to 1e-6
    form = max((t for t in runs if runs[t]["final_cost_rel_diff_vs_cholesky"] <= 1e-6),
               key=lambda t: runs[t]["value"])
    best = runs[form]
    ph = best["phases_ms"]
    P = 6 * p.n_poses
    flops = P ** 3 / 3.0 + 2.0 * P * P
    chol_ms = ph["cholesky"] + ph["solve"]
    ach = flops / (chol_ms * 1e-3) / 1e12
    line = {
        "metric": "LM iters/s & batched solves/s vs plain FP64 GPU solver at matched final cost",
        "value": best["value"], "unit": "LM iterations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": best["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": best["value"] / base["value"],
        "vs_baseline_note": "ratio to the FP64 Cholesky path on the same problem",
        "dtype": "f32" if "mixed" in form else "f64", "form": form,
        "data": "synthetic (seeded revisit generator: circle r=2 m traversed 1.25 times, 1 px noise)",
        "config": {"workload": "1024 keyframes (6144 x 6144 reduced system), 400000 landmarks, 12 stereo "
                               "observations each, the last 20% of the frames revisiting the first, LM; "
                               "CG to 1e-8 with the trained network or block Jacobi, Cholesky fallback",
                   "keyframes": p.n_poses, "landmarks": p.n_landmarks, "measurements": p.n_obs,
                   "max_span": int(np.max(p.meas_t - p.meas_a))},
        "training": {"systems": int(per * 3 * len(train_probs)), "problem_keyframes": list(C5_TRAIN_SIZES),
                     "steps": train_steps, "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
                     "train_s": t_train, "data_s": t_data},
        "held_out": held_out,
        "runs": runs,
        "roofline": {"kernel": "k_chol_panel + k_chol_update + k_solve_fwd + k_solve_bwd (dense)", "bound": "mfma",
                     "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
                     "traffic": None, "algorithmic_flops": flops,
                     "note": "dense-equivalent flops P^3/3 + 2P^2 over the factorisation and solve phases"},
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
