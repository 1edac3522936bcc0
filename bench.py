"""Benchmark: Levenberg-Marquardt iterations/s of the Schur-complement BA
solve on one MI355X, on the 1000-camera / 200k-point / ~1.2M-observation
ring problem (S = 6000 x 6000, ~30% dense), FP64 and mixed precision, with
the cameras in nested-dissection order inside the solver (the given order
is reported beside it).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config cfg2] [--order nd] [--no-cpu]

A "step" is one LM outer iteration (linearisation, Schur solve(s), cost
evaluation, accept/reject).  Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


# exactly K outer iterations: the rounding-level stopping tests are off (a
# negative epsilon never triggers), so that iterations/s does not depend on
# where a run happens to stop; rejected sub-iterations still count inside
# their iteration, as in the reference
def lm_settings(nsn, iters):
    return nsn.LevenbergMarquardtSettings(maxNumIt=iters, maxNumSubIt=50, epsIncSquare=-1.0, epsBVector=-1.0,
                                          epsChi2=-1.0)


def run_lm(nsn, g, precision, iters, warmup, refine=None, camera_order="given"):
    c0 = time.perf_counter()
    p = nsn.BAProblem.from_generated(g, precision=precision, camera_order=camera_order)
    create_s = time.perf_counter() - c0
    if refine is not None:
        p.set_refinement(refine)
    if warmup > 0:
        nsn.LevenbergMarquardt(lm_settings(nsn, warmup)).solve(p)
        p.set_state(g["cams0"], g["pts0"])
    lm = nsn.LevenbergMarquardt(lm_settings(nsn, iters))
    c0 = p.refinement_count() if precision == "mixed" else 0
    t0 = time.perf_counter()
    reason = lm.solve(p)
    t1 = time.perf_counter()
    r = lm.result()
    out = dict(reason=reason, iterations=r["iterations"], sub_iterations=r["sub_iterations"],
               final_chi2=r["final_chi2"], wall_s=t1 - t0, device_loop_ms=r["elapsed_ms"],
               it_per_s=r["iterations"] / (t1 - t0), ms_per_solve=1e3 * (t1 - t0) / max(r["sub_iterations"], 1),
               log=[(e[0], e[1], e[2], e[3]) for e in lm.stats()],
               create_ms=1e3 * create_s)
    if precision == "mixed":
        out["refinement_solves_per_increment"] = (p.refinement_count() - c0) / max(r["sub_iterations"], 1)
    return p, out


def chol_timing(nsn, p, reps=3):
    """Average device time of the factorisation and the other Schur phases,
    from HIP events recorded on the library's stream around each phase."""
    s = nsn.SolverCholeskySchur()
    p.update(True)
    s.init(p)
    acc = None
    for _ in range(reps):
        s.computeIncrement(0.0, 1.0)
        t = s.phase_times_ms()
        acc = t if acc is None else {k: acc[k] + t[k] for k in t}
    return {k: v / reps for k, v in acc.items()}


def sparse_timing(nsn, iters=10):
    """The sparse-S path on a banded problem: 4000 cameras along an open path,
    400k points, 2.4M observations (S of order 23 994 with about 3 % of its
    64x64 tiles inside the envelope): LM iterations/s in FP64 and mixed
    precision with the cameras shuffled and reordered by reverse
    Cuthill-McKee, the Schur phases, and the factorisation with the dense
    tile layout for comparison (NSN_DENSE_S=1)."""
    import numpy as np
    g = nsn.generate_ba(dict(nsn.CONFIGS["seq"]))
    # cameras renumbered at random (camera 0, the gauge, stays): the solver's
    # own ordering has to find the band
    rng = np.random.default_rng(3)
    C = g["n_cameras"]
    perm = np.r_[0, 1 + rng.permutation(C - 1)]
    g["cam_idx"] = perm[g["cam_idx"]].astype(np.int32)
    for k in ("cams_gt", "cams0"):
        c = np.zeros_like(g[k].reshape(C, 6))
        c[perm] = g[k].reshape(C, 6)
        g[k] = c.reshape(-1)
    out = {"workload": f"BA {C} cameras along an open path (shuffled; nested-dissection order inside the "
                       f"solver) / {g['n_points']} points / {g['n_edges']} observations, S {6 * (C - 1)}x{6 * (C - 1)}"}
    for prec in ("fp64", "mixed"):
        p = nsn.BAProblem.from_generated(g, precision=prec, camera_order="nd")
        out["tiles"], out["tiles_dense"] = p.n_tiles, p.n_tiles_dense
        out["chains"], out["chain_steps"] = p.n_chains, p.chain_steps
        _, r = run_lm(nsn, g, prec, iters, 1, camera_order="nd")
        out[prec] = {k: r[k] for k in ("it_per_s", "ms_per_solve", "iterations", "final_chi2")}
        out[prec]["phases_ms"] = chol_timing(nsn, p)
    # the same with reverse Cuthill-McKee (one chain over a narrow band) and
    # with the dense tile layout
    pr = nsn.BAProblem.from_generated(g, camera_order="rcm")
    _, r = run_lm(nsn, g, "fp64", iters, 1, camera_order="rcm")
    out["fp64_rcm"] = {"it_per_s": r["it_per_s"], "tiles": pr.n_tiles, "chain_steps": pr.chain_steps,
                       "phases_ms": chol_timing(nsn, pr)}
    old = os.environ.get("NSN_DENSE_S")
    os.environ["NSN_DENSE_S"] = "1"
    try:
        pd = nsn.BAProblem.from_generated(g, camera_order="rcm")
    finally:
        if old is None:
            del os.environ["NSN_DENSE_S"]
        else:
            os.environ["NSN_DENSE_S"] = old
    out["fp64_dense_layout_phases_ms"] = chol_timing(nsn, pd, reps=2)
    out["fp64_factorise_speedup_vs_dense_layout"] = (out["fp64_dense_layout_phases_ms"]["factorize"] /
                                                    out["fp64"]["phases_ms"]["factorize"])
    out["fp64_factorise_speedup_vs_rcm"] = (out["fp64_rcm"]["phases_ms"]["factorize"] /
                                           out["fp64"]["phases_ms"]["factorize"])
    return out


def adjoint_timing(nsn):
    """Differentiable BA (200 cameras / 20k points / 120k observations): LM to
    convergence, then the implicit-function-theorem adjoint of
    L = |t - t_gt|^2; its wall time against one forward LM iteration."""
    import numpy as np
    from nelson_amd.autograd import translation_loss_grad
    g = nsn.generate_ba(dict(nsn.CONFIGS["cfg5"]))
    p = nsn.BAProblem.from_generated(g)
    lm = nsn.LevenbergMarquardt(nsn.LevenbergMarquardtSettings(maxNumIt=50))
    t0 = time.perf_counter()
    reason = lm.solve(p)
    t1 = time.perf_counter()
    its = lm.numIterations()
    t = p.camera_translations()
    gc = translation_loss_grad(t, 2.0 * (t - g["cams_gt"].reshape(-1, 6)[:, 3:6]))
    out = {"workload": f"BA {g['n_cameras']} cameras / {g['n_points']} points / {g['n_edges']} observations, "
                       "L = |t - t_gt|^2", "lm_reason": reason, "lm_iterations": its,
           "lm_ms_per_iteration": 1e3 * (t1 - t0) / max(its, 1)}
    for steps in (0, 2):
        p.adjoint(g_cams=gc, hessian_steps=steps)  # warm
        reps = 5
        a0 = time.perf_counter()
        for _ in range(reps):
            p.adjoint(g_cams=gc, hessian_steps=steps)
        a1 = time.perf_counter()
        ms = 1e3 * (a1 - a0) / reps
        out[f"adjoint_ms_hessian_steps_{steps}"] = ms
        out[f"adjoint_over_lm_iteration_hessian_steps_{steps}"] = ms / out["lm_ms_per_iteration"]
    return out


def batch_timing(nsn, n_windows=512, iters=10):
    """Batched local BA: 512 independent windows (cameras U{5..40}, points
    U{200..3000}, 3-6 observations per point, 5% outliers at 20 px, Huber
    delta 2 px), mixed precision (no refinement, the batch default), LM
    with `iters` iterations on every window in one persistent launch; launch
    shape from the autotuner.  The same solve with refinement (at most 2
    steps per increment, tol 1e-10) and the FP64 batch's are timed beside
    it, with each mixed variant's final costs against the FP64 batch's."""
    import numpy as np
    wins = [nsn.generate_ba(nsn.window_config(w)) for w in range(n_windows)]
    t0 = time.perf_counter()
    b = nsn.BABatch(wins, precision="mixed")
    t1 = time.perf_counter()
    s = nsn.LevenbergMarquardtSettings(maxNumIt=iters)
    b.tune(s)
    cams0, pts0 = b.get_state()

    def timed(bb):
        times = []
        for _ in range(3):
            bb.set_state(cams0, pts0)
            chi2, ms = bb.lm_solve(s)
            times.append(ms)
        return chi2, min(times)

    chi2, ms = timed(b)
    reasons, its, subs = b.results()
    b.set_refinement(2, 1e-10)
    chi2_ir, ms_ir = timed(b)
    nref = b.refinement_counts()
    _, _, subs_ir = b.results()
    b64 = nsn.BABatch(wins, precision="fp64")
    b64.set_config(*[b.config()[k] for k in ("variant", "grid")])
    chi2_64, ms_64 = timed(b64)
    return {"workload": f"{n_windows} windows, {sum(g['n_cameras'] for g in wins)} cameras / "
                        f"{sum(g['n_points'] for g in wins)} points / {sum(g['n_edges'] for g in wins)} observations, "
                        f"mixed precision, Huber 2 px, LM {iters} iterations",
            "batch_solves_per_s": 1e3 / ms, "windows_per_s": n_windows * 1e3 / ms, "ms_per_batch_solve": ms,
            "mean_iterations": float(np.mean(its)), "mean_sub_iterations": float(np.mean(subs)),
            "final_cost_max_rel_err_vs_fp64": float(np.max(np.abs(chi2 - chi2_64) / chi2_64)),
            "refined": {"ms_per_batch_solve": ms_ir,
                        "refinement_solves_per_increment": float(np.sum(nref) / max(1, np.sum(subs_ir))),
                        "final_cost_max_rel_err_vs_fp64": float(np.max(np.abs(chi2_ir - chi2_64) / chi2_64))},
            "fp64_ms_per_batch_solve": ms_64,
            "config": b.config(), "create_s": t1 - t0}


LADYBUG = os.path.join(ROOT, "tests", "golden", "ladybug_49_7776.txt")


def ladybug_problem(nsn):
    """The reference's BA sample workload: a BA problem on the ladybug
    visibility graph (49 cameras / 7 776 points / 31 843 observations) with
    the sample's geometry (nelson/samples/BASchurPerfomances.cpp:40-92,236-248)."""
    return nsn.generate_on_graph(nsn.load_visibility(LADYBUG))


def ladybug_timing(nsn, g):
    """Gauss-Newton as the sample runs it (minNumIt = 5, other settings at
    their defaults; BASchurPerfomances.cpp:266-280) in FP64, and LM to
    convergence in FP64 and mixed precision, on the GPU."""
    def timed(make, precision):
        p = nsn.BAProblem.from_generated(g, precision=precision)
        make().solve(p)  # warm: graphs captured, buffers allocated
        p.set_state(g["cams0"], g["pts0"])
        d = make()
        t0 = time.perf_counter()
        reason = d.solve(p)
        t1 = time.perf_counter()
        its = d.numIterations()
        return dict(reason=reason, iterations=its, final_chi2=d.result()["final_chi2"],
                    wall_s=t1 - t0, it_per_s=its / (t1 - t0))
    gn = timed(lambda: nsn.GaussNewton(nsn.GaussNewtonSettings(minNumIt=5)), "fp64")
    lm64 = timed(lambda: nsn.LevenbergMarquardt(nsn.LevenbergMarquardtSettings(maxNumIt=30)), "fp64")
    lmm = timed(lambda: nsn.LevenbergMarquardt(nsn.LevenbergMarquardtSettings(maxNumIt=30)), "mixed")
    return {"workload": f"BA on the ladybug visibility graph, {g['n_cameras']} cameras / {g['n_points']} points / "
                        f"{g['n_edges']} observations, S 288x288", "gn_fp64": gn, "lm_fp64": lm64, "lm_mixed": lmm}


def measured_peaks(nsn):
    """The device's own ceilings: FP64 / FP32 16x16x4 MFMA rate on register
    operands and a 1 GiB streaming copy (nsn_measure_peaks)."""
    import ctypes as C
    lib = nsn.capi.lib()
    f64, f32, bw = C.c_double(0), C.c_double(0), C.c_double(0)
    nsn.capi.check(lib.nsn_measure_peaks(0, C.byref(f64), C.byref(f32), C.byref(bw)))
    return {"mfma_f64_tflops": f64.value, "mfma_f32_tflops": f32.value, "copy_gbps": bw.value,
            "datasheet": {"mfma_f64_tflops": FP64_PEAK_TFLOPS, "mfma_f32_tflops": 157.3, "hbm_gbps": 8000.0}}


def cpu_baseline_ladybug(g):
    """The oracle (one thread) running the sample's Gauss-Newton on the
    ladybug workload end to end."""
    from oracle.oracle import OracleBA
    o = OracleBA.from_generated(g)
    t0 = time.perf_counter()
    r = o.gn_solve(minNumIt=5)
    t1 = time.perf_counter()
    its = max(len(r["log"]) - 1, 1)
    return dict(value=its / (t1 - t0), unit="GN iter/s", cores=1, kind="port", iterations=its,
                reason=r["reason"], final_chi2=r["final_chi2"],
                sample=f"the whole Gauss-Newton solve ({t1 - t0:.2f} s) on {cpu_model()}, 1 thread")


def cpu_baseline(g, budget_s):
    """The oracle (C restatement of the reference path) timed on the host for
    one complete LM sub-iteration of the same workload: update(true), the
    Schur step (V solves, bS, S, pivoted LDL^T of the dense S, solves,
    back-substitution) and update(false).  With all host cores the LDL^T row
    updates run on every core (OpenMP; the rest of the oracle is serial) and
    the whole step is timed; with one thread (the reference's default,
    nelson/ParallelExecHelper.h:23) the LDL^T is timed on a leading k x k
    block of the same S and scaled by (n/k)^3 to stay within the budget."""
    import numpy as np
    from oracle import oracle as orc
    from oracle.oracle import OracleBA, ldlt
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    o = OracleBA.from_generated(g)
    # all cores
    cores = orc.lib().or_set_threads(nproc)
    t0 = time.perf_counter()
    o.update(True)
    t1 = time.perf_counter()
    st = o.schur_step(0.0, 1.0)
    t2 = time.perf_counter()
    o.update(False)
    t3 = time.perf_counter()
    per_it_all = t3 - t0
    # one thread: the serial parts measured above, the LDL^T extrapolated
    orc.lib().or_set_threads(1)
    r = o.schur_assemble(0.0, 1.0)
    S = r["S"]
    n = S.shape[0]
    k0 = min(n, 800)
    a = time.perf_counter()
    ldlt(np.ascontiguousarray(S[:k0, :k0]))
    t_k0 = time.perf_counter() - a
    target = max(1.0, 0.5 * budget_s)
    k = int(min(n, k0 * (target / max(t_k0, 1e-6)) ** (1.0 / 3.0)))
    a = time.perf_counter()
    ldlt(np.ascontiguousarray(S[:k, :k]))
    t_k = time.perf_counter() - a
    t_fact1 = t_k * (n / k) ** 3
    a = time.perf_counter()
    o.schur_assemble(0.0, 1.0)
    t_asm = time.perf_counter() - a
    per_it_1 = (t1 - t0) + (t3 - t2) + t_asm + t_fact1
    orc.lib().or_set_threads(nproc)
    assert st["ok"]
    return dict(value=1.0 / per_it_all, unit="LM iter/s", cores=cores, kind="port",
                sample=f"one LM sub-iteration of the same workload on {cpu_model()} ({cores} threads): "
                       f"update(true) {t1 - t0:.2f} s, Schur step (assembly, dense LDL^T of S n={n} on all cores, "
                       f"solves, back-substitution) {t2 - t1:.2f} s, update(false) {t3 - t2:.2f} s",
                one_thread={"value": 1.0 / per_it_1, "unit": "LM iter/s", "cores": 1,
                            "sample": f"same sub-iteration on one thread: updates {t1 - t0 + t3 - t2:.2f} s and "
                                      f"Schur assembly {t_asm:.2f} s measured, LDL^T {t_fact1:.1f} s extrapolated "
                                      f"as (n/k)^3 from the leading k={k} block ({t_k:.2f} s), solves not "
                                      f"included"})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--order", default="nd", choices=("given", "rcm", "nd"),
                    help="camera order inside the solver (nd: nested dissection, independent chains)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-sparse", action="store_true")
    ap.add_argument("--no-given", action="store_true",
                    help="skip the FP64 LM with the cameras in the given order (reported beside --order)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--only-ladybug", action="store_true",
                    help="only the BA sample's workload on the ladybug graph (its own JSON line)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of 1 GiB 5 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        a.fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": a.numel() * 8, "launches": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    import nelson_amd as nsn
    if args.only_ladybug:
        lb = ladybug_problem(nsn)
        out = ladybug_timing(nsn, lb)
        if not args.no_cpu:
            out["cpu_baseline"] = cpu_baseline_ladybug(lb)
        print(json.dumps(out))
        return
    cfg = dict(nsn.CONFIGS[args.config])
    g = nsn.generate_ba(cfg)
    log(f"bench: {args.config} C={g['n_cameras']} P={g['n_points']} E={g['n_edges']}")
    p64, r64 = run_lm(nsn, g, "fp64", args.steps, args.warmup, camera_order=args.order)
    log(f"bench: fp64 {r64}")
    ph = chol_timing(nsn, p64)
    log(f"bench: fp64 phases (ms) {ph}")
    given = None
    if args.order != "given" and not args.no_given:
        # the same FP64 LM with the cameras in the given order (one chain)
        pgv, rgv = run_lm(nsn, g, "fp64", args.steps, args.warmup)
        given = {"it_per_s": rgv["it_per_s"], "ms_per_solve": rgv["ms_per_solve"], "final_chi2": rgv["final_chi2"],
                 "tiles": pgv.n_tiles, "chain_steps": pgv.chain_steps, "phases_ms": chol_timing(nsn, pgv)}
        log(f"bench: fp64 given order {given}")
    n = p64.n_s
    # the factorisation's operations over the tile envelope of S (fewer than
    # the dense n^3/3 where tiles are zero; the library counts them)
    flops = p64.factor_flops
    chol_s = ph["factorize"] / 1e3
    achieved = flops / chol_s / 1e12
    traffic = None
    tname = f"traffic_k_chol_{args.config}_{args.order}_fp64.json"
    tfile = os.path.join(ROOT, "profiles", tname)
    if os.path.exists(tfile):
        with open(tfile) as f:
            t = json.load(f)
        traffic = t["hbm_read_bytes_per_launch"] + t["hbm_written_bytes_per_launch"]
    peaks = measured_peaks(nsn)
    log(f"bench: measured peaks {peaks}")
    roofline = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                "traffic_note": "HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated), from "
                                "profiles/" + tname,
                "kernel": "k_chol<double,64> (tile Cholesky of S over its tile envelope, n=%d, %d of %d tiles, "
                          "%.3g flop, %d diagonal chains, critical path %d of %d steps)"
                          % (n, p64.n_tiles, p64.n_tiles_dense, flops, p64.n_chains, p64.chain_steps,
                             (n + 64) // 64),
                "kernel_ms": ph["factorize"],
                "dense_equivalent_tflops": n ** 3 / 3.0 / chol_s / 1e12}
    mixed = None
    mixed_small = None
    mixed_ir1 = None
    mixed_ir0 = None
    if not args.no_mixed:
        # mixed precision (FP32 blocks / S / Cholesky, FP64 pivot tiles, cost, gain
        # ratio and the default iterative refinement: at most 2 steps, each
        # stopping on the device once the FP64 residual is at 1e-10 of the
        # gradient) on the same workload, and on the 10-camera / 500-point
        # problem where it is defined (20 iterations)
        pm, rm = run_lm(nsn, g, "mixed", args.steps, args.warmup, camera_order=args.order)
        log(f"bench: mixed {rm}")
        mixed = dict(rm, refinement="at most 2 steps, stopping once max|r| <= 1e-10 max|b| (FP64 residual)",
                     final_cost_rel_err_vs_fp64=abs(rm["final_chi2"] - r64["final_chi2"]) / r64["final_chi2"],
                     phases_ms=chol_timing(nsn, pm))
        # the same with one and with no refinement step (the library's
        # setting nsn_set_refinement): the speed / accuracy trade-off; every
        # one of them matches the FP64 final cost far below 1e-6
        def variant(k):
            _, rk = run_lm(nsn, g, "mixed", args.steps, args.warmup, refine=k, camera_order=args.order)
            return {"refinement_steps": k, "it_per_s": rk["it_per_s"], "ms_per_solve": rk["ms_per_solve"],
                    "it_per_s_ratio_vs_fp64": rk["it_per_s"] / r64["it_per_s"],
                    "final_cost_rel_err_vs_fp64": abs(rk["final_chi2"] - r64["final_chi2"]) / r64["final_chi2"]}
        mixed_ir1 = variant(1)
        log(f"bench: mixed, one refinement step {mixed_ir1}")
        mixed_ir0 = variant(0)
        log(f"bench: mixed, no refinement {mixed_ir0}")
        # the 10-camera / 500-point problem of the mixed-precision
        # configuration (20 iterations)
        gs = nsn.generate_ba(dict(nsn.CONFIGS["cfg1"]))
        _, s64 = run_lm(nsn, gs, "fp64", 20, 2)
        _, sm = run_lm(nsn, gs, "mixed", 20, 2)
        mixed_small = {"workload": "BA 10 cameras / 500 points / 2000 observations, LM 20 iterations",
                       "fp64_it_per_s": s64["it_per_s"], "mixed_it_per_s": sm["it_per_s"],
                       "ratio": sm["it_per_s"] / s64["it_per_s"],
                       "fp64_final_chi2": s64["final_chi2"], "mixed_final_chi2": sm["final_chi2"],
                       "final_cost_rel_err": abs(sm["final_chi2"] - s64["final_chi2"]) / s64["final_chi2"]}
        log(f"bench: mixed small {mixed_small}")
    batch = None
    if not args.no_batch:
        batch = batch_timing(nsn)
        log(f"bench: batch {batch}")
    sparse = None
    if not args.no_sparse:
        sparse = sparse_timing(nsn)
        log(f"bench: sparse {sparse}")
    lb = ladybug_problem(nsn)
    ladybug = ladybug_timing(nsn, lb)
    log(f"bench: ladybug {ladybug}")
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_timing(nsn)
        log(f"bench: adjoint {adjoint}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(g, args.cpu_budget)
        ladybug["cpu_baseline"] = cpu_baseline_ladybug(lb)
        ladybug["gn_gpu_over_cpu"] = ladybug["gn_fp64"]["it_per_s"] / ladybug["cpu_baseline"]["value"]
        log(f"bench: cpu {cpu} ladybug {ladybug['cpu_baseline']}")
    its = r64["iterations"]
    value = its / r64["wall_s"]
    mixed_vs_fp64 = None
    if mixed is not None:
        mixed_vs_fp64 = {"it_per_s_ratio": mixed["it_per_s"] / value,
                         "final_cost_rel_err": mixed["final_cost_rel_err_vs_fp64"],
                         "fp64_ms_per_solve": r64["ms_per_solve"], "mixed_ms_per_solve": mixed["ms_per_solve"],
                         "refinement_solves_per_increment": mixed["refinement_solves_per_increment"],
                         "one_step_ratio": mixed_ir1["it_per_s_ratio_vs_fp64"],
                         "one_step_final_cost_rel_err": mixed_ir1["final_cost_rel_err_vs_fp64"],
                         "no_refinement_ratio": mixed_ir0["it_per_s_ratio_vs_fp64"],
                         "no_refinement_final_cost_rel_err": mixed_ir0["final_cost_rel_err_vs_fp64"],
                         "small_problem_ratio": mixed_small["ratio"],
                         "small_problem_final_cost_rel_err": mixed_small["final_cost_rel_err"]}
    out = {
        "metric": "LM iter/s at matched final cost: mixed-precision vs FP64 GPU; batch solves/s",
        "value": value,
        "unit": "LM iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r64["wall_s"] / max(its, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"BA {g['n_cameras']} cameras / {g['n_points']} points / "
                               f"{g['n_edges']} observations (ring of cameras), S {n}x{n}, FP64 Schur LM",
                   "preset": args.config, "camera_order": args.order,
                   "cameras": g["n_cameras"], "points": g["n_points"], "observations": g["n_edges"],
                   "lm_iterations": args.steps},
        "mixed_vs_fp64": mixed_vs_fp64,
        "fp64": r64,
        "phases_ms": ph,
        "fp64_given_order": given,
        "measured_peaks": peaks,
        "mixed": mixed,
        "mixed_small": mixed_small,
        "mixed_ir1": mixed_ir1,
        "mixed_ir0": mixed_ir0,
        "adjoint": adjoint,
        "batch": batch,
        "sparse": sparse,
        "ladybug": ladybug,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
