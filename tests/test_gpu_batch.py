"""Batched independent windows (nsn_batch_*): one persistent launch runs the
LM solve of every window.  FP64 batch results match the single-problem GPU
path and the oracle (final cost 1e-9, same termination); the mixed-precision
batch reaches the FP64 final cost to 1e-6, and with two refinement steps
follows the FP64 trajectory (1e-9); the autotuner's pick and the state
round trip behave; solves replay one captured graph per launch shape.
"""
import numpy as np
import pytest

from nelson_amd import BABatch, BAProblem, CONFIGS, LevenbergMarquardt, LevenbergMarquardtSettings, generate_ba, \
    window_config
from oracle.oracle import OracleBA

pytestmark = pytest.mark.gpu


def _windows(n, seed=0, huber=False):
    rng = np.random.default_rng(seed)
    out = []
    for w in range(n):
        nc = int(rng.integers(3, 20))
        npt = int(rng.integers(30, 400))
        cfg = dict(CONFIGS["cfg1"], n_cameras=nc, n_points=npt, k_min=3, k_max=min(6, nc), seed=500 + 17 * w + seed)
        if huber:
            cfg.update(outlier_frac=0.05, outlier_sigma=20.0, huber_delta=2.0)
        out.append(generate_ba(cfg))
    return out


def _single(g, precision="fp64", **kw):
    p = BAProblem.from_generated(g, precision=precision)
    lm = LevenbergMarquardt(LevenbergMarquardtSettings(**kw))
    reason = lm.solve(p)
    return lm.result()["final_chi2"], reason, lm.numIterations()


@pytest.mark.parametrize("huber", [False, True])
def test_batch_fp64_matches_single_problem(huber):
    wins = _windows(24, seed=1 if huber else 0, huber=huber)
    b = BABatch(wins, precision="fp64")
    kw = dict(maxNumIt=15)
    chi2, ms = b.lm_solve(LevenbergMarquardtSettings(**kw))
    reasons, its, _ = b.results()
    for w, g in enumerate(wins):
        c1, r1, i1 = _single(g, **kw)
        assert abs(chi2[w] - c1) <= 1e-9 * c1, (w, chi2[w], c1)
        conv = {"SuccessSmallIncrement", "MaxIterationReached", "MaxSubIterationReached"}
        assert reasons[w] == r1 or {reasons[w], r1} <= conv, (w, reasons[w], r1)
    print(f"\nbatch of 24 windows (huber={huber}): {ms:.2f} ms, iterations {list(its)}")


def test_batch_vs_oracle():
    wins = _windows(6, seed=3)
    b = BABatch(wins, precision="fp64")
    chi2, _ = b.lm_solve(LevenbergMarquardtSettings(maxNumIt=20))
    for w, g in enumerate(wins):
        ro = OracleBA.from_generated(g).lm_solve(maxNumIt=20)
        assert abs(chi2[w] - ro["final_chi2"]) <= 1e-9 * ro["final_chi2"], (w, chi2[w], ro["final_chi2"])


def test_batch_mixed_reaches_fp64_cost():
    wins = [generate_ba(window_config(w)) for w in range(8)]
    s = LevenbergMarquardtSettings(maxNumIt=50)
    c64, _ = BABatch(wins, precision="fp64").lm_solve(s)
    bm = BABatch(wins, precision="mixed")
    cm, ms = bm.lm_solve(s)
    err = np.abs(cm - c64) / c64
    print(f"\nmixed vs FP64 batch (8 windows of the local-BA workload): max final-cost rel err {err.max():.2e}, "
          f"config {bm.config()}")
    assert err.max() <= 1e-6


@pytest.mark.parametrize("variant", [8, 16, 64])
def test_batch_mixed_refinement(variant):
    """Mixed-precision windows with iterative refinement: two fixed steps per
    increment (tol 0) make a 6-iteration LM trajectory equal the FP64
    batch's -- final costs to 1e-9, states to 1e-8, the same iteration and
    solve counts -- where the unrefined FP32 increments differ visibly; every
    solve runs exactly two refinement solves, none by default (refinement
    off), and the adaptive form (2 steps, tol 1e-10) runs at most two."""
    rng = np.random.default_rng(21)
    wins = []
    for w, nc in enumerate([5, 9, 14, 22, 31, 42]):
        npt = int(rng.integers(150, 600))
        wins.append(generate_ba(dict(CONFIGS["cfg1"], n_cameras=nc, n_points=npt, k_min=3, k_max=6, seed=1300 + w)))
    s = LevenbergMarquardtSettings(maxNumIt=6)
    b64 = BABatch(wins, precision="fp64")
    b64.set_config(variant, 6)
    c64, _ = b64.lm_solve(s)
    _, it64, sub64 = b64.results()
    cams64, pts64 = b64.get_state()
    errs = {}
    for mode in ("off", "fixed", "adaptive"):
        bm = BABatch(wins, precision="mixed")
        bm.set_config(variant, 6)
        if mode == "fixed":
            bm.set_refinement(2, 0.0)
        elif mode == "adaptive":
            bm.set_refinement(2, 1e-10)
        cm, ms = bm.lm_solve(s)
        _, itm, subm = bm.results()
        nref = bm.refinement_counts()
        cams, pts = bm.get_state()
        errs[mode] = float(np.max(np.abs(cm - c64) / c64))
        print(f"\nvariant {variant}, refinement {mode}: final cost vs FP64 {errs[mode]:.2e}, "
              f"state {max(np.abs(cams - cams64).max(), np.abs(pts - pts64).max()):.2e}, "
              f"refinement solves {list(nref)} for {list(subm)} solves, {ms:.2f} ms")
        if mode == "off":
            assert np.all(nref == 0)
            continue
        if mode == "fixed":
            assert np.array_equal(nref, 2 * subm), (nref, subm)
            assert errs[mode] <= 1e-9, errs[mode]
            assert np.array_equal(itm, it64) and np.array_equal(subm, sub64)
            scale = max(np.abs(cams64).max(), np.abs(pts64).max())
            assert np.abs(cams - cams64).max() <= 1e-8 * scale and np.abs(pts - pts64).max() <= 1e-8 * scale
        else:
            assert np.all(nref <= 2 * subm) and errs[mode] <= 1e-8, (nref, subm, errs[mode])
    assert errs["fixed"] <= errs["off"]


def test_batch_tune_and_state_roundtrip():
    wins = _windows(10, seed=5)
    b = BABatch(wins, precision="mixed")
    cams0, pts0 = b.get_state()
    # angle-axis read back through the rotation matrix, as BAProblem does
    ref = np.concatenate([BAProblem.from_generated(g).get_state()[0] for g in wins])
    assert np.allclose(cams0, ref, rtol=0, atol=1e-12)
    assert np.array_equal(pts0, np.concatenate([g["pts0"] for g in wins]))
    s = LevenbergMarquardtSettings(maxNumIt=10)
    nb, grid = b.tune(s)
    assert nb in (8, 16, 64) and grid in (256, 512, 1024)
    cams1, pts1 = b.get_state()
    assert np.array_equal(cams1, cams0) and np.array_equal(pts1, pts0)  # the tuner restores the state
    c_a, _ = b.lm_solve(s)
    b.set_state(cams0, pts0)
    b.set_config(8, 3)
    c_b, _ = b.lm_solve(s)
    assert np.allclose(c_a, c_b, rtol=1e-6)


def test_batch_variants_agree():
    """The tile (64) and panel (8, 16) factorisation variants give the same
    FP64 solves, on windows of one to four 64-row tiles."""
    rng = np.random.default_rng(7)
    wins = []
    for w, nc in enumerate([4, 11, 12, 25, 40, 43]):
        npt = int(rng.integers(100, 500))
        wins.append(generate_ba(dict(CONFIGS["cfg1"], n_cameras=nc, n_points=npt, k_min=3, k_max=6, seed=900 + w)))
    s = LevenbergMarquardtSettings(maxNumIt=8)
    res = {}
    for variant in (8, 16, 64):
        b = BABatch(wins, precision="fp64")
        b.set_config(variant, 4)
        res[variant], _ = b.lm_solve(s)
    for variant in (8, 16):
        assert np.allclose(res[variant], res[64], rtol=1e-9, atol=0), (variant, res[variant], res[64])
    for w, g in enumerate(wins):
        c1, _, _ = _single(g, maxNumIt=8)
        assert abs(res[64][w] - c1) <= 1e-9 * c1


def test_batch_graph_replays():
    """Every solve replays the same captured graph (counter reset + persistent
    kernel); replays separated by host work and state uploads give the same
    results bit for bit, and each matches the single-problem path."""
    wins = _windows(10, seed=5, huber=True)
    b = BABatch(wins, precision="fp64")
    b.set_config(16, 8)
    s = LevenbergMarquardtSettings(maxNumIt=6)
    cams0, pts0 = b.get_state()
    ref = [_single(g, maxNumIt=6)[0] for g in wins]
    first = None
    for rep in range(3):
        b.set_state(cams0, pts0)
        chi2, _ = b.lm_solve(s)
        cams, pts = b.get_state()
        if first is None:
            first = (chi2.copy(), cams.copy(), pts.copy())
        else:
            assert np.array_equal(chi2, first[0]) and np.array_equal(cams, first[1]) and np.array_equal(pts, first[2])
        for w in range(len(wins)):
            assert abs(chi2[w] - ref[w]) <= 1e-9 * ref[w], (rep, w, chi2[w], ref[w])


@pytest.mark.parametrize("variant", [8, 64])
def test_batch_degenerate_windows(variant):
    """Windows with no variable camera (S of order 0), with one variable
    camera and single observations, beside an ordinary one: the same final
    cost and termination as the single-problem path."""
    wins = []
    for w, (nc, npt) in enumerate([(1, 30), (2, 20), (5, 60)]):
        wins.append(generate_ba(dict(CONFIGS["cfg1"], n_cameras=nc, n_points=npt, k_min=1 if nc < 3 else 3,
                                     k_max=min(nc, 4), seed=300 + w)))
    b = BABatch(wins, precision="fp64")
    b.set_config(variant, 3)
    chi2, _ = b.lm_solve(LevenbergMarquardtSettings(maxNumIt=5))
    reasons, _, _ = b.results()
    for w, g in enumerate(wins):
        c1, r1, _ = _single(g, maxNumIt=5)
        assert abs(chi2[w] - c1) <= 1e-9 * c1, (w, chi2[w], c1)
        assert reasons[w] == r1, (w, reasons[w], r1)


def test_config4_all_512_windows():
    """The local-BA batch at its full size: all 512 windows of window_config
    (cameras U{5..40}, points U{200..3000}, 3-6 observations per point, 5 %
    outliers at 20 px, Huber 2 px) in one launch per solve.
      * FP64 batch, the workload's 10 LM iterations: every window's final cost
        equals the single-problem FP64 GPU path's (1e-9) with the same
        termination;
      * run to convergence: the mixed-precision batch reaches every window's
        FP64 final cost to 1e-6;
      * the oracle's LM on 8 windows, the largest among them: 1e-9."""
    wins = [generate_ba(window_config(w)) for w in range(512)]
    sizes = [(g["n_cameras"], g["n_points"], g["n_edges"]) for g in wins]
    print(f"\n512 windows: {sum(s[2] for s in sizes)} observations, largest {max(sizes, key=lambda s: s[2])}",
          flush=True)
    s10 = LevenbergMarquardtSettings(maxNumIt=10)
    b64 = BABatch(wins, precision="fp64")
    c10, ms10 = b64.lm_solve(s10)
    reasons, _, _ = b64.results()
    conv = {"SuccessSmallIncrement", "MaxIterationReached", "MaxSubIterationReached"}
    worst = 0.0
    for w, g in enumerate(wins):
        c1, r1, _ = _single(g, maxNumIt=10)
        worst = max(worst, abs(c10[w] - c1) / c1)
        assert abs(c10[w] - c1) <= 1e-9 * c1, (w, c10[w], c1)
        assert reasons[w] == r1 or {reasons[w], r1} <= conv, (w, reasons[w], r1)
    print(f"FP64 batch (10 iterations, {ms10:.1f} ms) vs single-problem path: max rel err {worst:.2e}", flush=True)
    # to convergence: mixed vs FP64
    sconv = LevenbergMarquardtSettings(maxNumIt=60)
    b64.set_state(*BABatch(wins, precision="fp64").get_state())
    cf64, _ = b64.lm_solve(sconv)
    bm = BABatch(wins, precision="mixed")
    cm, msm = bm.lm_solve(sconv)
    err = np.abs(cm - cf64) / cf64
    print(f"mixed vs FP64 batch to convergence: max final-cost rel err {err.max():.2e} "
          f"(median {np.median(err):.2e}), mixed config {bm.config()}", flush=True)
    assert err.max() <= 1e-6
    # the oracle on 8 windows including the largest
    big = int(np.argmax([s[2] for s in sizes]))
    pick = [big] + [int(w) for w in np.random.default_rng(4).choice(512, 7, replace=False) if w != big][:7]
    for w in pick:
        ro = OracleBA.from_generated(wins[w]).lm_solve(maxNumIt=60)
        assert abs(cf64[w] - ro["final_chi2"]) <= 1e-9 * ro["final_chi2"], (w, cf64[w], ro["final_chi2"])
