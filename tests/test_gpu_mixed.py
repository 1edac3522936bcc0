"""Mixed precision (FP32 Jacobians, Hessian blocks, S assembly and Cholesky
with FP64 pivot tiles; FP64 cost, gain ratio and iterative-refinement
residual) on Config 1's problem, against the FP64 oracle.

Reported figures (printed with -s): the increment error after 0..3
refinement steps, the increment error along an LM trajectory and the final
cost error.  Bars: with 2 refinement steps, and with the default adaptive
refinement (at most 2 steps, stopping on the device once the residual is at
1e-10 of the gradient), the increment meets the FP64 parity bar (1e-9); the
mixed LM reaches the FP64 final cost to 1e-6.
"""
import numpy as np
import pytest

from nelson_amd import BAProblem, LevenbergMarquardt, LevenbergMarquardtSettings, SolverCholeskySchur
from oracle.oracle import OracleBA

from helpers import relerr

pytestmark = pytest.mark.gpu


def _step(p, rel, ab):
    s = SolverCholeskySchur()
    s.init(p)
    ok = s.computeIncrement(rel, ab)
    return ok, s.incrementVector()


@pytest.mark.parametrize("ab", [1e-3, 1.0, 62.0])
def test_refinement_recovers_fp64_increment(cfg1, ab):
    o = OracleBA.from_generated(cfg1)
    o.update(True)
    ro = o.schur_step(0.0, ab)
    errs = []
    for ir in range(4):
        p = BAProblem.from_generated(cfg1, precision="mixed")
        p.set_refinement(ir)
        p.update(True)
        ok, (dxu, dxv) = _step(p, 0.0, ab)
        assert ok
        errs.append(max(relerr(dxu, ro["xU"]), relerr(dxv, ro["xV"])))
    print(f"\nmixed increment error vs FP64 oracle (abs damping {ab:g}), refinement steps 0..3: {errs}")
    assert errs[0] < 1e-3  # FP32 solve alone
    assert errs[1] < errs[0]
    assert errs[2] <= 1e-9  # the default: FP64 parity bar
    assert errs[3] <= 1e-9


def test_mixed_increment_along_trajectory(cfg1):
    """Per-iteration increment error: at each state of a 10-step trajectory
    (taken with the mixed increments, damping halved per step), the mixed
    increment and the FP64 GPU increment vs the oracle's FP64 increment at
    the same state and damping.  Near the minimum b is at rounding level and
    any two FP64 implementations differ by more than 1e-9 relative, so the
    bar is: mixed error <= max(1e-9, 10 x the FP64 GPU path's error), with
    two refinement steps."""
    pm = BAProblem.from_generated(cfg1, precision="mixed")
    pm.set_refinement(2)
    p64 = BAProblem.from_generated(cfg1)
    o = OracleBA.from_generated(cfg1)
    mu = 1.0
    rows = []
    for it in range(10):
        cams, pts = pm.get_state_rt()
        o.set_state_rt(cams, pts)
        p64.set_state_rt(cams, pts)
        pm.update(True)
        p64.update(True)
        o.update(True)
        okm, (um, vm) = _step(pm, 0.0, mu)
        ok64, (u64, v64) = _step(p64, 0.0, mu)
        ro = o.schur_step(0.0, mu)
        assert okm and ok64 and ro["ok"]
        em = max(relerr(um, ro["xU"]), relerr(vm, ro["xV"]))
        e64 = max(relerr(u64, ro["xU"]), relerr(v64, ro["xV"]))
        rows.append((em, e64, float(np.sqrt(np.sum(ro["xU"] ** 2) + np.sum(ro["xV"] ** 2)))))
        pm.oplus()
        mu *= 0.5
    print("\nper iteration: mixed err, FP64-GPU err, |dx| oracle: " +
          ", ".join("(%.1e, %.1e, %.1e)" % r for r in rows))
    for em, e64, _ in rows:
        assert em <= max(1e-9, 10 * e64)


def test_mixed_lm_final_cost(cfg1):
    """Config 3: LM 20 iterations in mixed precision reaches the FP64
    oracle's final cost to 1e-6 (in practice far below)."""
    o = OracleBA.from_generated(cfg1)
    ro = o.lm_solve(maxNumIt=20)
    for ir in (0, 2):
        p = BAProblem.from_generated(cfg1, precision="mixed")
        p.set_refinement(ir)
        lm = LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=20))
        lm.solve(p)
        err = abs(lm.result()["final_chi2"] - ro["final_chi2"]) / ro["final_chi2"]
        print(f"\nmixed LM ({ir} refinement steps): final cost {lm.result()['final_chi2']:.10g} vs oracle "
              f"{ro['final_chi2']:.10g}, rel err {err:.2e}, iterations {lm.numIterations()} vs {ro['iterations']}")
        assert err <= 1e-6


def test_refinement_without_variable_points(cfg1):
    """Every point fixed (no V blocks, S = U): the refinement's camera update
    then runs without the back-substitution kernel.  After two refinement
    steps the increment meets the FP64 bar against the oracle, and the mixed
    LM reaches the oracle's final cost (bar 1e-6)."""
    pf = np.ones(cfg1["n_points"], np.uint8)
    o = OracleBA.from_generated(cfg1, pt_fixed=pf)
    o.update(True)
    ro = o.schur_step(0.0, 1.0)
    p = BAProblem.from_generated(cfg1, precision="mixed", pt_fixed=pf)
    p.update(True)
    ok, (dxu, dxv) = _step(p, 0.0, 1.0)
    assert ok and np.asarray(dxv).size == 0
    err = relerr(dxu, ro["xU"])
    print(f"no variable points: increment error {err:.2e}")
    assert err <= 1e-9
    lm = LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=30))
    lm.solve(p)
    rl = OracleBA.from_generated(cfg1, pt_fixed=pf).lm_solve(maxNumIt=30)
    cost = lm.result()["final_chi2"]
    assert abs(cost - rl["final_chi2"]) <= 1e-6 * rl["final_chi2"], (cost, rl["final_chi2"])


@pytest.mark.parametrize("ab", [1e-3, 1.0])
def test_adaptive_refinement(cfg1, ab):
    """The default refinement stops on the device once the FP64 residual's
    largest entry is at most 1e-10 of the gradient's (at most 2 steps): it
    runs 1..2 solves per increment and meets the FP64 parity bar (1e-9)
    against the oracle.  A tolerance above the FP32 solve's residual runs no
    solve and gives the unrefined increment bit for bit; tolerance 0 gives
    the fixed-step increment bit for bit."""
    o = OracleBA.from_generated(cfg1)
    o.update(True)
    ro = o.schur_step(0.0, ab)

    def run(setup):
        p = BAProblem.from_generated(cfg1, precision="mixed")
        setup(p)
        p.update(True)
        c0 = p.refinement_count()
        ok, (dxu, dxv) = _step(p, 0.0, ab)
        assert ok
        return np.concatenate([dxu, dxv]), p.refinement_count() - c0

    x, n = run(lambda p: None)
    e = relerr(x, np.concatenate([ro["xU"], ro["xV"]]))
    print(f"\nadaptive refinement (abs damping {ab:g}): {n} solves, increment error vs oracle {e:.2e}")
    assert 1 <= n <= 2 and e <= 1e-9, (n, e)
    x0, n0 = run(lambda p: p.set_refinement(0))
    xl, nl = run(lambda p: p.set_refinement(3, tol=1e-2))
    assert (n0, nl) == (0, 0) and np.array_equal(x0, xl)
    x2, n2 = run(lambda p: p.set_refinement(2))
    x2t, n2t = run(lambda p: p.set_refinement(2, tol=0.0))
    assert (n2, n2t) == (2, 2) and np.array_equal(x2, x2t)
