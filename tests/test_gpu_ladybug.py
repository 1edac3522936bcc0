"""The reference's BA sample workload on the GPU: Gauss-Newton over a BA
problem on the ladybug visibility graph (49 cameras, 7 776 points, 31 843
observations; camera 0 and point 0 fixed; nelson/samples/BASchurPerfomances.cpp:236-288,
GaussNewton with minNumIt = 5), built with the sample's geometry
(nsn_generate_on_graph) from the graph in tests/golden/ladybug_49_7776.txt.
S is 288 x 288: five tile rows, the last one half full.

Against the oracle on the same inputs: linearisation 1e-12, Schur
increments 1e-9 at three dampings, the Gauss-Newton cost log 1e-9 with the
same termination reason, the LM log (cost and mu) over 12 iterations 1e-9,
the LM final cost 1e-9; mixed precision reaches it to 1e-6."""
import os

import pytest

from helpers import oracle_blocks, relerr
from nelson_amd import BAProblem, GaussNewton, GaussNewtonSettings, LevenbergMarquardt, \
    LevenbergMarquardtSettings, SolverCholeskySchur, generate_on_graph, load_visibility
from oracle.oracle import OracleBA

pytestmark = pytest.mark.gpu

LADYBUG = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ladybug_49_7776.txt")
TOL = 1e-9


def ladybug_fixture():
    return generate_on_graph(load_visibility(LADYBUG))


@pytest.fixture(scope="module")
def ladybug():
    return ladybug_fixture()


def test_ladybug_linearization_and_steps(ladybug):
    p = BAProblem.from_generated(ladybug)
    o = OracleBA.from_generated(ladybug)
    chi2 = p.update(True)
    o.update(True)
    Ho = oracle_blocks(o)
    Hg = p.export_hessian()
    assert abs(chi2 - Ho["chi2"]) <= 1e-12 * Ho["chi2"]
    for k in ("U", "V", "W", "bU", "bV"):
        assert relerr(Hg[k], Ho[k]) <= 1e-12, k
    s = SolverCholeskySchur()
    s.init(p)
    for rel, ab in ((0.0, 0.0), (0.0, 1e-3 * s.maxAbsHDiag()), (1e-2, 0.0)):
        assert s.computeIncrement(rel, ab)
        ro = o.schur_step(rel, ab)
        dxu, dxv = s.incrementVector()
        assert relerr(dxu, ro["xU"]) <= TOL and relerr(dxv, ro["xV"]) <= TOL, (rel, ab)


def test_ladybug_gauss_newton_as_the_sample(ladybug):
    """GaussNewton with minNumIt = 5 (BASchurPerfomances.cpp:278), else defaults."""
    p = BAProblem.from_generated(ladybug)
    o = OracleBA.from_generated(ladybug)
    gn = GaussNewton(GaussNewtonSettings(minNumIt=5))
    reason = gn.solve(p)
    ro = o.gn_solve(minNumIt=5)
    print(f"\nladybug GN: {reason}, {len(gn.stats())} iterations, chi2 {gn.stats()[-1][2]:.10g} "
          f"(oracle {ro['reason']}, {len(ro['log'])}, {ro['final_chi2']:.10g})")
    assert reason == ro["reason"]
    assert len(gn.stats()) == len(ro["log"])
    for a, b in zip(gn.stats(), ro["log"]):
        assert abs(a[2] - b[2]) <= TOL * abs(b[2]), (gn.stats(), ro["log"])


def test_ladybug_levenberg_marquardt_log():
    """12 iterations, before the cost reaches its rounding floor: the same log
    (cost and mu per sub-iteration) and reason."""
    p = BAProblem.from_generated(ladybug_fixture())
    lm = LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=12))
    reason = lm.solve(p)
    ro = OracleBA.from_generated(ladybug_fixture()).lm_solve(maxNumIt=12)
    assert reason == ro["reason"] == "MaxIterationReached"
    assert len(lm.stats()) == len(ro["log"])
    for a, b in zip(lm.stats(), ro["log"]):
        assert a[:2] == b[:2] and abs(a[2] - b[2]) <= TOL * b[2] and abs(a[3] - b[3]) <= TOL * abs(b[3]), (a, b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ladybug_levenberg_marquardt(ladybug, precision):
    """To convergence: the final cost.  Near the minimum the cost changes in
    its 14th digit, so which steps LM accepts there, and with them its
    termination reason, follow rounding on either side (measured: the logs
    agree to 1e-14 up to iteration 17 of 19)."""
    p = BAProblem.from_generated(ladybug, precision=precision)
    lm = LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=30))
    reason = lm.solve(p)
    ro = OracleBA.from_generated(ladybug).lm_solve(maxNumIt=30)
    c, co = lm.result()["final_chi2"], ro["final_chi2"]
    print(f"\nladybug LM {precision}: {reason} {c:.12g}, oracle {ro['reason']} {co:.12g}")
    assert abs(c - co) <= (TOL if precision == "fp64" else 1e-6) * co
