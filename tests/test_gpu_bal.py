"""The BAL camera model on the GPU (per-camera f, k1, k2; the camera looks
down -z) against the oracle on the same inputs: the 10-camera problem turned
into the BAL model (helpers.bal_problem), written to and read back from a BAL
file.  Linearisation 1e-12, Schur increments 1e-9 at three dampings, the LM
log and final cost 1e-9, mixed precision's final cost 1e-6, the
Gauss-Newton adjoint 1e-9 against the FP64 reference; batched windows
reject the model."""
import numpy as np
import pytest

from helpers import bal_problem, dense_system, oracle_blocks, relerr, write_bal
from nelson_amd import BABatch, BAProblem, CONFIGS, GaussNewton, GaussNewtonSettings, LevenbergMarquardt, \
    LevenbergMarquardtSettings, NelsonError, SolverCholeskySchur, load_bal
from nelson_amd.autograd import translation_loss_grad
from oracle.oracle import OracleBA
from test_gpu_parity import _compare_logs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bal(tmp_path_factory):
    g = bal_problem(CONFIGS["cfg1"])
    f = tmp_path_factory.mktemp("bal") / "cfg1_bal.txt"
    write_bal(f, g)
    d = load_bal(f)
    d["cams_gt"] = g["cams_gt"]
    return d


def test_bal_linearization_and_steps(bal):
    p = BAProblem.from_generated(bal)
    o = OracleBA.from_generated(bal)
    chi2 = p.update(True)
    o.update(True)
    Ho = oracle_blocks(o)
    Hg = p.export_hessian()
    assert abs(chi2 - Ho["chi2"]) <= 1e-12 * Ho["chi2"]
    for k in ("U", "V", "W", "bU", "bV"):
        assert relerr(Hg[k], Ho[k]) <= 1e-12, k
    s = SolverCholeskySchur()
    s.init(p)
    for rel, ab in ((0.0, 62.0), (1e-3, 0.0), (0.0, 1e-4)):
        assert s.computeIncrement(rel, ab)
        ro = o.schur_step(rel, ab)
        dxu, dxv = s.incrementVector()
        assert relerr(dxu, ro["xU"]) <= 1e-9 and relerr(dxv, ro["xV"]) <= 1e-9, (rel, ab)


def test_bal_levenberg_marquardt(bal):
    p = BAProblem.from_generated(bal)
    lm = LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=20))
    reason = lm.solve(p)
    ro = OracleBA.from_generated(bal).lm_solve(maxNumIt=20)
    detail = _compare_logs(lm.stats(), ro["log"], "bal lm")
    c, co = lm.result()["final_chi2"], ro["final_chi2"]
    print(f"\nBAL LM: {reason} {c:.12g}, oracle {ro['reason']} {co:.12g}")
    assert abs(c - co) <= 1e-9 * co, detail
    pm = BAProblem.from_generated(bal, precision="mixed")
    lmm = LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=20))
    lmm.solve(pm)
    assert abs(lmm.result()["final_chi2"] - co) <= 1e-6 * co


def test_bal_gauss_newton_adjoint(bal):
    p = BAProblem.from_generated(bal)
    LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=200)).solve(p)
    GaussNewton(GaussNewtonSettings(maxNumIt=8, epsIncSquare=0.0, epsBVector=0.0, epsChi2=0.0)).solve(p)
    t = p.camera_translations()
    gc = translation_loss_grad(t, 2.0 * (t - bal["cams_gt"].reshape(-1, 6)[:, 3:6]))
    dobs, dw = p.adjoint(g_cams=gc, hessian_steps=0)
    o = OracleBA.from_generated(bal)
    cams, pts = p.get_state_rt()
    o.set_state_rt(cams, pts)
    o.update(True)
    A, _ = dense_system(oracle_blocks(o), o.NU, o.NV)
    cf = p.cam_fixed.astype(bool)
    lam = np.linalg.solve(A, np.concatenate([gc.reshape(-1, 6)[~cf].reshape(-1), np.zeros(3 * o.NV)]))
    lu, lv = lam[:6 * o.NU], lam[6 * o.NU:]
    r, Jc, Jp = o.edge_jacobians()
    cvar = np.cumsum(~cf) - 1
    pvar = np.cumsum(~p.pt_fixed.astype(bool)) - 1
    a = np.zeros((len(r), 2))
    for e in range(len(r)):
        c, q = p.cam_idx[e], p.pt_idx[e]
        if not p.cam_fixed[c]:
            a[e] += Jc[e] @ lu[6 * cvar[c]:6 * cvar[c] + 6]
        if not p.pt_fixed[q]:
            a[e] += Jp[e] @ lv[3 * pvar[q]:3 * pvar[q] + 3]
    rd, rw = a, -np.sum(a * r, axis=1)
    scale = max(np.abs(rd).max(), np.abs(rw).max())
    err = max(np.abs(dobs - rd).max(), np.abs(dw - rw).max()) / scale
    print(f"\nBAL GN adjoint vs FP64 reference: {err:.2e}")
    assert err <= 1e-9


def test_batch_rejects_bal_model(bal):
    with pytest.raises(NelsonError):
        BABatch([bal], precision="fp64")
