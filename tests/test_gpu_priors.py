"""Unary priors (nsn_set_priors; nelson's EdgeUnary, EdgeUnary.h:76-93, for
the BA section) on the GPU against the oracle: camera priors on angle-axis
and translation, point priors on position, on fixed and variable nodes.
Linearisation 1e-12, Schur increments 1e-9, LM final cost 1e-9 (FP64);
mixed precision reaches the FP64 final cost to 1e-6 and its refined
increment the oracle's to 1e-9; the Gauss-Newton adjoint includes the
priors' Hessian (1e-9 against the FP64 reference) and the exact-Hessian
adjoint their second-order terms (1e-6 against central differences of the
oracle's gradient, priors included)."""
import numpy as np
import pytest

from nelson_amd import BAProblem, CONFIGS, GaussNewton, GaussNewtonSettings, LevenbergMarquardt, \
    LevenbergMarquardtSettings, SolverCholeskySchur, generate_ba
from nelson_amd.autograd import translation_loss_grad
from oracle.oracle import OracleBA

from helpers import dense_system, oracle_blocks, relerr

pytestmark = pytest.mark.gpu


def _priors(g, seed=4):
    rng = np.random.default_rng(seed)
    cam_ids = np.array([0, 2, 5, 7], np.int32)  # camera 0 is fixed: a constant cost term
    cam_p0 = np.asarray(g["cams0"]).reshape(-1, 6)[cam_ids] + rng.normal(0, [0.02] * 3 + [0.1] * 3, (4, 6))
    cam_w = rng.uniform(0.5, 3.0, (4, 6))
    pt_ids = np.array([0, 3, 17, 250, 499], np.int32)  # point 0 is fixed
    pt_p0 = np.asarray(g["pts0"]).reshape(-1, 3)[pt_ids] + rng.normal(0, 0.2, (5, 3))
    pt_w = rng.uniform(0.5, 3.0, (5, 3))
    return dict(cam_ids=cam_ids, cam_p0=cam_p0, cam_w=cam_w, pt_ids=pt_ids, pt_p0=pt_p0, pt_w=pt_w)


def _pair(precision="fp64"):
    g = generate_ba(CONFIGS["cfg1"])
    pri = _priors(g)
    p = BAProblem.from_generated(g, precision=precision)
    p.set_priors(**pri)
    o = OracleBA.from_generated(g)
    o.set_priors(**pri)
    return g, pri, p, o


def test_priors_linearization_and_step():
    g, pri, p, o = _pair()
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


def test_priors_replace_and_remove():
    """A second call replaces the priors; empty arrays remove them."""
    g, pri, p, o = _pair()
    c1 = p.update(True)
    p.set_priors()
    c0 = p.update(True)
    o0 = OracleBA.from_generated(g)
    assert abs(c0 - o0.update(True)) <= 1e-12 * c0
    p.set_priors(**pri)
    assert abs(p.update(True) - c1) <= 1e-15 * c1


def test_priors_lm_matches_oracle():
    g, pri, p, o = _pair()
    lm = LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=20))
    reason = lm.solve(p)
    ro = o.lm_solve(maxNumIt=20)
    c, co = lm.result()["final_chi2"], ro["final_chi2"]
    print(f"\nLM with priors: GPU {c:.12g} ({reason}), oracle {co:.12g} ({ro['reason']})")
    assert abs(c - co) <= 1e-9 * co


def test_priors_mixed_precision():
    g, pri, p64, o = _pair()
    _, _, pm, _ = _pair("mixed")
    lm64 = LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=20))
    lm64.solve(p64)
    lmm = LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=20))
    lmm.solve(pm)
    c64, cm = lm64.result()["final_chi2"], lmm.result()["final_chi2"]
    assert abs(cm - c64) <= 1e-6 * c64, (cm, c64)
    # one refined mixed step at the initial state against the oracle
    _, _, q, o2 = _pair("mixed")
    q.update(True)
    o2.update(True)
    s = SolverCholeskySchur()
    s.init(q)
    assert s.computeIncrement(0.0, 1e-3)
    ro = o2.schur_step(0.0, 1e-3)
    dxu, dxv = s.incrementVector()
    print(f"\nmixed step with priors vs oracle: {relerr(dxu, ro['xU']):.2e} / {relerr(dxv, ro['xV']):.2e}")
    assert relerr(dxu, ro["xU"]) <= 1e-9 and relerr(dxv, ro["xV"]) <= 1e-9


def test_priors_gauss_newton_adjoint():
    """The adjoint's Gauss-Newton Hessian includes the priors: lambda from the
    oracle's blocks (priors included) and the per-edge formula."""
    g, pri, p, o = _pair()
    LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=200)).solve(p)
    GaussNewton(GaussNewtonSettings(maxNumIt=8, epsIncSquare=0.0, epsBVector=0.0, epsChi2=0.0)).solve(p)
    t = p.camera_translations()
    t_gt = g["cams_gt"].reshape(-1, 6)[:, 3:6]
    gc = translation_loss_grad(t, 2.0 * (t - t_gt))
    dobs, dw = p.adjoint(g_cams=gc, hessian_steps=0)
    cams, pts = p.get_state_rt()
    o.set_state_rt(cams, pts)
    o.update(True)
    H = oracle_blocks(o)
    A, _ = dense_system(H, o.NU, o.NV)
    cf = p.cam_fixed.astype(bool)
    gu = gc.reshape(-1, 6)[~cf].reshape(-1)
    lam = np.linalg.solve(A, np.concatenate([gu, np.zeros(3 * o.NV)]))
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
    ar = np.sum(a * r, axis=1)
    rd, rw = a, -ar
    scale = max(np.abs(rd).max(), np.abs(rw).max())
    print(f"\nGN adjoint with priors vs FP64 reference: {np.abs(dobs - rd).max() / scale:.2e}, "
          f"{np.abs(dw - rw).max() / scale:.2e}")
    assert np.abs(dobs - rd).max() <= 1e-9 * scale
    assert np.abs(dw - rw).max() <= 1e-9 * scale


def test_priors_exact_hessian_adjoint():
    g = generate_ba(dict(CONFIGS["cfg1"], n_cameras=6, n_points=80, k_min=3, k_max=6, seed=7))
    rng = np.random.default_rng(9)
    cam_ids = np.array([0, 2, 4], np.int32)
    pt_ids = np.array([0, 5, 40, 79], np.int32)
    pri = dict(cam_ids=cam_ids,
               cam_p0=g["cams0"].reshape(-1, 6)[cam_ids] + rng.normal(0, [0.05] * 3 + [0.2] * 3, (3, 6)),
               cam_w=rng.uniform(0.5, 3.0, (3, 6)), pt_ids=pt_ids,
               pt_p0=g["pts0"].reshape(-1, 3)[pt_ids] + rng.normal(0, 0.3, (4, 3)),
               pt_w=rng.uniform(0.5, 3.0, (4, 3)))
    p = BAProblem.from_generated(g)
    p.set_priors(**pri)
    LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=200)).solve(p)
    GaussNewton(GaussNewtonSettings(maxNumIt=8, epsIncSquare=0.0, epsBVector=0.0, epsChi2=0.0)).solve(p)
    t = p.camera_translations()
    gc = translation_loss_grad(t, 2.0 * (t - g["cams_gt"].reshape(-1, 6)[:, 3:6]))
    dobs, dw = p.adjoint(g_cams=gc, hessian_steps=3)
    o = OracleBA.from_generated(g)
    o.set_priors(**pri)
    cams, pts = p.get_state_rt()
    NU, NV = o.NU, o.NV
    n = 6 * NU + 3 * NV
    B = np.zeros((n, n))
    h = 1e-6
    for i in range(n):
        cols = []
        for sgn in (1.0, -1.0):
            o.set_state_rt(cams, pts)
            d = np.zeros(n)
            d[i] = sgn * h
            o.oplus(d[:6 * NU], d[6 * NU:])
            o.update(True)
            H = oracle_blocks(o)
            cols.append(np.concatenate([H["bU"], H["bV"]]))
        B[:, i] = (cols[0] - cols[1]) / (2 * h)
    B = 0.5 * (B + B.T)
    o.set_state_rt(cams, pts)
    o.update(True)
    cf = p.cam_fixed.astype(bool)
    lam = np.linalg.solve(B, np.concatenate([gc.reshape(-1, 6)[~cf].reshape(-1), np.zeros(3 * NV)]))
    lu, lv = lam[:6 * NU], lam[6 * NU:]
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
    print(f"\nexact-Hessian adjoint with priors vs FP64 reference: {err:.2e}")
    assert err <= 1e-6
