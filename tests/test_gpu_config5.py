"""The differentiable-BA configuration at its full size: 200 cameras, 20 000
points, 6 observations per point (120 000 observations), run to
convergence, loss L = |refined camera translations - ground truth|^2.

  * Gauss-Newton adjoint (hessian_steps = 0) against an FP64 reference
    gradient: the oracle's Hessian blocks at the GPU's solution, reduced to
    the 1194 x 1194 camera system with scipy's sparse products, a dense
    solve, back-substitution for the points and the per-edge formula -> 1e-9
    of the largest entry;
  * exact-Hessian adjoint (hessian_steps = 3) against central finite
    differences of L through re-solves (h = 1e-6) on 100 random coordinates
    (70 observation coordinates, 30 edge weights) -> 1e-4 of the largest
    entry;
  * the mixed-precision forward: the same solve in NSN_MIXED (FP32 assembly
    and factorisation, FP64 refinement), whose exact-Hessian adjoint matches
    the FP64 one.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from nelson_amd import BAProblem, CONFIGS, GaussNewton, GaussNewtonSettings, LevenbergMarquardt, \
    LevenbergMarquardtSettings, generate_ba
from nelson_amd.autograd import translation_loss_grad
from oracle.oracle import OracleBA

from helpers import oracle_blocks

pytestmark = pytest.mark.gpu

POLISH = dict(maxNumIt=8, epsIncSquare=0.0, epsBVector=0.0, epsChi2=0.0)


@pytest.fixture(scope="module")
def cfg5():
    g = generate_ba(CONFIGS["cfg5"])
    assert (g["n_cameras"], g["n_points"], g["n_edges"]) == (200, 20000, 120000)
    return g


def _solve(g, precision="fp64", p=None):
    p = p or BAProblem.from_generated(g, precision=precision)
    reason = LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=200)).solve(p)
    GaussNewton(GaussNewtonSettings(**POLISH)).solve(p)
    return p, reason


def _loss_grad(p, g):
    t = p.camera_translations()
    t_gt = g["cams_gt"].reshape(-1, 6)[:, 3:6]
    return float(np.sum((t - t_gt) ** 2)), translation_loss_grad(t, 2.0 * (t - t_gt))


def _reference_gn_adjoint(p, g, gc):
    """FP64 reference of the Gauss-Newton adjoint through the Schur
    complement: lambda_u = S^-1 g_u, lambda_v = -V^-1 W^T lambda_u (g_v = 0),
    then dL/dobs_e = a_e, dL/dw_e = -a_e . r_e with a_e = Jc lambda_c + Jp
    lambda_p (unit weights, no Huber)."""
    o = OracleBA.from_generated(g)
    cams, pts = p.get_state_rt()
    o.set_state_rt(cams, pts)
    o.update(True)
    H = oracle_blocks(o)
    NU, NV = o.NU, o.NV
    nu, nv = 6 * NU, 3 * NV
    Wi, Wj, W = H["Wi"], H["Wj"], H["W"]
    rows = (6 * Wi[:, None, None] + np.arange(6)[None, :, None] + 0 * np.arange(3)[None, None, :]).reshape(-1)
    cols = (3 * Wj[:, None, None] + 0 * np.arange(6)[None, :, None] + np.arange(3)[None, None, :]).reshape(-1)
    Ws = sp.csr_matrix((W.reshape(-1), (rows, cols)), shape=(nu, nv))
    Vinv = np.linalg.inv(H["V"])
    vr = (3 * np.arange(NV)[:, None, None] + np.arange(3)[None, :, None] + 0 * np.arange(3)[None, None, :]).reshape(-1)
    vc = (3 * np.arange(NV)[:, None, None] + 0 * np.arange(3)[None, :, None] + np.arange(3)[None, None, :]).reshape(-1)
    Vis = sp.csr_matrix((Vinv.reshape(-1), (vr, vc)), shape=(nv, nv))
    S = -(Ws @ Vis @ Ws.T).toarray()
    for u in range(NU):
        S[6 * u:6 * u + 6, 6 * u:6 * u + 6] += H["U"][u]
    cf = p.cam_fixed.astype(bool)
    gu = gc.reshape(-1, 6)[~cf].reshape(-1)
    lam_u = np.linalg.solve(S, gu)
    lam_v = -(Vis @ (Ws.T @ lam_u))
    r, Jc, Jp = o.edge_jacobians()
    cvar = np.cumsum(~cf) - 1
    pvar = np.cumsum(~p.pt_fixed.astype(bool)) - 1
    ci, qi = p.cam_idx, p.pt_idx
    lu = np.where(cf[ci][:, None], 0.0, lam_u.reshape(-1, 6)[np.maximum(cvar[ci], 0)])
    lv = np.where(p.pt_fixed.astype(bool)[qi][:, None], 0.0, lam_v.reshape(-1, 3)[np.maximum(pvar[qi], 0)])
    a = np.einsum("eij,ej->ei", Jc, lu) + np.einsum("eij,ej->ei", Jp, lv)
    return a, -np.sum(a * r, axis=1)


def test_config5_gauss_newton_adjoint_vs_fp64_reference(cfg5):
    p, reason = _solve(cfg5)
    L, gc = _loss_grad(p, cfg5)
    dobs, dw = p.adjoint(g_cams=gc, hessian_steps=0)
    rd, rw = _reference_gn_adjoint(p, cfg5, gc)
    scale = max(np.abs(rd).max(), np.abs(rw).max())
    eo, ew = np.abs(dobs - rd).max() / scale, np.abs(dw - rw).max() / scale
    print(f"\nconfig 5 ({reason}, L = {L:.6e}): GN adjoint vs FP64 reference: dobs {eo:.2e}, dw {ew:.2e}")
    assert eo <= 1e-9 and ew <= 1e-9


def test_config5_exact_adjoint_vs_central_differences_and_mixed_forward(cfg5):
    base, _ = _solve(cfg5)
    L0, gc = _loss_grad(base, cfg5)
    dobs, dw = base.adjoint(g_cams=gc, hessian_steps=3)
    cams_star, pts_star = base.get_state()
    E = cfg5["n_edges"]
    obs0 = cfg5["obs"].reshape(-1, 2).copy()
    w0 = np.ones(E)
    h = 1e-6

    def loss_at(obs, w):
        base.set_observations(obs, w)
        base.set_state(cams_star, pts_star)  # warm start at x*
        GaussNewton(GaussNewtonSettings(**POLISH)).solve(base)
        return _loss_grad(base, cfg5)[0]

    rng = np.random.default_rng(55)
    fd, an = [], []
    for _ in range(70):
        e, k = int(rng.integers(E)), int(rng.integers(2))
        op, om = obs0.copy(), obs0.copy()
        op[e, k] += h
        om[e, k] -= h
        fd.append((loss_at(op, w0) - loss_at(om, w0)) / (2 * h))
        an.append(dobs[e, k])
    for _ in range(30):
        e = int(rng.integers(E))
        wp, wm = w0.copy(), w0.copy()
        wp[e] += h
        wm[e] -= h
        fd.append((loss_at(obs0, wp) - loss_at(obs0, wm)) / (2 * h))
        an.append(dw[e])
    scale = max(np.abs(dobs).max(), np.abs(dw).max())
    efd = np.abs(np.array(fd) - np.array(an)).max() / scale
    print(f"\nconfig 5 exact adjoint vs central differences on 100 coordinates: {efd:.2e} of max |grad| {scale:.3e}")
    assert efd <= 1e-4
    # the mixed-precision forward and its adjoint
    pm, reason_m = _solve(cfg5, precision="mixed")
    Lm, gcm = _loss_grad(pm, cfg5)
    dom, dwm = pm.adjoint(g_cams=gcm, hessian_steps=3)
    loss_at(obs0, w0)  # the unperturbed problem at its solution again
    chi_m, chi_64 = pm.update(False), base.update(False)
    em = max(np.abs(dom - dobs).max(), np.abs(dwm - dw).max()) / scale
    print(f"mixed forward ({reason_m}): final cost rel err {abs(chi_m - chi_64) / chi_64:.2e}, "
          f"L rel err {abs(Lm - L0) / L0:.2e}, exact adjoint vs FP64 {em:.2e}")
    assert abs(chi_m - chi_64) <= 1e-6 * chi_64
    assert em <= 1e-6
