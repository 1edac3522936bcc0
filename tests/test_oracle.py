"""Pins the CPU oracle independently of its own code: against numpy/scipy
linear algebra, closed forms, central finite differences, the reference's
own property tests (nelson/tests/testMatrix*.cpp) and its linear toy problem
(nelson/tests/testDoubleSectionCommon.h)."""
import numpy as np
import pytest
import scipy.linalg as sla
from scipy.spatial.transform import Rotation

from oracle import oracle as orc
from oracle.oracle import OracleBA, PyProblem

from helpers import CONFIGS, dense_system, generate_ba, oracle_blocks, relerr, small_cfg


# ---------------------------------------------------------------------------
# dense pivoted LDL^T (Eigen::LDLT restated)
# ---------------------------------------------------------------------------
def test_ldlt_solves_spd_and_indefinite():
    rng = np.random.default_rng(1)
    for n, spd in [(1, True), (7, True), (60, True), (33, False)]:
        A = rng.standard_normal((n, n))
        A = A @ A.T + n * np.eye(n) if spd else A + A.T
        b = rng.standard_normal(n)
        ok, f, tp = orc.ldlt(A)
        assert ok
        x = orc.ldlt_solve(f, tp, b)
        assert relerr(A @ x, b) < 1e-11
        assert relerr(x, np.linalg.solve(A, b)) < 1e-9


def test_ldlt_reconstruction_and_pivot_order():
    """P A P^T = L D L^T with the pivot chosen as the largest remaining
    |diagonal| (first occurrence)."""
    A = np.array([[1.0, 0.2, 0.1], [0.2, 5.0, 0.3], [0.1, 0.3, 3.0]])
    ok, f, tp = orc.ldlt(A)
    assert ok and tp[0] == 1
    n = 3
    L = np.tril(np.asarray(f), -1) + np.eye(n)
    D = np.diag(np.diag(np.asarray(f)))
    P = np.eye(n)
    for k in range(n):
        P[[k, tp[k]]] = P[[tp[k], k]]
    assert relerr(P @ A @ P.T, L @ D @ L.T) < 1e-14
    # scipy's LDL of the same matrix gives the same D up to ordering
    _, d2, _ = sla.ldl(A)
    assert np.allclose(sorted(np.diag(D)), sorted(np.diag(d2)), rtol=0.5)  # same inertia, positive
    assert np.all(np.diag(d2) > 0)


def test_ldlt_failure_cases():
    # zero pivot followed by a non-zero column fails (Eigen ComputationInfo::NumericalIssue)
    A = np.array([[0.0, 1.0], [1.0, 0.0]])
    ok, _, _ = orc.ldlt(A)
    assert not ok
    # a zero matrix "succeeds" (nothing to eliminate), like Eigen's LDLT
    ok, _, _ = orc.ldlt(np.zeros((3, 3)))
    assert ok


# ---------------------------------------------------------------------------
# L4 kernels: the reference's property tests on its own patterns
# ---------------------------------------------------------------------------
def _sym(rng, B):
    m = rng.standard_normal((B, B))
    return m @ m.T + B * np.eye(B)


def test_diag_inv_and_mult():  # nelson/tests/testMatrixVinv.cpp:20-64
    rng = np.random.default_rng(2)
    NV, BV = 5, 3
    V = np.stack([_sym(rng, BV) for _ in range(NV)])
    Vinv = np.zeros_like(V)
    assert orc.lib().or_diag_inv(NV, BV, orc.dp(V), 0.1, 0.5, orc.dp(Vinv))
    for p in range(NV):
        M = np.triu(V[p]) + np.triu(V[p], 1).T
        M = M + np.diag(np.diag(M) * 0.1 + 0.5)
        assert relerr(Vinv[p], np.linalg.inv(M)) < 1e-12
    v = rng.standard_normal(NV * BV)
    res = np.zeros_like(v)
    orc.lib().or_diag_inv_mult(NV, BV, orc.dp(Vinv), orc.dp(v), orc.dp(res))
    ref = np.concatenate([Vinv[p] @ v[BV * p:BV * p + BV] for p in range(NV)])
    assert relerr(res, ref) < 1e-14


# the fixed 3x5 block pattern of nelson/tests/testMatrixWVinvMultiplier.cpp:45-52
_PATTERN = [(0, 0), (0, 3), (0, 4), (1, 0), (1, 1), (1, 3), (2, 1), (2, 2)]


def _dense_w(W, Wi, Wj, NR, NC, BR, BC):
    D = np.zeros((NR * BR, NC * BC))
    for k in range(len(Wi)):
        D[Wi[k] * BR:Wi[k] * BR + BR, Wj[k] * BC:Wj[k] * BC + BC] = W[k]
    return D


@pytest.mark.parametrize("BR,BC", [(6, 3), (2, 3), (3, 2)])
def test_wvinv_and_rightmult(BR, BC):  # testMatrixWVinvMultiplier.cpp:17-116
    rng = np.random.default_rng(3)
    NR, NC = 3, 5
    Wi = np.array([a for a, _ in _PATTERN], np.int32)
    Wj = np.array([b for _, b in _PATTERN], np.int32)
    W = rng.standard_normal((len(Wi), BR, BC))
    Vinv = rng.standard_normal((NC, BC, BC))
    Y = np.zeros_like(W)
    orc.lib().or_wvinv(len(Wi), BR, BC, orc.ip(Wj), orc.dp(W), orc.dp(Vinv), orc.dp(Y))
    Dv = sla.block_diag(*Vinv)
    Dw = _dense_w(W, Wi, Wj, NR, NC, BR, BC)
    assert relerr(_dense_w(Y, Wi, Wj, NR, NC, BR, BC), Dw @ Dv) < 1e-14
    b = rng.standard_normal(NC * BC)
    bv = rng.standard_normal(NR * BR)
    res = bv.copy()
    orc.lib().or_wvinv_mult_sub(len(Wi), BR, BC, orc.ip(Wi), orc.ip(Wj), orc.dp(Y), orc.dp(b), orc.dp(res))
    assert relerr(res, bv - (Dw @ Dv) @ b) < 1e-13


def test_wwt():  # testMatrixWWtMultiplier.cpp:17-103, U with an extra (0,2) block
    rng = np.random.default_rng(4)
    NR, NC, BR, BC = 3, 5, 6, 3
    Wi = np.array([a for a, _ in _PATTERN], np.int32)
    Wj = np.array([b for _, b in _PATTERN], np.int32)
    A = rng.standard_normal((len(Wi), BR, BC))
    B = rng.standard_normal((len(Wi), BR, BC))
    nu = NR * BR
    U = np.zeros((nu, nu))
    for i in range(NR):
        m = _sym(rng, BR)
        U[i * BR:i * BR + BR, i * BR:i * BR + BR] = m
    U[0:BR, 2 * BR:3 * BR] = rng.standard_normal((BR, BR))
    S = np.zeros((nu, nu))
    orc.lib().or_wwt(NR, BR, BC, len(Wi), orc.ip(Wi), orc.ip(Wj), orc.dp(A), orc.dp(B), orc.dp(U), orc.dp(S))
    ref = np.triu(U - _dense_w(A, Wi, Wj, NR, NC, BR, BC) @ _dense_w(B, Wi, Wj, NR, NC, BR, BC).T)
    # only upper blocks are produced, within-block lower parts of diagonal blocks excluded
    mask = np.zeros((nu, nu), bool)
    for i in range(NR):
        for j in range(i, NR):
            mask[i * BR:i * BR + BR, j * BR:j * BR + BR] = True
    mask &= np.triu(np.ones((nu, nu), bool))
    assert relerr(S[mask], ref[mask]) < 1e-13


def test_wtx():  # testMatrixWtXMultiplier.cpp:17-92
    rng = np.random.default_rng(5)
    NR, NC, BR, BC = 3, 5, 6, 3
    Wi = np.array([a for a, _ in _PATTERN], np.int32)
    Wj = np.array([b for _, b in _PATTERN], np.int32)
    W = rng.standard_normal((len(Wi), BR, BC))
    x = rng.standard_normal(NR * BR)
    bV = rng.standard_normal(NC * BC)
    res = bV.copy()
    orc.lib().or_wtx_sub(len(Wi), BR, BC, orc.ip(Wi), orc.ip(Wj), orc.dp(W), orc.dp(x), orc.dp(res))
    assert relerr(res, bV - _dense_w(W, Wi, Wj, NR, NC, BR, BC).T @ x) < 1e-13


# ---------------------------------------------------------------------------
# BA edge, rotations, assembly
# ---------------------------------------------------------------------------
def test_rodrigues_and_se3_update_vs_scipy():
    rng = np.random.default_rng(6)
    for w in [np.zeros(3), 1e-9 * rng.standard_normal(3), 1e-4 * rng.standard_normal(3), rng.standard_normal(3)]:
        R = np.zeros(9)
        orc.lib().or_rodrigues(orc.dp(w), orc.dp(R))
        assert relerr(R.reshape(3, 3), Rotation.from_rotvec(w).as_matrix()) < 1e-13
        # left update exp(delta) * T against the 4x4 matrix exponential of the twist
        d = np.concatenate([w, rng.standard_normal(3)])
        R0 = Rotation.from_rotvec(rng.standard_normal(3)).as_matrix()
        t0 = rng.standard_normal(3)
        Rm, tm = R0.reshape(-1).copy(), t0.copy()
        orc.lib().or_se3_left_update(orc.dp(d), orc.dp(Rm), orc.dp(tm))
        X = np.zeros((4, 4))
        X[:3, :3] = np.array([[0, -d[2], d[1]], [d[2], 0, -d[0]], [-d[1], d[0], 0]])
        X[:3, 3] = d[3:]
        T0 = np.eye(4)
        T0[:3, :3], T0[:3, 3] = R0, t0
        T1 = sla.expm(X) @ T0
        assert relerr(Rm.reshape(3, 3), T1[:3, :3]) < 1e-12
        assert relerr(tm, T1[:3, 3]) < 1e-12


def _numpy_residuals(g, cams_rt, pts):
    E = g["n_edges"]
    r = np.zeros((E, 2))
    for e in range(E):
        c, q = g["cam_idx"][e], g["pt_idx"][e]
        R = cams_rt[12 * c:12 * c + 9].reshape(3, 3)
        t = cams_rt[12 * c + 9:12 * c + 12]
        Xc = R @ pts[3 * q:3 * q + 3] + t
        r[e] = g["focal"] * Xc[:2] / Xc[2] + np.array([g["cx"], g["cy"]]) - g["obs"][2 * e:2 * e + 2]
    return r


def test_edge_residual_and_jacobians_fd():
    g = generate_ba(small_cfg(n_cameras=4, n_points=20, k_min=2, k_max=3, seed=7))
    o = OracleBA.from_generated(g)
    cams, pts = o.get_state_rt()
    # residuals against an independent numpy projection with scipy's Rodrigues
    cr = np.zeros(12 * g["n_cameras"])
    for c in range(g["n_cameras"]):
        cr[12 * c:12 * c + 9] = Rotation.from_rotvec(g["cams0"][6 * c:6 * c + 3]).as_matrix().reshape(-1)
        cr[12 * c + 9:12 * c + 12] = g["cams0"][6 * c + 3:6 * c + 6]
    r, Jc, Jp = o.edge_jacobians()
    assert relerr(r, _numpy_residuals(g, cr, g["pts0"])) < 1e-10
    # central differences of the residual w.r.t. the left perturbation and the point
    h = 1e-6
    for e in range(g["n_edges"]):
        c, q = g["cam_idx"][e], g["pt_idx"][e]
        for k in range(6):
            d = np.zeros(6)
            d[k] = h
            rp, rm = [], []
            for s, store in ((1, rp), (-1, rm)):
                Rm, tm = cams[12 * c:12 * c + 9].copy(), cams[12 * c + 9:12 * c + 12].copy()
                orc.lib().or_se3_left_update(orc.dp(s * d), orc.dp(Rm), orc.dp(tm))
                cc = cams.copy()
                cc[12 * c:12 * c + 9], cc[12 * c + 9:12 * c + 12] = Rm, tm
                store.append(_numpy_residuals(g, cc, pts)[e])
            fd = (rp[0] - rm[0]) / (2 * h)
            assert np.allclose(Jc[e][:, k], fd, rtol=1e-5, atol=1e-4)
        for k in range(3):
            pp, pm = pts.copy(), pts.copy()
            pp[3 * q + k] += h
            pm[3 * q + k] -= h
            fd = (_numpy_residuals(g, cams, pp)[e] - _numpy_residuals(g, cams, pm)[e]) / (2 * h)
            assert np.allclose(Jp[e][:, k], fd, rtol=1e-5, atol=1e-4)


def test_hessian_equals_dense_jtj():
    """U, V, W, bU, bV and chi2 equal J^T J, J^T r, |r|^2 assembled densely
    from the per-edge Jacobians (fixed camera 0 / point 0 columns dropped)."""
    g = generate_ba(small_cfg(n_cameras=5, n_points=40, k_min=2, k_max=4, seed=8))
    o = OracleBA.from_generated(g)
    chi2 = o.update(True)
    r, Jc, Jp = o.edge_jacobians()
    NU, NV = o.NU, o.NV
    J = np.zeros((2 * g["n_edges"], 6 * NU + 3 * NV))
    for e in range(g["n_edges"]):
        c, q = g["cam_idx"][e], g["pt_idx"][e]
        if c > 0:
            J[2 * e:2 * e + 2, 6 * (c - 1):6 * c] = Jc[e]
        if q > 0:
            J[2 * e:2 * e + 2, 6 * NU + 3 * (q - 1):6 * NU + 3 * q] = Jp[e]
    Hd, b = dense_system(oracle_blocks(o), NU, NV)
    assert relerr(Hd, J.T @ J) < 1e-12
    assert relerr(b, J.T @ r.reshape(-1)) < 1e-12
    assert abs(chi2 - np.sum(r ** 2)) < 1e-12 * chi2


# ---------------------------------------------------------------------------
# Schur step vs brute force
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("rel,ab", [(0.0, 1e3), (0.0, 1.0), (0.05, 0.3), (0.0, 0.0)])
def test_schur_step_equals_dense_solve(cfg1, rel, ab):
    """dx of SolverCholeskySchur == solve((H + diag(rel*H_ii + abs)) dx = -b)."""
    o = OracleBA.from_generated(cfg1)
    o.update(True)
    Hd, b = dense_system(oracle_blocks(o), o.NU, o.NV)
    D = np.diag(np.diag(Hd) * rel + ab)
    x = np.linalg.solve(Hd + D, -b)
    s = o.schur_step(rel, ab)
    assert s["ok"]
    assert relerr(np.concatenate([s["xU"], s["xV"]]), x) < 1e-10


def test_schur_matrix_is_the_schur_complement(cfg1):
    o = OracleBA.from_generated(cfg1)
    o.update(True)
    Hd, b = dense_system(oracle_blocks(o), o.NU, o.NV)
    nu = 6 * o.NU
    mu = 2.0
    A = Hd + mu * np.eye(len(b))
    Sref = A[:nu, :nu] - A[:nu, nu:] @ np.linalg.solve(A[nu:, nu:], A[nu:, :nu])
    bSref = -b[:nu] - A[:nu, nu:] @ np.linalg.solve(A[nu:, nu:], -b[nu:])
    s = o.schur_step(0.0, mu, want_s=True)
    assert relerr(np.triu(s["S"]), np.triu(Sref)) < 1e-11
    assert relerr(s["bS"], bSref) < 1e-11


def test_rho_closed_form(cfg1):
    o = OracleBA.from_generated(cfg1)
    old = o.update(True)
    s = o.schur_step(0.0, 5.0)
    h = o.hessian_struct()
    new = old - 10.0
    den = s["xU"] @ (5.0 * s["xU"] - o.hessian()["bU"]) + s["xV"] @ (5.0 * s["xV"] - o.hessian()["bV"])
    rho = orc.lib().or_rho(orc.C.byref(h), 5.0, orc.dp(s["xU"]), orc.dp(s["xV"]), old, new)
    assert abs(rho - 10.0 / den) < 1e-12 * abs(rho)


# ---------------------------------------------------------------------------
# drivers on the linear toy problem of nelson/tests/testDoubleSectionCommon.h
# ---------------------------------------------------------------------------
class Toy(PyProblem):
    """5 2-D points (U), 3 3-D points (V); unary, U-U and U-V difference
    edges with identity Jacobians; the measurements are exact for the ground
    truth, so the optimum has chi2 = 0."""
    N2, N3 = 5, 3

    def __init__(self, seed=0, sigma=0.5):
        Wi = [i for i in range(self.N2) for j in range(self.N3)]
        Wj = [j for i in range(self.N2) for j in range(self.N3)]
        super().__init__(2, 3, self.N2, self.N3, Wi, Wj)
        rng = np.random.default_rng(seed)
        self.gt2 = rng.uniform(-1, 1, (self.N2, 2))
        self.gt3 = rng.uniform(-1, 1, (self.N3, 3))
        self.p2 = self.gt2 + rng.uniform(-1, 1, (self.N2, 2)) * sigma
        self.p3 = self.gt3 + rng.uniform(-1, 1, (self.N3, 3)) * sigma
        self.nupdates = 0

    def update(self, hess):
        self.nupdates += 1
        v = self.views()
        chi2 = 0.0
        if hess:
            for k in ("U", "V", "W", "bU", "bV"):
                v[k][...] = 0
        U, V, W, bU, bV = v["U"], v["V"], v["W"], v["bU"], v["bV"]
        for i in range(self.N2):  # unary U
            e = self.p2[i] - self.gt2[i]
            chi2 += e @ e
            if hess:
                U[2 * i:2 * i + 2, 2 * i:2 * i + 2] += np.eye(2)
                bU[2 * i:2 * i + 2] += e
        for j in range(self.N3):  # unary V
            e = self.p3[j] - self.gt3[j]
            chi2 += e @ e
            if hess:
                V[j] += np.eye(3)
                bV[3 * j:3 * j + 3] += e
        for i in range(self.N2):  # U-U
            for j in range(i + 1, self.N2):
                e = (self.p2[i] - self.p2[j]) - (self.gt2[i] - self.gt2[j])
                chi2 += e @ e
                if hess:
                    U[2 * i:2 * i + 2, 2 * i:2 * i + 2] += np.eye(2)
                    U[2 * j:2 * j + 2, 2 * j:2 * j + 2] += np.eye(2)
                    U[2 * i:2 * i + 2, 2 * j:2 * j + 2] -= np.eye(2)
                    bU[2 * i:2 * i + 2] += e
                    bU[2 * j:2 * j + 2] -= e
        k = 0
        for i in range(self.N2):  # U-V (first two coordinates)
            for j in range(self.N3):
                e = (self.p2[i] - self.p3[j][:2]) - (self.gt2[i] - self.gt3[j][:2])
                chi2 += e @ e
                if hess:
                    U[2 * i:2 * i + 2, 2 * i:2 * i + 2] += np.eye(2)
                    bU[2 * i:2 * i + 2] += e
                    V[j] += np.diag([1.0, 1.0, 0.0])
                    bV[3 * j:3 * j + 3] -= np.array([e[0], e[1], 0.0])
                    W[k] += np.array([[-1.0, 0, 0], [0, -1.0, 0]])
                k += 1
        self.h.chi2 = chi2

    def oplus(self, xu, xv):
        self.p2 += xu.reshape(self.N2, 2)
        self.p3 += xv.reshape(self.N3, 3)

    def backup(self):
        self.bk = (self.p2.copy(), self.p3.copy())

    def rollback(self):
        self.p2, self.p3 = self.bk[0].copy(), self.bk[1].copy()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_toy_problem_converges_gn_and_lm(seed):
    t = Toy(seed)
    r = t.gn_solve()
    assert t.h.chi2 < 1e-12, r
    t = Toy(seed)
    r = t.lm_solve()
    assert t.h.chi2 < 1e-12, r


def test_lm_mu_schedule_on_linear_problem():
    """A linear problem makes the quadratic model exact: rho = 1 and
    alpha = clamp(1 - 1^3, 1/3, 2/3) = 1/3, so mu shrinks x1/3 per accepted
    step from mu0 = tau * max|diag H| (LevenbergMarquardt.hpp:58,109-112)."""
    t = Toy(3)
    t.update(True)
    mu0 = 1e-3 * max(np.max(np.abs(np.diag(t.views()["U"]))),
                     max(np.max(np.abs(np.diag(t.views()["V"][j]))) for j in range(3)))
    r = Toy(3).lm_solve(maxNumIt=6, epsIncSquare=0.0, epsChi2=0.0, epsBVector=0.0)
    mus = [e[3] for e in r["log"][1:]]
    assert abs(mus[0] - mu0) < 1e-15 * mu0
    for a, b in zip(mus, mus[1:]):
        assert abs(b - a / 3.0) < 1e-12 * a


class Rejector(Toy):
    """Every candidate's cost is reported as higher than the last accepted
    one, so every step is rejected: mu grows x2, x4, x8, ... and the solve
    ends with MaxSubIterationReached after maxNumSubIt tries."""

    def update(self, hess):
        super().update(hess)
        if not hess:
            self.h.chi2 += 1e6


def test_lm_rejections_and_max_subiterations():
    t = Rejector(4)
    r = t.lm_solve(maxNumSubIt=5)
    assert r["reason"] == "MaxSubIterationReached"
    mus = [e[3] for e in r["log"][1:]]
    assert len(mus) == 5
    ratios = [b / a for a, b in zip(mus, mus[1:])]
    assert np.allclose(ratios, [2.0, 4.0, 8.0, 16.0])


def test_success_at_start_and_gn_failure():
    assert Toy(5).lm_solve(epsBVector=1e300)["reason"] == "SuccessAtStart"
    # removing the unary V edges and the U-V edges leaves V = 0: GN fails
    class Singular(Toy):
        def update(self, hess):
            super().update(hess)
            if hess:
                self.views()["V"][...] = 0.0
    assert Singular(6).gn_solve()["reason"] == "SolverFailure"


def test_lm_converges_cfg1(cfg1):
    """The restated LM brings the cfg1 problem from its perturbed start to a
    cost at the noise level: chi2 ~ 2E - dof for N(0, 1 px) noise."""
    o = OracleBA.from_generated(cfg1)
    r = o.lm_solve(maxNumIt=20)
    dof = 6 * o.NU + 3 * o.NV
    expected = 2 * cfg1["n_edges"] - dof
    assert r["log"][0][2] > 50 * expected
    assert 0.8 * expected < r["final_chi2"] < 1.2 * expected
    # costs of accepted steps decrease monotonically
    acc = [e[2] for e in r["log"][1:] if e[2] >= 0]
    assert all(b <= a * (1 + 1e-12) or b > a for a, b in zip(acc, acc[1:]))


# ---- unary priors -------------------------------------------------------------

def test_so3_log_and_inverse_left_jacobian():
    """or_so3_log against scipy's rotation vector (angles from 1e-9 to
    pi - 1e-9); or_so3_jl_inv against the central difference of
    log(exp(d) R) in d."""
    rng = np.random.default_rng(3)
    for k in range(600):
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        th = [rng.uniform(0, np.pi), np.pi - 10 ** rng.uniform(-9, -2), 10 ** rng.uniform(-9, -2)][k % 3]
        R = Rotation.from_rotvec(th * ax).as_matrix()
        assert np.abs(orc.so3_log(R) - Rotation.from_rotvec(th * ax).as_rotvec()).max() <= 1e-10
    for w in (np.array([0.3, -0.2, 0.5]), np.array([1e-7, -2e-7, 3e-7]), np.array([1.0, 2.0, -1.5])):
        R = Rotation.from_rotvec(w).as_matrix()
        J = np.zeros((3, 3))
        h = 1e-6
        for i in range(3):
            d = np.zeros(3)
            d[i] = h
            J[:, i] = (orc.so3_log(Rotation.from_rotvec(d).as_matrix() @ R)
                       - orc.so3_log(Rotation.from_rotvec(-d).as_matrix() @ R)) / (2 * h)
        assert np.abs(orc.so3_jl_inv(w) - J).max() <= 1e-7, (w, orc.so3_jl_inv(w), J)


def _priors_cfg1():
    g = generate_ba(CONFIGS["cfg1"])
    rng = np.random.default_rng(4)
    cam_ids = np.array([0, 2, 5, 7], np.int32)          # camera 0 is fixed (constant cost only)
    cam_p0 = np.asarray(g["cams0"]).reshape(-1, 6)[cam_ids] + rng.normal(0, [0.02] * 3 + [0.1] * 3, (4, 6))
    cam_w = rng.uniform(0.5, 3.0, (4, 6))
    pt_ids = np.array([0, 3, 17, 250, 499], np.int32)   # point 0 is fixed
    pt_p0 = np.asarray(g["pts0"]).reshape(-1, 3)[pt_ids] + rng.normal(0, 0.2, (5, 3))
    pt_w = rng.uniform(0.5, 3.0, (5, 3))
    return g, dict(cam_ids=cam_ids, cam_p0=cam_p0, cam_w=cam_w, pt_ids=pt_ids, pt_p0=pt_p0, pt_w=pt_w)


def _prior_residuals(cams_rt, pts, pri):
    """The priors' residuals from the state, with scipy's log (independent
    of the oracle)."""
    r = []
    for k, c in enumerate(pri["cam_ids"]):
        R = cams_rt[12 * c:12 * c + 9].reshape(3, 3)
        t = cams_rt[12 * c + 9:12 * c + 12]
        e = np.concatenate([Rotation.from_matrix(R).as_rotvec(), t]) - pri["cam_p0"][k]
        r.append(pri["cam_w"][k] * e)
    for k, q in enumerate(pri["pt_ids"]):
        r.append(pri["pt_w"][k] * (pts[3 * q:3 * q + 3] - pri["pt_p0"][k]))
    return np.concatenate(r)


def test_priors_cost_gradient_hessian_vs_finite_differences():
    """The priors add sum |r|^2 to the cost, J^T r to b and J^T J to the
    diagonal blocks of the variable nodes, J the central difference of the
    residuals under the section's oplus (cameras exp(d) T, points additive)."""
    g, pri = _priors_cfg1()
    o0 = OracleBA.from_generated(g)
    o0.update(True)
    H0 = oracle_blocks(o0)
    o = OracleBA.from_generated(g)
    o.set_priors(**pri)
    o.update(True)
    H1 = oracle_blocks(o)
    cams, pts = o.get_state_rt()
    r = _prior_residuals(cams, pts, pri)
    assert abs((H1["chi2"] - H0["chi2"]) - float(r @ r)) <= 1e-12 * H1["chi2"]
    NU, NV = o.NU, o.NV
    J = np.zeros((len(r), 6 * NU + 3 * NV))
    h = 1e-6
    for i in range(6 * NU + 3 * NV):
        for s in (1, -1):
            o.set_state_rt(cams, pts)
            xu, xv = np.zeros(6 * NU), np.zeros(3 * NV)
            (xu if i < 6 * NU else xv)[i if i < 6 * NU else i - 6 * NU] = s * h
            o.oplus(xu, xv)
            c2, p2 = o.get_state_rt()
            J[:, i] += s * _prior_residuals(c2, p2, pri) / (2 * h)
    o.set_state_rt(cams, pts)
    dU = np.array([H1["U"][u] - H0["U"][u] for u in range(NU)])
    JtJ = J.T @ J
    for u in range(NU):
        assert np.abs(dU[u] - JtJ[6 * u:6 * u + 6, 6 * u:6 * u + 6]).max() <= 1e-6 * max(1.0, np.abs(dU[u]).max())
    for q in range(NV):
        blk = JtJ[6 * NU + 3 * q:6 * NU + 3 * q + 3, 6 * NU + 3 * q:6 * NU + 3 * q + 3]
        assert np.abs((H1["V"][q] - H0["V"][q]) - blk).max() <= 1e-6 * max(1.0, np.abs(blk).max())
    Jtr = J.T @ r
    assert np.abs((H1["bU"] - H0["bU"]) - Jtr[:6 * NU]).max() <= 1e-6 * max(1.0, np.abs(Jtr).max())
    assert np.abs((H1["bV"] - H0["bV"]) - Jtr[6 * NU:]).max() <= 1e-6 * max(1.0, np.abs(Jtr).max())
    assert np.array_equal(H1["W"], H0["W"])


# ---------------------------------------------------------------------------
# per-edge weights and the Huber loss of the BA edge, pinned independently:
# the cost from a numpy projection with a numpy Huber, its gradient by central
# differences under the section's oplus, the blocks from a dense assembly
# with the IRLS weight w * psi computed in numpy
# ---------------------------------------------------------------------------
def _huber_rho(s, delta):
    """rho(s) of the squared residual norm s: s inside delta^2, the linear
    branch 2 delta sqrt(s) - delta^2 outside."""
    s = np.asarray(s, np.float64)
    return np.where(s <= delta * delta, s, 2.0 * delta * np.sqrt(s) - delta * delta)


def _huber_psi(s, delta):
    """rho'(s): 1 inside, delta / sqrt(s) outside."""
    s = np.asarray(s, np.float64)
    return np.where(s <= delta * delta, 1.0, delta / np.sqrt(np.maximum(s, 1e-300)))


def _weighted_huber_problem():
    g = generate_ba(small_cfg(n_cameras=5, n_points=40, k_min=2, k_max=4, seed=11, outlier_frac=0.3,
                              outlier_sigma=8.0))
    # at the ground truth: N(0, 1 px) residuals mostly inside delta, the
    # outliers outside
    g = dict(g, cams0=g["cams_gt"].copy(), pts0=g["pts_gt"].copy())
    rng = np.random.default_rng(12)
    w = rng.uniform(0.3, 2.0, g["n_edges"])
    return g, w, 2.0


def _cams_rt_from_aa(cams6):
    C = len(cams6) // 6
    cr = np.zeros(12 * C)
    for c in range(C):
        cr[12 * c:12 * c + 9] = Rotation.from_rotvec(cams6[6 * c:6 * c + 3]).as_matrix().reshape(-1)
        cr[12 * c + 9:12 * c + 12] = cams6[6 * c + 3:6 * c + 6]
    return cr


def _numpy_cost(g, w, delta, cams_rt, pts):
    r = _numpy_residuals(g, cams_rt, pts)
    return float(np.sum(w * _huber_rho(np.sum(r * r, axis=1), delta)))


def test_weighted_huber_cost_and_residual_split():
    """The oracle's cost equals sum_e w_e rho(|r_e|^2) from an independent
    numpy projection and Huber function, with residuals on both sides of delta."""
    g, w, delta = _weighted_huber_problem()
    o = OracleBA.from_generated(g, weight=w, huber_delta=delta)
    cr = _cams_rt_from_aa(g["cams0"])
    r = _numpy_residuals(g, cr, g["pts0"])
    s = np.sum(r * r, axis=1)
    assert np.sum(s > delta * delta) >= 10 and np.sum(s <= delta * delta) >= 10
    assert abs(o.update(False) - _numpy_cost(g, w, delta, cr, g["pts0"])) < 1e-12 * o.update(False)


def test_weighted_huber_gradient_is_half_the_cost_gradient():
    """b = sum_e w_e psi_e J_e^T r_e is half the gradient of the cost under
    the section's perturbations (left SE(3) update of the variable cameras,
    additive points): central differences (h = 1e-6) of the numpy cost."""
    g, w, delta = _weighted_huber_problem()
    o = OracleBA.from_generated(g, weight=w, huber_delta=delta)
    o.update(True)
    H = o.hessian()
    cams, pts = o.get_state_rt()
    h = 1e-6
    fd_u = np.zeros(6 * o.NU)
    for c in range(1, g["n_cameras"]):  # camera 0 is fixed
        for k in range(6):
            vals = []
            for sgn in (1.0, -1.0):
                d = np.zeros(6)
                d[k] = sgn * h
                Rm, tm = cams[12 * c:12 * c + 9].copy(), cams[12 * c + 9:12 * c + 12].copy()
                orc.lib().or_se3_left_update(orc.dp(d), orc.dp(Rm), orc.dp(tm))
                cc = cams.copy()
                cc[12 * c:12 * c + 9], cc[12 * c + 9:12 * c + 12] = Rm, tm
                vals.append(_numpy_cost(g, w, delta, cc, pts))
            fd_u[6 * (c - 1) + k] = (vals[0] - vals[1]) / (2 * h)
    fd_v = np.zeros(3 * o.NV)
    for q in range(1, g["n_points"]):  # point 0 is fixed
        for k in range(3):
            pp, pm = pts.copy(), pts.copy()
            pp[3 * q + k] += h
            pm[3 * q + k] -= h
            fd_v[3 * (q - 1) + k] = (_numpy_cost(g, w, delta, cams, pp) - _numpy_cost(g, w, delta, cams, pm)) / (2 * h)
    grad = 2.0 * np.concatenate([H["bU"], H["bV"]])
    assert relerr(grad, np.concatenate([fd_u, fd_v])) < 1e-6


def test_weighted_huber_blocks_equal_dense_irls_assembly():
    """U, V, W, bU, bV equal J^T Omega J and J^T Omega r with
    Omega = diag(w_e psi(|r_e|^2)) computed in numpy (J from the oracle's
    per-edge Jacobians, which central differences pin separately)."""
    g, w, delta = _weighted_huber_problem()
    o = OracleBA.from_generated(g, weight=w, huber_delta=delta)
    o.update(True)
    r, Jc, Jp = o.edge_jacobians()
    cr = _cams_rt_from_aa(g["cams0"])
    assert relerr(r, _numpy_residuals(g, cr, g["pts0"])) < 1e-10
    om = w * _huber_psi(np.sum(r * r, axis=1), delta)
    NU, NV = o.NU, o.NV
    J = np.zeros((2 * g["n_edges"], 6 * NU + 3 * NV))
    for e in range(g["n_edges"]):
        c, q = g["cam_idx"][e], g["pt_idx"][e]
        if c > 0:
            J[2 * e:2 * e + 2, 6 * (c - 1):6 * c] = Jc[e]
        if q > 0:
            J[2 * e:2 * e + 2, 6 * NU + 3 * (q - 1):6 * NU + 3 * q] = Jp[e]
    Om = np.repeat(om, 2)
    Hd, b = dense_system(oracle_blocks(o), NU, NV)
    assert relerr(Hd, J.T @ (Om[:, None] * J)) < 1e-12
    assert relerr(b, J.T @ (Om * r.reshape(-1))) < 1e-12
    # an unweighted, Huber-free build of the same problem gives different blocks
    o2 = OracleBA.from_generated(g)
    o2.update(True)
    Hd2, _ = dense_system(oracle_blocks(o2), NU, NV)
    assert relerr(Hd2, Hd) > 1e-3


@pytest.mark.parametrize("wt", [1.0, 0.7])
def test_huber_continuous_at_delta(wt):
    """rho and the gradient are continuous where |r| crosses delta: one
    camera, one point, the observation placed at |r| = delta (1 -+ 1e-9)."""
    delta = 2.0
    cams6 = np.array([0.1, -0.2, 0.05, 0.3, -0.1, 6.0])
    pt = np.array([0.2, 0.1, 0.5])
    cr = _cams_rt_from_aa(cams6)
    Xc = cr[:9].reshape(3, 3) @ pt + cr[9:]
    proj = 500.0 * Xc[:2] / Xc[2] + np.array([320.0, 240.0])
    vals = []
    for eps in (-1e-9, 1e-9):
        z = proj - delta * (1.0 + eps) * np.array([0.6, 0.8])
        o = OracleBA(2, 2, [1], [1], z, 500.0, 320.0, 240.0, np.concatenate([np.zeros(6), cams6]),
                     np.concatenate([np.zeros(3), pt]), weight=np.array([wt]), huber_delta=delta)
        c = o.update(True)
        vals.append((c, o.hessian()["bV"].copy(), o.hessian()["bU"].copy()))
    (c0, bv0, bu0), (c1, bv1, bu1) = vals
    assert abs(c0 - wt * delta ** 2) < 1e-8 and abs(c1 - wt * delta ** 2) < 1e-8
    assert relerr(bv1, bv0) < 1e-8 and relerr(bu1, bu0) < 1e-8


def test_schur_assemble_is_the_steps_reduction(cfg1):
    """or_schur_assemble and or_schur_step share one reduction: the same S
    and bS bit for bit."""
    o = OracleBA.from_generated(cfg1, huber_delta=2.0)
    o.update(True)
    a = o.schur_assemble(0.01, 3.0)
    s = o.schur_step(0.01, 3.0, want_s=True)
    assert a["ok"] and s["ok"]
    assert np.array_equal(a["S"], s["S"]) and np.array_equal(a["bS"], s["bS"])


# ---------------------------------------------------------------------------
# V with off-diagonal blocks (V-V edges, nelson/DoubleSection.h:498-510):
# the Schur step with the whole V factorised (SolverCholeskySchur.hpp:97-147)
# equals a dense solve of the damped full system
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("BU,BV,rel,ab", [(2, 3, 0.0, 0.0), (6, 3, 1e-3, 0.2), (3, 2, 0.5, 0.0)])
def test_schur_step_with_vv_blocks_equals_dense_solve(BU, BV, rel, ab):
    rng = np.random.default_rng(BU * 10 + BV)
    NU, NV = 5, 7
    wpairs = [(i, j) for i in range(NU) for j in range(NV) if (i + 2 * j) % 3 != 0]
    vpairs = [(i, j) for i in range(NV) for j in range(i + 1, NV) if (i * j + i) % 4 != 1]
    p = PyProblem(BU, BV, NU, NV, [a for a, _ in wpairs], [b for _, b in wpairs], vpairs)
    nu, nv = BU * NU, BV * NV
    n = nu + nv
    H = np.zeros((n, n))
    J = rng.standard_normal((3 * n, n))
    dense = J.T @ J + n * np.eye(n)
    mask = np.zeros((n, n), bool)  # the block pattern: U dense, V blocks, W pairs
    mask[:nu, :nu] = True
    for q in range(NV):
        mask[nu + BV * q:nu + BV * q + BV, nu + BV * q:nu + BV * q + BV] = True
    for i, j in wpairs:
        mask[BU * i:BU * i + BU, nu + BV * j:nu + BV * j + BV] = True
    for i, j in vpairs:
        mask[nu + BV * i:nu + BV * i + BV, nu + BV * j:nu + BV * j + BV] = True
    mask = mask | mask.T
    H[mask] = dense[mask]
    H += 4 * n * np.eye(n)  # diagonally dominant: SPD with every pattern
    b = rng.standard_normal(n)
    v = p.views()
    v["U"][...] = np.triu(H[:nu, :nu])
    for q in range(NV):
        v["V"][q] = H[nu + BV * q:nu + BV * q + BV, nu + BV * q:nu + BV * q + BV]
    for k, (i, j) in enumerate(wpairs):
        v["W"][k] = H[BU * i:BU * i + BU, nu + BV * j:nu + BV * j + BV]
    for k, (i, j) in enumerate(vpairs):
        v["Vo"][k] = H[nu + BV * i:nu + BV * i + BV, nu + BV * j:nu + BV * j + BV]
    v["bU"][:] = b[:nu]
    v["bV"][:] = b[nu:]
    xU, xV = np.zeros(nu), np.zeros(nv)
    S, bS = np.zeros((nu, nu)), np.zeros(nu)
    ok = orc.lib().or_schur_step(orc.C.byref(p.h), rel, ab, orc.dp(xU), orc.dp(xV), orc.dp(S), orc.dp(bS))
    assert ok
    Hd = H + np.diag(rel * np.diag(H) + ab)
    x = np.linalg.solve(Hd, -b)
    assert relerr(np.concatenate([xU, xV]), x) < 1e-12
    # S and bS are the Schur complement of the damped V block and its reduced
    # right-hand side
    A, B, D = Hd[:nu, :nu], Hd[:nu, nu:], Hd[nu:, nu:]
    Sref = A - B @ np.linalg.solve(D, B.T)
    assert relerr(np.triu(S), np.triu(Sref)) < 1e-12
    assert relerr(bS, -b[:nu] + B @ np.linalg.solve(D, b[nu:])) < 1e-12
    # a V-V block left out changes the step: the blocks are used
    v["Vo"][0] = 0.0
    ok = orc.lib().or_schur_step(orc.C.byref(p.h), rel, ab, orc.dp(xU), orc.dp(xV), None, None)
    assert ok and relerr(np.concatenate([xU, xV]), x) > 1e-6
