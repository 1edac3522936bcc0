"""Nested-dissection camera order (camera_order="nd"): the camera graph is
cut into independent parts and separators, aligned to the 64-row tiles of
S, so that the tile Cholesky runs the independent parts' diagonal chains at
the same time (several chain workgroups in one launch).  The order changes
the elimination order, so results equal the given order's to rounding:
  * a 300-camera ring (two or more chains): increments against the oracle's
    own solve (its pivoted LDL^T of the dense S) to 1e-9 at two dampings, the
    LM log to 1e-9, the state after LM in camera order;
  * a 600-camera open path (several chains, separators reaching over their
    subtrees): increments solve the oracle's damped normal equations to a
    componentwise backward error of 1e-12 (FP64) / 1e-10 (mixed, two
    refinement steps)
    and equal the given order's to 1e-10;
  * the 1000-camera ring and the 4000-camera path at full size: the chains
    the plan reports, the backward error of the increment (1e-12 FP64,
    1e-10 mixed with two refinement steps), and the increment against the
    given order's;
  * mixed-precision LM reaches the FP64 final cost (1e-6), the Gauss-Newton
    adjoint equals the given order's (1e-9)."""
import numpy as np
import pytest

from helpers import oracle_blocks, relerr
from nelson_amd import (CONFIGS, BAProblem, LevenbergMarquardt, LevenbergMarquardtSettings, SolverCholeskySchur,
                        generate_ba)
from oracle.oracle import OracleBA
from test_gpu_fullsize import _backward_error

pytestmark = pytest.mark.gpu


def _increments(p, dampings):
    p.update(True)
    s = SolverCholeskySchur()
    s.init(p)
    out = []
    for rel, ab in dampings:
        assert s.computeIncrement(rel, ab * s.maxAbsHDiag())
        out.append(s.incrementVector())
    return out


def _ring(n_cameras, arc, n_points, seed):
    return generate_ba(dict(CONFIGS["cfg2"], n_cameras=n_cameras, arc=arc, n_points=n_points, seed=seed))


def _path(n_cameras, arc, n_points, seed):
    return generate_ba(dict(CONFIGS["seq"], n_cameras=n_cameras, arc=arc, n_points=n_points, seed=seed))


def test_nd_ring_increments_and_lm_vs_oracle():
    g = _ring(300, 40, 30000, 31)
    info, order = BAProblem.plan(g, "nd")
    p = BAProblem.from_generated(g, camera_order="nd")
    print(f"\nring 300: {p.n_chains} chains, {p.chain_steps} steps of {p.n_s // 64 + 1}, {p.n_tiles} tiles")
    assert p.n_chains >= 2 and p.chain_steps < p.n_s // 64 + 1
    assert (info["n_chains"], info["chain_steps"], info["n_tiles"]) == (p.n_chains, p.chain_steps, p.n_tiles)
    o = OracleBA.from_generated(g)
    o.update(True)
    damp = [(0.0, 1e-4), (1e-3, 0.0)]
    got = _increments(p, damp)
    mdiag = SolverCholeskySchur()
    mdiag.init(p)
    for (rel, ab), (dxu, dxv) in zip(damp, got):
        st = o.schur_step(rel, ab * mdiag.maxAbsHDiag())
        assert st["ok"]
        e = relerr(np.concatenate([dxu, dxv]), np.concatenate([st["xU"], st["xV"]]))
        print(f"rel={rel} abs={ab}: increment vs oracle {e:.2e}")
        assert e <= 1e-9, e
    ro = OracleBA.from_generated(g).lm_solve(maxNumIt=3, dx_cap=3 * (6 * o.NU + 3 * o.NV))
    q = BAProblem.from_generated(g, camera_order="nd")
    lm = LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=3))
    lm.solve(q)
    log = lm.stats()
    assert len(log) == len(ro["log"])
    for a, b in zip(log, ro["log"]):
        assert a[0] == b[0] and a[1] == b[1]
        assert abs(a[2] - b[2]) <= 1e-9 * abs(b[2]) and abs(a[3] - b[3]) <= 1e-9 * abs(b[3])
    oo = OracleBA.from_generated(g)
    oo.lm_solve(maxNumIt=3, dx_cap=3 * (6 * o.NU + 3 * o.NV))
    cg, pg = q.get_state_rt()
    co, po = oo.get_state_rt()
    assert relerr(cg, co) <= 1e-8 and relerr(pg, po) <= 1e-8, (relerr(cg, co), relerr(pg, po))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nd_open_path_solves_normal_equations(precision):
    g = _path(600, 30, 60000, 33)
    given = BAProblem.from_generated(g, precision=precision)
    nd = BAProblem.from_generated(g, precision=precision, camera_order="nd")
    if precision == "mixed":
        given.set_refinement(2)
        nd.set_refinement(2)
    print(f"\npath 600: {nd.n_chains} chains, {nd.chain_steps} of {given.chain_steps} steps, "
          f"tiles {nd.n_tiles} (given {given.n_tiles})")
    assert nd.n_chains >= 4 and nd.chain_steps < given.chain_steps // 2
    o = OracleBA.from_generated(g)
    o.update(True)
    Ho = oracle_blocks(o)
    damp = [(0.0, 1e-4), (1e-3, 0.0)]
    s = SolverCholeskySchur()
    nd.update(True)
    s.init(nd)
    scale = s.maxAbsHDiag()
    for (rel, ab), (a, b) in zip(damp, zip(_increments(given, damp), _increments(nd, damp))):
        w = _backward_error(Ho, b[0], b[1], rel, ab * scale)
        e = relerr(np.concatenate(b), np.concatenate(a))
        print(f"{precision} rel={rel}: backward error {w:.2e}, vs given order {e:.2e}")
        assert w <= (1e-12 if precision == "fp64" else 1e-10), w
        assert e <= (1e-10 if precision == "fp64" else 1e-8), e


@pytest.mark.parametrize("name,expect_chains", [("cfg2", 4), ("seq", 16)])
def test_nd_full_size(name, expect_chains):
    g = generate_ba(dict(CONFIGS[name]))
    o = OracleBA.from_generated(g)
    o.update(True)
    Ho = oracle_blocks(o)
    given = BAProblem.from_generated(g)
    ref = _increments(given, [(0.0, 1e-4)])[0]
    for precision in ("fp64", "mixed"):
        p = BAProblem.from_generated(g, precision=precision, camera_order="nd")
        if precision == "mixed":
            p.set_refinement(2)
        print(f"\n{name} {precision}: {p.n_chains} chains, {p.chain_steps} of {given.chain_steps} steps, "
              f"{p.n_tiles} tiles (given {given.n_tiles})")
        assert p.n_chains == expect_chains
        p.update(True)
        s = SolverCholeskySchur()
        s.init(p)
        mu = 1e-4 * s.maxAbsHDiag()
        assert s.computeIncrement(0.0, mu)
        dxu, dxv = s.incrementVector()
        w = _backward_error(Ho, dxu, dxv, 0.0, mu)
        e = relerr(np.concatenate([dxu, dxv]), np.concatenate(ref))
        print(f"backward error {w:.2e}, vs given order {e:.2e}")
        assert w <= (1e-12 if precision == "fp64" else 1e-10), w
        assert e <= (1e-9 if precision == "fp64" else 1e-7), e


def test_nd_mixed_lm_and_adjoint():
    from nelson_amd.autograd import translation_loss_grad
    g = _ring(300, 40, 30000, 35)
    s = LevenbergMarquardtSettings(maxNumIt=30)
    p64 = BAProblem.from_generated(g, camera_order="nd")
    l64 = LevenbergMarquardt(s)
    l64.solve(p64)
    pm = BAProblem.from_generated(g, precision="mixed", camera_order="nd")
    lm = LevenbergMarquardt(s)
    lm.solve(pm)
    c64, cm = l64.result()["final_chi2"], lm.result()["final_chi2"]
    print(f"\nfinal cost fp64 {c64:.12e} mixed {cm:.12e}")
    assert abs(cm - c64) <= 1e-6 * c64
    # the adjoint at the FP64 solution, ND against the given order
    pg = BAProblem.from_generated(g)
    LevenbergMarquardt(s).solve(pg)
    t = pg.camera_translations()
    gc = translation_loss_grad(t, 2.0 * (t - g["cams_gt"].reshape(-1, 6)[:, 3:6]))
    cams, pts = pg.get_state()
    p64.set_state(cams, pts)
    a = pg.adjoint(g_cams=gc, hessian_steps=0)
    b = p64.adjoint(g_cams=gc, hessian_steps=0)
    for x, y in zip(a, b):
        assert relerr(y, x) <= 1e-9, relerr(y, x)
