"""The sparse-S path: S stored and factorised over its tile envelope (tile
row i from its first non-zero tile column; the factor has no fill outside
it).
  * The tiles outside the envelope are exact zeros of S and of the factor,
    so the envelope layout and the dense one (NSN_DENSE_S=1) give the same
    increments to rounding (1e-12 normwise; the dense layout is the one
    pinned against the oracle at full size): on the 1000-camera ring (about
    half the tiles) and on an open camera path.  Not bit for bit: a sum of
    tile products longer than 24 terms is cut into chunks added in order,
    and the dense layout's sums (from column 0) are cut where the
    envelope's (from their first non-zero tile) are not.
  * An open path of 4000 cameras / 400k points / 2.4M observations (n =
    23 994, about 3 % of the dense tiles): the FP64 increment solves the
    oracle's damped normal equations to a componentwise backward error of
    1e-12, the mixed one to 1e-10 after two refinement steps (the size-independent
    check of the full-size tests).
  * A 300-camera open path against the oracle's own solve (its pivoted LDL^T
    of the dense S): increments to 1e-9 and the LM log to 1e-9."""
import os

import numpy as np
import pytest

from helpers import oracle_blocks, relerr
from nelson_amd import CONFIGS, BAProblem, SolverCholeskySchur, generate_ba
from oracle.oracle import OracleBA
from test_gpu_fullsize import _backward_error

pytestmark = pytest.mark.gpu


def _problem(g, precision="fp64", dense=False):
    old = os.environ.get("NSN_DENSE_S")
    os.environ["NSN_DENSE_S"] = "1" if dense else "0"
    try:
        return BAProblem.from_generated(g, precision=precision)
    finally:
        if old is None:
            del os.environ["NSN_DENSE_S"]
        else:
            os.environ["NSN_DENSE_S"] = old


def _increments(p, dampings):
    p.update(True)
    s = SolverCholeskySchur()
    s.init(p)
    out = []
    for rel, ab in dampings:
        assert s.computeIncrement(rel, ab * s.maxAbsHDiag())
        out.append(np.concatenate(s.incrementVector()))
    return out


@pytest.mark.parametrize("name,cfg", [
    ("ring", dict(CONFIGS["cfg2"])),
    ("path", dict(CONFIGS["seq"], n_cameras=600, n_points=60000, arc=30, seed=21)),
])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_envelope_equals_dense(name, cfg, precision):
    g = generate_ba(cfg)
    pe, pd = _problem(g, precision), _problem(g, precision, dense=True)
    assert pd.n_tiles == pd.n_tiles_dense
    assert pe.n_tiles < pd.n_tiles, (pe.n_tiles, pd.n_tiles)
    damp = [(0.0, 1e-4), (1e-3, 0.0)]
    for a, b in zip(_increments(pe, damp), _increments(pd, damp)):
        e = relerr(a, b)
        print(f"\n{name} {precision}: {pe.n_tiles} of {pd.n_tiles} tiles, envelope vs dense {e:.2e}")
        assert e <= (1e-12 if precision == "fp64" else 1e-6), e


@pytest.fixture(scope="module")
def seq():
    g = generate_ba(CONFIGS["seq"])
    o = OracleBA.from_generated(g)
    o.update(True)
    return g, oracle_blocks(o)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_open_path_4000_cameras_solves_normal_equations(seq, precision):
    g, Ho = seq
    p = BAProblem.from_generated(g, precision=precision)
    if precision == "mixed":
        p.set_refinement(2)
    assert p.n_tiles < 0.05 * p.n_tiles_dense, (p.n_tiles, p.n_tiles_dense)
    p.update(True)
    s = SolverCholeskySchur()
    s.init(p)
    mu = 1e-4 * s.maxAbsHDiag()
    for rel, ab in ((0.0, mu), (1e-3, 0.0)):
        assert s.computeIncrement(rel, ab)
        dxu, dxv = s.incrementVector()
        w = _backward_error(Ho, dxu, dxv, rel, ab)
        print(f"\n{precision} n={p.n_s} tiles {p.n_tiles}/{p.n_tiles_dense} rel={rel}: backward error {w:.2e}")
        assert w <= (1e-12 if precision == "fp64" else 1e-10), w


def test_open_path_increment_and_lm_vs_oracle():
    from nelson_amd import LevenbergMarquardt, LevenbergMarquardtSettings
    g = generate_ba(dict(CONFIGS["seq"], n_cameras=300, n_points=30000, arc=20, seed=23))
    o = OracleBA.from_generated(g)
    ro = o.lm_solve(maxNumIt=3, dx_cap=3 * (6 * o.NU + 3 * o.NV))
    p = BAProblem.from_generated(g)
    assert p.n_tiles < 0.5 * p.n_tiles_dense
    p.update(True)
    s = SolverCholeskySchur()
    s.init(p)
    mu0 = 1e-3 * s.maxAbsHDiag()
    assert abs(mu0 - ro["log"][1][3]) <= 1e-12 * mu0
    assert s.computeIncrement(0.0, mu0)
    got = np.concatenate(s.incrementVector())
    assert relerr(got, ro["dx"][0]) <= 1e-9, relerr(got, ro["dx"][0])
    p2 = BAProblem.from_generated(g)
    lm = LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=3))
    lm.solve(p2)
    log = lm.stats()
    assert len(log) == len(ro["log"])
    for a, b in zip(log, ro["log"]):
        assert a[0] == b[0] and a[1] == b[1]
        assert abs(a[2] - b[2]) <= 1e-9 * abs(b[2]) and abs(a[3] - b[3]) <= 1e-9 * abs(b[3])


def _shuffled(g, seed):
    """The same problem with the cameras renumbered at random (camera 0, the
    fixed gauge, keeps its number): the open path's banded S becomes a
    scattered one in the given order."""
    rng = np.random.default_rng(seed)
    C = g["n_cameras"]
    perm = np.r_[0, 1 + rng.permutation(C - 1)]  # new index of each camera
    h = dict(g)
    h["cam_idx"] = perm[g["cam_idx"]].astype(np.int32)
    for k in ("cams_gt", "cams0"):
        c = np.zeros_like(g[k].reshape(C, 6))
        c[perm] = g[k].reshape(C, 6)
        h[k] = c.reshape(-1)
    return h


def test_rcm_order_recovers_the_envelope_and_keeps_the_interface_order():
    """Reverse Cuthill-McKee on a camera-shuffled open path: the envelope of S
    shrinks to that of the unshuffled problem (within 2x) from near dense; the
    Hessian exported in camera order equals the oracle's; the increments (in
    camera order) solve the oracle's damped normal equations and equal the
    given-order solve's to rounding; the LM log follows the oracle's."""
    from nelson_amd import LevenbergMarquardt, LevenbergMarquardtSettings
    g0 = generate_ba(dict(CONFIGS["seq"], n_cameras=600, n_points=60000, arc=30, seed=25))
    g = _shuffled(g0, 5)
    base = BAProblem.from_generated(g0)
    given = BAProblem.from_generated(g)
    rcm = BAProblem.from_generated(g, camera_order="rcm")
    print(f"\ntiles: open path {base.n_tiles}, shuffled {given.n_tiles}, shuffled + RCM {rcm.n_tiles} "
          f"(dense {rcm.n_tiles_dense})")
    assert given.n_tiles > 0.5 * given.n_tiles_dense
    assert rcm.n_tiles <= 2 * base.n_tiles
    o = OracleBA.from_generated(g)
    o.update(True)
    Ho = oracle_blocks(o)
    rcm.update(True)
    Hg = rcm.export_hessian()
    for k in ("U", "V", "W", "bU", "bV"):
        assert relerr(Hg[k], Ho[k]) <= 1e-11, k
    assert np.array_equal(Hg["Wi"], Ho["Wi"]) and np.array_equal(Hg["Wj"], Ho["Wj"])
    given.update(True)
    s1, s2 = SolverCholeskySchur(), SolverCholeskySchur()
    s1.init(given)
    s2.init(rcm)
    mu = 1e-4 * s1.maxAbsHDiag()
    assert s1.computeIncrement(0.0, mu) and s2.computeIncrement(0.0, mu)
    a, b = s1.incrementVector(), s2.incrementVector()
    assert relerr(np.concatenate(b), np.concatenate(a)) <= 1e-10
    w = _backward_error(Ho, b[0], b[1], 0.0, mu)
    assert w <= 1e-12, w
    ro = o.lm_solve(maxNumIt=3, dx_cap=3 * (6 * o.NU + 3 * o.NV))
    p = BAProblem.from_generated(g, camera_order="rcm")
    lm = LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=3))
    lm.solve(p)
    log = lm.stats()
    assert len(log) == len(ro["log"])
    for x, y in zip(log, ro["log"]):
        assert x[0] == y[0] and x[1] == y[1]
        assert abs(x[2] - y[2]) <= 1e-9 * abs(y[2]) and abs(x[3] - y[3]) <= 1e-9 * abs(y[3])
    # the state comes back in camera order (the oracle's after the same LM)
    cg, pg = p.get_state_rt()
    co, po = o.get_state_rt()
    assert relerr(cg, co) <= 1e-8 and relerr(pg, po) <= 1e-8, (relerr(cg, co), relerr(pg, po))
