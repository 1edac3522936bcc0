"""The BAL camera model and file format on the host: the oracle's BAL edge
(or_ba_set_intrinsics) pinned against a numpy restatement of the model,
central differences of its residual, and the identity (to rounding) that
with k1 = k2 = 0 it is the pinhole of focal -f at principal point (0, 0);
nsn_bal_load against the file written from the same arrays, and its error
cases."""
import numpy as np
import pytest

from helpers import bal_problem, bal_project, relerr, small_cfg, write_bal
from nelson_amd import NelsonError, load_bal
from oracle import oracle as orc
from oracle.oracle import OracleBA


@pytest.fixture(scope="module")
def bal():
    return bal_problem(small_cfg(n_cameras=5, n_points=40, k_min=2, k_max=4, seed=8))


def _cams_rt(o):
    cams, pts = o.get_state_rt()
    return cams, pts


def test_oracle_bal_residuals_and_jacobians(bal):
    g = bal
    o = OracleBA.from_generated(g)
    r, Jc, Jp = o.edge_jacobians()
    uv = bal_project(g["cams0"], g["intrinsics"], g["pts0"], g["cam_idx"], g["pt_idx"])
    assert relerr(r, uv - g["obs"].reshape(-1, 2)) < 1e-12
    cams, pts = o.get_state_rt()
    h = 1e-6
    for e in range(g["n_edges"]):
        c, q = g["cam_idx"][e], g["pt_idx"][e]
        for k in range(6):
            res = []
            for s in (1.0, -1.0):
                d = np.zeros(6)
                d[k] = s * h
                Rm, tm = cams[12 * c:12 * c + 9].copy(), cams[12 * c + 9:12 * c + 12].copy()
                orc.lib().or_se3_left_update(orc.dp(d), orc.dp(Rm), orc.dp(tm))
                cc = cams.copy()
                cc[12 * c:12 * c + 9], cc[12 * c + 9:12 * c + 12] = Rm, tm
                o.set_state_rt(cc, pts)
                res.append(o.edge_jacobians()[0][e])
            assert np.allclose(Jc[e][:, k], (res[0] - res[1]) / (2 * h), rtol=1e-6, atol=1e-5)
        for k in range(3):
            res = []
            for s in (1.0, -1.0):
                pp = pts.copy()
                pp[3 * q + k] += s * h
                o.set_state_rt(cams, pp)
                res.append(o.edge_jacobians()[0][e])
            assert np.allclose(Jp[e][:, k], (res[0] - res[1]) / (2 * h), rtol=1e-6, atol=1e-5)
        o.set_state_rt(cams, pts)


def test_oracle_bal_without_distortion_is_a_pinhole(bal):
    g = dict(bal)
    intr = g["intrinsics"].reshape(-1, 3).copy()
    intr[:, 0] = 450.0
    intr[:, 1:] = 0.0
    a = OracleBA.from_generated(g, intrinsics=intr)
    b = OracleBA.from_generated(dict(g, focal=-450.0, cx=0.0, cy=0.0), intrinsics=None)
    ca, cb = a.update(True), b.update(True)
    assert abs(ca - cb) <= 1e-14 * cb
    ra, Jca, Jpa = a.edge_jacobians()
    rb, Jcb, Jpb = b.edge_jacobians()
    assert relerr(ra, rb) < 1e-14 and relerr(Jca, Jcb) < 1e-14 and relerr(Jpa, Jpb) < 1e-14


def test_bal_file_round_trip(bal, tmp_path):
    f = tmp_path / "problem.txt"
    write_bal(f, bal)
    d = load_bal(f)
    assert (d["n_cameras"], d["n_points"], d["n_edges"]) == (bal["n_cameras"], bal["n_points"], bal["n_edges"])
    for k in ("cam_idx", "pt_idx", "obs", "cams0", "pts0", "intrinsics"):
        assert np.array_equal(d[k], np.asarray(bal[k]).reshape(d[k].shape)), k


def test_bal_file_errors(tmp_path):
    ok = "2 2 2\n0 0 1.5 2.5\n1 1 -1 0.25\n" + "\n".join(["0.1"] * 18) + "\n" + "\n".join(["1"] * 6) + "\n"
    p = tmp_path / "ok.txt"
    p.write_text(ok)
    d = load_bal(p)
    assert d["obs"].tolist() == [1.5, 2.5, -1.0, 0.25] and d["pts0"].tolist() == [1.0] * 6
    bad = {
        "missing": None,
        "header": "2 2\n",
        "camera range": ok.replace("1 1 -1", "2 1 -1"),
        "point range": ok.replace("1 1 -1", "1 2 -1"),
        "short observation": "2 2 2\n0 0 1.5 2.5\n1 1 -1\n",
        "short cameras": "2 2 2\n0 0 1.5 2.5\n1 1 -1 0.25\n" + "\n".join(["0.1"] * 17) + "\n",
        "short points": ok[:ok.rstrip().rfind("\n")] + "\n",
        "malformed": ok.replace("0.25", "x"),
    }
    for name, text in bad.items():
        q = tmp_path / f"bad_{name.replace(' ', '_')}.txt"
        if text is not None:
            q.write_text(text)
        with pytest.raises(NelsonError):
            load_bal(q)
