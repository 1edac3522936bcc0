"""A user section with user edges through the C++ interface
(include/nelson_amd.hpp: DoubleSection / EdgeUnaryU / EdgeUnaryV /
EdgeBinaryUU / EdgeBinaryUV, SolverCholeskySchur::init(H, b) /
computeIncrement(H, b, rel, abs), the host LevenbergMarquardt / GaussNewton
loops), on the linear problem of the reference's own tests
(nelson/tests/testDoubleSectionCommon.h, with the block-diagonal second
section of its Schur tests, nelson/tests/testDoubleSectionSchurCommon.hpp):
the program tests/cpp/test_section reads the problem written here and its
results must match the oracle's drivers on the same problem, converge to
chi2 < 1e-12 as the reference's tests require, and not depend on the internal
order of the U parameters (reversed, minimum degree)."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import PyProblem

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_section")


class Toy(PyProblem):
    """The oracle's copy of the problem: 2-D points (U), 3-D points (V),
    unary edges on both, U-U difference edges on every pair (uu), V-V
    difference edges on every pair (vv: V is then not block diagonal), U-V
    edges on every pair (wall) or on (i, i) only; edges in the reference's
    order."""

    def __init__(self, N2, N3, uu, wall, seed, vv=False, sigma=0.5):
        self.N2, self.N3, self.uu, self.wall, self.vv = N2, N3, uu, wall, vv
        self.pairs = [(i, j) for i in range(N2) for j in range(N3) if wall or i == j]
        self.vpairs = [(i, j) for i in range(N3) for j in range(i + 1, N3)] if vv else []
        super().__init__(2, 3, N2, N3, [p[0] for p in self.pairs], [p[1] for p in self.pairs], self.vpairs)
        rng = np.random.default_rng(seed)
        self.gt2 = rng.uniform(-1, 1, (N2, 2))
        self.gt3 = rng.uniform(-1, 1, (N3, 3))
        self.p2 = self.gt2 + rng.uniform(-1, 1, (N2, 2)) * sigma
        self.p3 = self.gt3 + rng.uniform(-1, 1, (N3, 3)) * sigma

    def write(self, path):
        with open(path, "w") as f:
            f.write(f"{self.N2} {self.N3} {int(self.uu)} {int(self.wall)} {int(self.vv)}\n")
            for a in (self.gt2, self.gt3, self.p2, self.p3):
                for row in a:
                    f.write(" ".join(f"{v:.17g}" for v in row) + "\n")

    def update(self, hess):
        v = self.views()
        chi2 = 0.0
        if hess:
            for k in ("U", "V", "W", "bU", "bV") + (("Vo",) if self.vv else ()):
                v[k][...] = 0
        U, V, W, bU, bV = v["U"], v["V"], v["W"], v["bU"], v["bV"]
        for i in range(self.N2):
            e = self.p2[i] - self.gt2[i]
            chi2 += e @ e
            if hess:
                U[2 * i:2 * i + 2, 2 * i:2 * i + 2] += np.eye(2)
                bU[2 * i:2 * i + 2] += e
        for j in range(self.N3):
            e = self.p3[j] - self.gt3[j]
            chi2 += e @ e
            if hess:
                V[j] += np.eye(3)
                bV[3 * j:3 * j + 3] += e
        if self.uu:
            for i in range(self.N2):
                for j in range(i + 1, self.N2):
                    e = (self.p2[i] - self.p2[j]) - (self.gt2[i] - self.gt2[j])
                    chi2 += e @ e
                    if hess:
                        U[2 * i:2 * i + 2, 2 * i:2 * i + 2] += np.eye(2)
                        U[2 * j:2 * j + 2, 2 * j:2 * j + 2] += np.eye(2)
                        U[2 * i:2 * i + 2, 2 * j:2 * j + 2] -= np.eye(2)
                        bU[2 * i:2 * i + 2] += e
                        bU[2 * j:2 * j + 2] -= e
        for k, (i, j) in enumerate(self.vpairs):
            e = (self.p3[i] - self.p3[j]) - (self.gt3[i] - self.gt3[j])
            chi2 += e @ e
            if hess:
                V[i] += np.eye(3)
                V[j] += np.eye(3)
                v["Vo"][k] -= np.eye(3)
                bV[3 * i:3 * i + 3] += e
                bV[3 * j:3 * j + 3] -= e
        for k, (i, j) in enumerate(self.pairs):
            e = (self.p2[i] - self.p3[j][:2]) - (self.gt2[i] - self.gt3[j][:2])
            chi2 += e @ e
            if hess:
                U[2 * i:2 * i + 2, 2 * i:2 * i + 2] += np.eye(2)
                bU[2 * i:2 * i + 2] += e
                V[j] += np.diag([1.0, 1.0, 0.0])
                bV[3 * j:3 * j + 3] -= np.array([e[0], e[1], 0.0])
                W[k] += np.array([[-1.0, 0, 0], [0, -1.0, 0]])
        self.h.chi2 = chi2

    def oplus(self, xu, xv):
        self.p2 += xu.reshape(self.N2, 2)
        self.p3 += xv.reshape(self.N3, 3)

    def backup(self):
        self.bk = (self.p2.copy(), self.p3.copy())

    def rollback(self):
        self.p2, self.p3 = self.bk[0].copy(), self.bk[1].copy()

    def dense(self):
        """H (upper U read as symmetric) and b of the current linearisation"""
        v = self.views()
        nu, nv = 2 * self.N2, 3 * self.N3
        H = np.zeros((nu + nv, nu + nv))
        Uu = np.triu(v["U"])
        H[:nu, :nu] = Uu + np.triu(Uu, 1).T
        for j in range(self.N3):
            H[nu + 3 * j:nu + 3 * j + 3, nu + 3 * j:nu + 3 * j + 3] = v["V"][j]
        for k, (i, j) in enumerate(self.pairs):
            H[2 * i:2 * i + 2, nu + 3 * j:nu + 3 * j + 3] = v["W"][k]
            H[nu + 3 * j:nu + 3 * j + 3, 2 * i:2 * i + 2] = v["W"][k].T
        for k, (i, j) in enumerate(self.vpairs):
            H[nu + 3 * i:nu + 3 * i + 3, nu + 3 * j:nu + 3 * j + 3] = v["Vo"][k]
            H[nu + 3 * j:nu + 3 * j + 3, nu + 3 * i:nu + 3 * i + 3] = v["Vo"][k].T
        return H, np.concatenate([v["bU"], v["bV"]])


# (N2, N3, U-U edges, U-V on every pair, seed, V-V edges)
CASES = [(5, 3, True, True, 0, False), (5, 3, False, True, 1, False), (5, 3, True, False, 2, False),
         (7, 4, True, True, 3, False), (5, 3, False, False, 4, False),
         (5, 3, True, True, 5, True), (4, 6, False, True, 6, True), (5, 4, True, False, 7, True),
         (3, 7, False, False, 8, True)]


def _run(toy, tmp_path):
    path = tmp_path / "toy.txt"
    toy.write(str(path))
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (nelson_amd.build builds it with the library)")
    r = subprocess.run([EXE, str(path)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return {d["solve"]: d for d in (json.loads(line) for line in r.stdout.splitlines() if line.strip())}


SUCCESS = {"SuccessSmallIncrement", "SuccessEpsB", "SuccessSmallResiduals"}


def _same_logs(a, b, what, exact_length=False):
    """entries equal in (it, subit) and in chi2 / mu to 1e-9 relative while
    the cost is above the rounding floor (1e-20); the linear problem reaches
    chi2 ~ 1e-30 in one step, after which the rounding-level stopping tests
    (|dx|^2 <= 1e-30, max |b| <= 1e-15) may stop either run one entry
    earlier"""
    if exact_length:
        assert len(a) == len(b), (what, a, b)
    for x, y in zip(a, b):
        assert x[0] == y[0] and x[1] == y[1], (what, a, b)
        assert abs(x[2] - y[2]) <= 1e-9 * abs(y[2]) + 1e-20, (what, a, b)
        assert abs(x[3] - y[3]) <= 1e-9 * abs(y[3]) + 1e-300, (what, a, b)
        if y[2] < 1e-20:
            return
    assert len(a) == len(b), (what, a, b)


@pytest.mark.parametrize("N2,N3,uu,wall,seed,vv", CASES)
def test_user_section_matches_oracle(N2, N3, uu, wall, seed, vv, tmp_path):
    res = _run(Toy(N2, N3, uu, wall, seed, vv), tmp_path)
    for name, solve in (("gn", "gn_solve"), ("lm", "lm_solve")):
        t = Toy(N2, N3, uu, wall, seed, vv)
        ro = getattr(t, solve)()
        d = res[name]
        assert d["reason"] == ro["reason"] or {d["reason"], ro["reason"]} <= SUCCESS, (name, d["reason"], ro["reason"])
        _same_logs(d["log"], ro["log"], name)
        assert d["final_chi2"] < 1e-12, d  # the reference's REQUIRE(chi2 < dummy_precision)
        assert np.allclose(d["p2"], t.p2.ravel(), rtol=0, atol=1e-12)
        assert np.allclose(d["p3"], t.p3.ravel(), rtol=0, atol=1e-12)
    # stopping tests off: the damping schedule of a linear problem (rho = 1,
    # mu x 1/3 per accepted step, LevenbergMarquardt.hpp:109-112)
    t = Toy(N2, N3, uu, wall, seed, vv)
    ro = t.lm_solve(maxNumIt=3, epsIncSquare=-1.0, epsChi2=0.0, epsBVector=0.0)
    d = res["lm_nostop"]
    assert d["reason"] == ro["reason"] == "MaxIterationReached"
    _same_logs(d["log"], ro["log"], "lm_nostop", exact_length=True)
    mus = [e[3] for e in d["log"][1:]]
    for a, b in zip(mus, mus[1:]):
        assert abs(b - a / 3.0) <= 1e-12 * a


@pytest.mark.parametrize("N2,N3,uu,wall,seed,vv", CASES)
def test_internal_order_of_u(N2, N3, uu, wall, seed, vv, tmp_path):
    """setUser2InternalIndexesU (reversed) and permuteAMD_SchurU change the
    internal order of bU and of the increment's U part, not the solution."""
    res = _run(Toy(N2, N3, uu, wall, seed, vv), tmp_path)
    base = res["lm"]
    rev = res["lm_reversed"]
    assert rev["u2i"] == list(range(N2 - 1, -1, -1))
    md = res["lm_min_degree"]
    assert sorted(md["u2i"]) == list(range(N2))
    for d in (rev, md):
        assert d["reason"] == base["reason"] or {d["reason"], base["reason"]} <= SUCCESS
        _same_logs(d["log"], base["log"], "order")
        assert np.allclose(d["p2"], base["p2"], rtol=0, atol=1e-12)
        assert np.allclose(d["p3"], base["p3"], rtol=0, atol=1e-12)
    # one damped step in the reversed order against a dense solve of the
    # oracle's Hessian: (H + diag(rel H_ii + abs)) dx = -b
    t = Toy(N2, N3, uu, wall, seed, vv)
    t.update(True)
    H, b = t.dense()
    x = np.linalg.solve(H + np.diag(0.05 * np.diag(H) + 0.3), -b)
    s = res["step"]
    assert s["ok"]
    du = np.array(s["du"]).reshape(N2, 2)
    du_user = np.concatenate([du[s["u2i"][i]] for i in range(N2)])
    got = np.concatenate([du_user, s["dv"]])
    assert np.abs(got - x).max() <= 1e-12 * np.abs(x).max()
    assert abs(s["sq_norm"] - x @ x) <= 1e-12 * (x @ x)
    assert abs(s["chi2"] - t.h.chi2) <= 1e-14 * t.h.chi2
    diag = np.abs(np.diag(H)).max()
    assert abs(s["max_abs_h_diag"] - diag) <= 1e-15 * diag
    assert abs(s["max_abs_b"] - np.abs(b).max()) <= 1e-15 * np.abs(b).max()
