"""Per-parameter block sizes through the C++ interface (the reference's
mat::Variable sections: parameterUSize() / parameterVSize(),
nelson/DoubleSection.h:76-110; its tests testDoubleSectionSchurDiagBlockV*.cpp):
tests/cpp/test_varsize reads the problem written here (U parameters of 1 to
3 coordinates, V parameters of 1 to 3, unary edges, difference edges on the
coordinates two parameters share, with and without V-V edges).

Checked against: the oracle's drivers on the same problem with every
parameter embedded in a 3-coordinate block (identity on the unused
diagonal, zero gradient there: the same increment on the used coordinates,
zero on the rest, the same costs), logs to 1e-9; chi2 < 1e-12 at the end
as the reference's tests require; one damped step against a dense solve of
the problem in its own sizes (1e-12)."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import PyProblem

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_varsize")


class VarToy(PyProblem):
    def __init__(self, NU, NV, vv, seed, sigma=0.5):
        rng = np.random.default_rng(seed)
        self.NU_, self.NV_, self.vv = NU, NV, vv
        self.su = [1 + (i % 3) for i in range(NU)]
        self.sv = [1 + ((j + 1) % 3) for j in range(NV)]
        self.wp = [(i, j) for i in range(NU) for j in range(NV) if (i + j) % 2 == 0]
        self.vp = [(j, j + 1) for j in range(NV - 1)] if vv else []
        super().__init__(3, 3, NU, NV, [p[0] for p in self.wp], [p[1] for p in self.wp], self.vp)
        self.gtU = [rng.uniform(-1, 1, s) for s in self.su]
        self.gtV = [rng.uniform(-1, 1, s) for s in self.sv]
        self.pU = [g + rng.uniform(-1, 1, g.size) * sigma for g in self.gtU]
        self.pV = [g + rng.uniform(-1, 1, g.size) * sigma for g in self.gtV]

    def write(self, path):
        with open(path, "w") as f:
            f.write(f"{self.NU_} {self.NV_} {int(self.vv)}\n")
            f.write(" ".join(map(str, self.su)) + "\n" + " ".join(map(str, self.sv)) + "\n")
            for a in (self.gtU, self.gtV, self.pU, self.pV):
                for row in a:
                    f.write(" ".join(f"{v:.17g}" for v in row) + "\n")

    def _edges(self):
        """(kind, a, b, m, residual) in the program's insertion order"""
        out = []
        for i in range(self.NU_):
            out.append(("u", i, None, self.su[i], self.pU[i] - self.gtU[i]))
        for j in range(self.NV_):
            out.append(("v", j, None, self.sv[j], self.pV[j] - self.gtV[j]))
        for i in range(self.NU_ - 1):
            m = min(self.su[i], self.su[i + 1])
            out.append(("uu", i, i + 1, m, (self.pU[i][:m] - self.pU[i + 1][:m]) - (self.gtU[i][:m] - self.gtU[i + 1][:m])))
        for (j, k) in self.vp:
            m = min(self.sv[j], self.sv[k])
            out.append(("vv", j, k, m, (self.pV[j][:m] - self.pV[k][:m]) - (self.gtV[j][:m] - self.gtV[k][:m])))
        for (i, j) in self.wp:
            m = min(self.su[i], self.sv[j])
            out.append(("uv", i, j, m, (self.pU[i][:m] - self.pV[j][:m]) - (self.gtU[i][:m] - self.gtV[j][:m])))
        return out

    def update(self, hess):
        v = self.views()
        if hess:
            for k in ("U", "V", "W", "bU", "bV") + (("Vo",) if self.vp else ()):
                v[k][...] = 0
            for i in range(self.NU_):  # the embedding: identity on the unused coordinates
                for k in range(self.su[i], 3):
                    v["U"][3 * i + k, 3 * i + k] = 1.0
            for j in range(self.NV_):
                for k in range(self.sv[j], 3):
                    v["V"][j][k, k] = 1.0
        chi2 = 0.0
        wk = {p: n for n, p in enumerate(self.wp)}
        vk = {p: n for n, p in enumerate(self.vp)}
        for kind, a, b, m, e in self._edges():
            chi2 += float(e @ e)
            if not hess:
                continue
            I = np.eye(m)
            if kind == "u":
                v["U"][3 * a:3 * a + m, 3 * a:3 * a + m] += I
                v["bU"][3 * a:3 * a + m] += e
            elif kind == "v":
                v["V"][a][:m, :m] += I
                v["bV"][3 * a:3 * a + m] += e
            elif kind == "uu":
                v["U"][3 * a:3 * a + m, 3 * a:3 * a + m] += I
                v["U"][3 * b:3 * b + m, 3 * b:3 * b + m] += I
                v["U"][3 * a:3 * a + m, 3 * b:3 * b + m] -= I
                v["bU"][3 * a:3 * a + m] += e
                v["bU"][3 * b:3 * b + m] -= e
            elif kind == "vv":
                v["V"][a][:m, :m] += I
                v["V"][b][:m, :m] += I
                v["Vo"][vk[(a, b)]][:m, :m] -= I
                v["bV"][3 * a:3 * a + m] += e
                v["bV"][3 * b:3 * b + m] -= e
            else:
                v["U"][3 * a:3 * a + m, 3 * a:3 * a + m] += I
                v["V"][b][:m, :m] += I
                v["W"][wk[(a, b)]][:m, :m] -= I
                v["bU"][3 * a:3 * a + m] += e
                v["bV"][3 * b:3 * b + m] -= e
        self.h.chi2 = chi2

    def oplus(self, xu, xv):
        for i in range(self.NU_):
            self.pU[i] = self.pU[i] + xu[3 * i:3 * i + self.su[i]]
        for j in range(self.NV_):
            self.pV[j] = self.pV[j] + xv[3 * j:3 * j + self.sv[j]]

    def backup(self):
        self.bk = ([x.copy() for x in self.pU], [x.copy() for x in self.pV])

    def rollback(self):
        self.pU, self.pV = [x.copy() for x in self.bk[0]], [x.copy() for x in self.bk[1]]

    def dense(self):
        """H and b in the parameters' own sizes (user order)"""
        off = np.cumsum([0] + self.su + self.sv)
        n = off[-1]
        H, b = np.zeros((n, n)), np.zeros(n)
        for kind, a, c, m, e in self._edges():
            ia = off[a] if kind in ("u", "uu", "uv") else off[self.NU_ + a]
            if kind in ("u", "v"):
                H[ia:ia + m, ia:ia + m] += np.eye(m)
                b[ia:ia + m] += e
                continue
            ic = off[c] if kind == "uu" else off[self.NU_ + c]
            J = np.zeros((m, n))
            J[:, ia:ia + m] = np.eye(m)
            J[:, ic:ic + m] = -np.eye(m)
            H += J.T @ J
            b += J.T @ e
        return H, b

    def x(self):
        return np.concatenate([np.concatenate(self.pU), np.concatenate(self.pV)])


CASES = [(6, 5, False, 0), (7, 6, True, 1), (5, 7, True, 2)]


def _run(toy, tmp_path):
    path = tmp_path / "var.txt"
    toy.write(str(path))
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (nelson_amd.build builds it with the library)")
    r = subprocess.run([EXE, str(path)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return {d["solve"]: d for d in (json.loads(line) for line in r.stdout.splitlines() if line.strip())}


@pytest.mark.parametrize("NU,NV,vv,seed", CASES)
def test_variable_block_sizes(NU, NV, vv, seed, tmp_path):
    res = _run(VarToy(NU, NV, vv, seed), tmp_path)
    for name, solve in (("gn", "gn_solve"), ("lm", "lm_solve")):
        t = VarToy(NU, NV, vv, seed)
        ro = getattr(t, solve)()
        d = res[name]
        assert d["final_chi2"] < 1e-12, d  # the reference's REQUIRE(chi2 < dummy_precision)
        for x, y in zip(d["log"], ro["log"]):
            assert x[0] == y[0] and x[1] == y[1], (name, d["log"], ro["log"])
            assert abs(x[2] - y[2]) <= 1e-9 * abs(y[2]) + 1e-20, (name, d["log"], ro["log"])
            assert abs(x[3] - y[3]) <= 1e-9 * abs(y[3]) + 1e-300, (name, d["log"], ro["log"])
            if y[2] < 1e-20:
                break
        assert np.allclose(d["x"], t.x(), rtol=0, atol=1e-12)
    # the U parameters alone as a single section (parameterSize()): LM
    # reaches the ground truth
    d = res["single_lm"]
    assert d["final_chi2"] < 1e-12, d
    assert np.allclose(d["x"], np.concatenate(t.gtU), rtol=0, atol=1e-6)
    # one damped step in the parameters' own sizes against a dense solve
    t = VarToy(NU, NV, vv, seed)
    H, b = t.dense()
    x = np.linalg.solve(H + np.diag(0.05 * np.diag(H) + 0.3), -b)
    s = res["step"]
    assert s["ok"]
    assert s["u_size"] == sum(t.su) and s["v_size"] == sum(t.sv)
    got = np.array(s["dx"])
    assert np.abs(got - x).max() <= 1e-12 * np.abs(x).max()
    assert abs(s["sq_norm"] - x @ x) <= 1e-12 * (x @ x)
