"""The 2-D scan alignment problem of the reference's single-section test
(nelson/tests/testGaussNewton2DScansAlign.cpp), shared by the GPU test of
the C++ single section and the oracle's copy: SE(2) poses (theta, tx, ty),
one fixed; point-to-line edges between scans; the scans are exact views of
one set of world points, so the optimum has chi2 = 0."""
import numpy as np

from oracle.oracle import PyProblem


def se2_exp_left(pose, d):
    """pose <- exp(delta) pose, delta = (rho_x, rho_y, phi)"""
    th, tx, ty = pose
    rx, ry, ph = d
    if abs(ph) < 1e-8:
        a, b = 1.0 - ph * ph / 6.0, ph / 2.0
    else:
        a, b = np.sin(ph) / ph, (1.0 - np.cos(ph)) / ph
    c, s = np.cos(ph), np.sin(ph)
    return np.array([th + ph, c * tx - s * ty + a * rx - b * ry, s * tx + c * ty + b * rx + a * ry])


def make(n_poses=15, n_points=300, full=True, seed=0, noise=(0.01, 0.01, np.pi / 1000)):
    """world points 100 U(-1, 1)^2 with random unit normals; n_poses + 1 scan
    poses exp(diag(10, 10, pi) U(-1, 1)^3); scans and normals in each scan's
    frame; the start: pose 0 exact (fixed), the others perturbed by
    exp(noise .* U(-1, 1)^3); edges from scan 0 to every scan and, with
    `full`, between every pair (i < j), else a few random extra pairs"""
    rng = np.random.default_rng(seed)
    W = 100.0 * rng.uniform(-1, 1, (n_points, 2))
    N = rng.uniform(-1, 1, (n_points, 2))
    N /= np.linalg.norm(N, axis=1)[:, None]
    NP = n_poses + 1
    poses = [se2_exp_left(np.zeros(3), np.array([10.0, 10.0, np.pi]) * rng.uniform(-1, 1, 3)) for _ in range(NP)]
    scans = []
    for th, tx, ty in poses:
        R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        loc = (W - np.array([tx, ty])) @ R        # R^T (w - t)
        nl = N @ R
        scans.append(np.concatenate([loc, nl], axis=1))
    start = [poses[0]] + [se2_exp_left(p, np.array(noise) * rng.uniform(-1, 1, 3)) for p in poses[1:]]
    if full:
        edges = [(i, j) for i in range(NP) for j in range(i + 1, NP)]
    else:
        edges = [(0, j) for j in range(1, NP)]
        for i in range(1, NP - 1):
            for _ in range(int(rng.integers(0, NP - i - 1))):
                j = int(rng.integers(i + 1, NP))
                edges.append((i, j))
    return dict(NP=NP, NPTS=n_points, start=np.array(start), scans=scans, edges=edges)


def write(d, path):
    with open(path, "w") as f:
        f.write(f"{d['NP']} {d['NPTS']} {len(d['edges'])}\n")
        for p in d["start"]:
            f.write(" ".join(f"{v:.17g}" for v in p) + "\n")
        for s in d["scans"]:
            for row in s:
                f.write(" ".join(f"{v:.17g}" for v in row) + "\n")
        for i, j in d["edges"]:
            f.write(f"{i} {j}\n")


class ScanProblem(PyProblem):
    """The oracle's copy: one section of NP - 1 variable SE(2) poses (H = U
    dense, no V), binary edges in insertion order, updaters in reverse."""

    def __init__(self, d):
        self.d = d
        self.NU = d["NP"] - 1
        super().__init__(3, 3, self.NU, 0, [], [])
        self.poses = d["start"][1:].copy()
        self.fixed = d["start"][0].copy()

    def pose(self, i):
        return self.fixed if i == 0 else self.poses[i - 1]

    def update(self, hess):
        v = self.views()
        if hess:
            v["U"][...] = 0
            v["bU"][...] = 0
        chi2 = 0.0
        blocks = []
        for i, j in self.d["edges"]:
            m, t = self.pose(i), self.pose(j)
            sm, st = self.d["scans"][i], self.d["scans"][j]
            cm, snm, ct, snt = np.cos(m[0]), np.sin(m[0]), np.cos(t[0]), np.sin(t[0])
            mx = cm * sm[:, 0] - snm * sm[:, 1] + m[1]
            my = snm * sm[:, 0] + cm * sm[:, 1] + m[2]
            nx = cm * sm[:, 2] - snm * sm[:, 3]
            ny = snm * sm[:, 2] + cm * sm[:, 3]
            px = ct * st[:, 0] - snt * st[:, 1] + t[1]
            py = snt * st[:, 0] + ct * st[:, 1] + t[2]
            e = (mx - px) * nx + (my - py) * ny
            chi2 += float(e @ e)
            if hess:
                C = px * ny - py * nx
                J = np.stack([nx, ny, C], 1)
                blocks.append((i, j, J.T @ J, J.T @ e))
        self.h.chi2 = chi2
        if not hess:
            return
        U, bU = v["U"], v["bU"]
        for i, j, Hs, bs in reversed(blocks):
            if i > 0:
                a = 3 * (i - 1)
                U[a:a + 3, a:a + 3] += Hs
                bU[a:a + 3] += bs
            if j > 0:
                c = 3 * (j - 1)
                U[c:c + 3, c:c + 3] += Hs
                bU[c:c + 3] -= bs
            if i > 0 and j > 0:
                U[a:a + 3, c:c + 3] -= Hs

    def oplus(self, xu, xv):
        for k in range(self.NU):
            self.poses[k] = se2_exp_left(self.poses[k], xu[3 * k:3 * k + 3])

    def backup(self):
        self.bk = self.poses.copy()

    def rollback(self):
        self.poses = self.bk.copy()
