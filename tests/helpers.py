"""Shared helpers of the parity tests."""
import numpy as np

from nelson_amd import BAProblem, CONFIGS, generate_ba
from oracle.oracle import OracleBA


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(float(np.abs(b).max()) if b.size else 0.0, 1e-300)
    return float(np.abs(a - b).max()) / scale if a.size else 0.0


def small_cfg(**kw):
    c = dict(CONFIGS["cfg1"])
    c.update(kw)
    return c


def pair(g, **kw):
    """The GPU problem and the oracle problem on the same arrays and state."""
    p = BAProblem.from_generated(g, **kw)
    okw = {k: v for k, v in kw.items() if k in ("weight", "cam_fixed", "pt_fixed", "huber_delta")}
    o = OracleBA.from_generated(g, **okw)
    return p, o


def oracle_blocks(o):
    """Oracle Hessian in the export layout of the C ABI."""
    H = o.hessian()
    NU = o.NU
    Uf = np.triu(H["U"]) + np.triu(H["U"], 1).T
    U = np.stack([Uf[6 * u:6 * u + 6, 6 * u:6 * u + 6] for u in range(NU)]) if NU else np.zeros((0, 6, 6))
    return dict(U=U, V=H["V"], W=H["W"], Wi=H["Wi"], Wj=H["Wj"], bU=H["bU"], bV=H["bV"], chi2=H["chi2"])


def dense_system(H, NU, NV):
    """Dense H (full symmetric) and b from block arrays in export layout."""
    nu, nv = 6 * NU, 3 * NV
    Hd = np.zeros((nu + nv, nu + nv))
    for u in range(NU):
        Hd[6 * u:6 * u + 6, 6 * u:6 * u + 6] = H["U"][u]
    for q in range(NV):
        Hd[nu + 3 * q:nu + 3 * q + 3, nu + 3 * q:nu + 3 * q + 3] = H["V"][q]
    for k in range(len(H["Wi"])):
        i, j = H["Wi"][k], H["Wj"][k]
        Hd[6 * i:6 * i + 6, nu + 3 * j:nu + 3 * j + 3] = H["W"][k]
        Hd[nu + 3 * j:nu + 3 * j + 3, 6 * i:6 * i + 6] = H["W"][k].T
    b = np.concatenate([H["bU"], H["bV"]])
    return Hd, b


def bal_project(cams, intr, pts, cam_idx, pt_idx):
    """The BAL camera model in numpy (independent of the library and the
    oracle): Xc = R X + t (scipy's Rodrigues), q = -Xc_xy / Xc_z,
    pixel = f (1 + k1 |q|^2 + k2 |q|^4) q."""
    from scipy.spatial.transform import Rotation
    c6 = np.asarray(cams, np.float64).reshape(-1, 6)
    K = np.asarray(intr, np.float64).reshape(-1, 3)
    R = Rotation.from_rotvec(c6[cam_idx, :3]).as_matrix()
    Xc = np.einsum("eij,ej->ei", R, np.asarray(pts).reshape(-1, 3)[pt_idx]) + c6[cam_idx, 3:]
    q = -Xc[:, :2] / Xc[:, 2:3]
    s = np.sum(q * q, axis=1)
    k = K[cam_idx]
    return k[:, 0:1] * (1.0 + k[:, 1:2] * s[:, None] + k[:, 2:3] * (s * s)[:, None]) * q


def bal_problem(cfg, seed=3, sigma=0.5):
    """A problem in the BAL camera model from the seeded generator: its
    cameras turned by pi about x (so they look down -z), per-camera
    f in U(400, 600), k1 in U(-0.3, 0.3), k2 in U(-0.1, 0.1), observations
    re-projected with bal_project plus N(0, sigma) noise."""
    from scipy.spatial.transform import Rotation
    g = generate_ba(cfg)
    rng = np.random.default_rng(seed)
    F = np.diag([1.0, -1.0, -1.0])
    out = dict(g)
    for key in ("cams_gt", "cams0"):
        c6 = g[key].reshape(-1, 6).copy()
        c6[:, :3] = Rotation.from_matrix(F @ Rotation.from_rotvec(c6[:, :3]).as_matrix()).as_rotvec()
        c6[:, 3:] = c6[:, 3:] @ F.T
        out[key] = c6.reshape(-1)
    C = g["n_cameras"]
    out["intrinsics"] = np.stack([rng.uniform(400, 600, C), rng.uniform(-0.3, 0.3, C), rng.uniform(-0.1, 0.1, C)],
                                 1).reshape(-1)
    uv = bal_project(out["cams_gt"], out["intrinsics"], g["pts_gt"], g["cam_idx"], g["pt_idx"])
    out["obs"] = (uv + sigma * rng.standard_normal(uv.shape)).reshape(-1)
    out["focal"], out["cx"], out["cy"] = 0.0, 0.0, 0.0
    return out


def write_bal(path, g):
    """Writes g (the dict of bal_problem) as a BAL problem file."""
    with open(path, "w") as f:
        f.write(f"{g['n_cameras']} {g['n_points']} {g['n_edges']}\n")
        obs = g["obs"].reshape(-1, 2)
        for e in range(g["n_edges"]):
            f.write(f"{g['cam_idx'][e]} {g['pt_idx'][e]} {obs[e, 0]:.17g} {obs[e, 1]:.17g}\n")
        c6, K = g["cams0"].reshape(-1, 6), g["intrinsics"].reshape(-1, 3)
        for c in range(g["n_cameras"]):
            for v in list(c6[c]) + list(K[c]):
                f.write(f"{v:.17g}\n")
        for v in g["pts0"]:
            f.write(f"{v:.17g}\n")


__all__ = ["relerr", "small_cfg", "pair", "oracle_blocks", "dense_system", "generate_ba", "CONFIGS", "bal_project",
           "bal_problem", "write_bal"]
