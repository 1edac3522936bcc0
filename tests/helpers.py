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


__all__ = ["relerr", "small_cfg", "pair", "oracle_blocks", "dense_system", "generate_ba", "CONFIGS"]
