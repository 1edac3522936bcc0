"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_U8 = C.POINTER(C.c_ubyte)


class Hessian(C.Structure):
    _fields_ = [("BU", C.c_int), ("BV", C.c_int), ("NU", C.c_int), ("NV", C.c_int), ("nW", C.c_int),
                ("U", _D), ("V", _D), ("W", _D), ("bU", _D), ("bV", _D), ("Wi", _I), ("Wj", _I),
                ("chi2", C.c_double), ("nVo", C.c_int), ("Vo", _D), ("Voi", _I), ("Voj", _I)]


UPDATE_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int)
OPLUS_CB = C.CFUNCTYPE(None, C.c_void_p, _D, _D)
VOID_CB = C.CFUNCTYPE(None, C.c_void_p)
HESS_CB = C.CFUNCTYPE(C.c_void_p, C.c_void_p)


class Ops(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("update", UPDATE_CB), ("oplus", OPLUS_CB), ("backup", VOID_CB),
                ("rollback", VOID_CB), ("hessian", HESS_CB)]


class LMSettings(C.Structure):
    _fields_ = [("maxNumIt", C.c_int), ("maxNumSubIt", C.c_int), ("minNumIt", C.c_int),
                ("epsBVector", C.c_double), ("epsIncSquare", C.c_double), ("epsChi2", C.c_double),
                ("tau", C.c_double), ("penaltyV", C.c_double), ("initV", C.c_double),
                ("maxScaleFactor", C.c_double), ("minScaleFactor", C.c_double)]


class GNSettings(C.Structure):
    _fields_ = [("maxNumIt", C.c_int), ("minNumIt", C.c_int), ("epsBVector", C.c_double),
                ("epsIncSquare", C.c_double), ("epsChi2", C.c_double), ("absLambda", C.c_double),
                ("relLambda", C.c_double)]


class LogEntry(C.Structure):
    _fields_ = [("it", C.c_int), ("subit", C.c_int), ("chi2", C.c_double), ("mu", C.c_double)]


class LMResult(C.Structure):
    _fields_ = [("reason", C.c_int), ("iterations", C.c_int), ("n_log", C.c_int), ("final_chi2", C.c_double),
                ("log", C.POINTER(LogEntry)), ("log_cap", C.c_int), ("dx_log", _D), ("dx_log_cap", C.c_int),
                ("dx_len", C.c_int)]


REASONS = ["SuccessAtStart", "SuccessSmallIncrement", "SuccessEpsB", "SuccessSmallResiduals",
           "SolverFailure", "MaxIterationReached", "MaxSubIterationReached"]

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "or_hessian_alloc": (C.c_int, [C.POINTER(Hessian), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "or_hessian_free": (None, [C.POINTER(Hessian)]),
            "or_hessian_alloc_voff": (C.c_int, [C.POINTER(Hessian), C.c_int]),
            "or_hessian_max_abs_b": (C.c_double, [C.POINTER(Hessian)]),
            "or_ldlt_factor": (C.c_int, [_D, C.c_int, _I]),
            "or_ldlt_solve": (None, [_D, C.c_int, _I, _D]),
            "or_schur_step": (C.c_int, [C.POINTER(Hessian), C.c_double, C.c_double, _D, _D, _D, _D]),
            "or_schur_assemble": (C.c_int, [C.POINTER(Hessian), C.c_double, C.c_double, _D, _D]),
            "or_max_abs_h_diag": (C.c_double, [C.POINTER(Hessian)]),
            "or_rho": (C.c_double, [C.POINTER(Hessian), C.c_double, _D, _D, C.c_double, C.c_double]),
            "or_diag_inv": (C.c_int, [C.c_int, C.c_int, _D, C.c_double, C.c_double, _D]),
            "or_diag_inv_mult": (None, [C.c_int, C.c_int, _D, _D, _D]),
            "or_wvinv": (None, [C.c_int, C.c_int, C.c_int, _I, _D, _D, _D]),
            "or_wvinv_mult_sub": (None, [C.c_int, C.c_int, C.c_int, _I, _I, _D, _D, _D]),
            "or_wwt": (None, [C.c_int, C.c_int, C.c_int, C.c_int, _I, _I, _D, _D, _D, _D]),
            "or_wtx_sub": (None, [C.c_int, C.c_int, C.c_int, _I, _I, _D, _D, _D]),
            "or_lm_default_settings": (None, [C.POINTER(LMSettings)]),
            "or_gn_default_settings": (None, [C.POINTER(GNSettings)]),
            "or_lm_solve": (C.c_int, [C.POINTER(LMSettings), C.POINTER(Ops), C.POINTER(LMResult)]),
            "or_gn_solve": (C.c_int, [C.POINTER(GNSettings), C.POINTER(Ops), C.POINTER(LMResult)]),
            "or_ba_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, _I, _I, _D, _D, _U8, _U8, C.c_double,
                                          C.c_double, C.c_double, C.c_double, _D, _D]),
            "or_ba_destroy": (None, [C.c_void_p]),
            "or_ba_ops": (Ops, [C.c_void_p]),
            "or_ba_hessian": (C.POINTER(Hessian), [C.c_void_p]),
            "or_ba_update": (None, [C.c_void_p, C.c_int]),
            "or_ba_oplus": (None, [C.c_void_p, _D, _D]),
            "or_ba_get_state": (None, [C.c_void_p, _D, _D]),
            "or_ba_set_state_rt": (None, [C.c_void_p, _D, _D]),
            "or_ba_edge_jacobians": (None, [C.c_void_p, _D, _D, _D]),
            "or_ba_num_var_cams": (C.c_int, [C.c_void_p]),
            "or_ba_num_var_pts": (C.c_int, [C.c_void_p]),
            "or_ba_set_priors": (None, [C.c_void_p, C.c_int, _I, _D, _D, C.c_int, _I, _D, _D]),
            "or_ba_set_intrinsics": (None, [C.c_void_p, _D]),
            "or_so3_log": (None, [_D, _D]),
            "or_so3_jl_inv": (None, [_D, _D]),
            "or_rodrigues": (None, [_D, _D]),
            "or_se3_left_update": (None, [_D, _D, _D]),
            "or_set_threads": (C.c_int, [C.c_int]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def ip(a):
    return None if a is None else a.ctypes.data_as(_I)


def so3_log(R):
    """The oracle's principal log of a 3x3 rotation (angle-axis)."""
    R = np.ascontiguousarray(R, np.float64).reshape(9)
    w = np.zeros(3)
    lib().or_so3_log(dp(R), dp(w))
    return w


def so3_jl_inv(w):
    """The oracle's inverse left Jacobian of SO(3) at w (3x3)."""
    w = np.ascontiguousarray(w, np.float64)
    J = np.zeros(9)
    lib().or_so3_jl_inv(dp(w), dp(J))
    return J.reshape(3, 3)


def hessian_views(h):
    """numpy views of an or_hessian's arrays."""
    BU, BV, NU, NV, nW = h.BU, h.BV, h.NU, h.NV, h.nW
    nu = BU * NU
    d = dict(
        U=np.ctypeslib.as_array(h.U, shape=(nu * nu,)).reshape(nu, nu),
        V=np.ctypeslib.as_array(h.V, shape=(max(NV * BV * BV, 1),))[:NV * BV * BV].reshape(NV, BV, BV),
        W=np.ctypeslib.as_array(h.W, shape=(max(nW * BU * BV, 1),))[:nW * BU * BV].reshape(nW, BU, BV),
        bU=np.ctypeslib.as_array(h.bU, shape=(max(nu, 1),))[:nu],
        bV=np.ctypeslib.as_array(h.bV, shape=(max(NV * BV, 1),))[:NV * BV],
        Wi=np.ctypeslib.as_array(h.Wi, shape=(max(nW, 1),))[:nW],
        Wj=np.ctypeslib.as_array(h.Wj, shape=(max(nW, 1),))[:nW],
    )
    if h.nVo > 0:  # off-diagonal V blocks (V-V edges)
        n = h.nVo
        d["Vo"] = np.ctypeslib.as_array(h.Vo, shape=(n * BV * BV,)).reshape(n, BV, BV)
        d["Voi"] = np.ctypeslib.as_array(h.Voi, shape=(n,))
        d["Voj"] = np.ctypeslib.as_array(h.Voj, shape=(n,))
    return d


def ldlt(a):
    """Pivoted LDL^T of a symmetric matrix (lower triangle read); returns
    (ok, factor (column-major lower, as Fortran array), transpositions)."""
    n = a.shape[0]
    f = np.asfortranarray(np.tril(a).astype(np.float64))
    tp = np.zeros(max(n, 1), np.int32)
    ok = lib().or_ldlt_factor(f.ctypes.data_as(_D), n, ip(tp))
    return bool(ok), f, tp


def ldlt_solve(f, tp, b):
    x = np.array(b, np.float64, copy=True)
    lib().or_ldlt_solve(f.ctypes.data_as(_D), f.shape[0], ip(tp), dp(x))
    return x


def lm_defaults(**kw):
    s = LMSettings()
    lib().or_lm_default_settings(C.byref(s))
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def gn_defaults(**kw):
    s = GNSettings()
    lib().or_gn_default_settings(C.byref(s))
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def _run_driver(fn, settings, ops, log_cap=4096, dx_cap=0):
    log = (LogEntry * log_cap)()
    res = LMResult()
    res.log = log
    res.log_cap = log_cap
    dxbuf = np.zeros(max(dx_cap, 1))
    res.dx_log = dp(dxbuf) if dx_cap else None
    res.dx_log_cap = dx_cap
    fn(C.byref(settings), C.byref(ops), C.byref(res))
    entries = [(log[i].it, log[i].subit, log[i].chi2, log[i].mu) for i in range(min(res.n_log, log_cap))]
    out = dict(reason=REASONS[res.reason], iterations=res.iterations, log=entries, final_chi2=res.final_chi2)
    if dx_cap and res.dx_len:
        nsolve = dx_cap // res.dx_len
        out["dx"] = dxbuf[:nsolve * res.dx_len].reshape(nsolve, res.dx_len)
    return out


class OracleBA:
    """The oracle's BA section built from the same arrays as BAProblem."""

    def __init__(self, n_cameras, n_points, cam_idx, pt_idx, obs, focal, cx, cy, cams0, pts0, weight=None,
                 cam_fixed=None, pt_fixed=None, huber_delta=0.0, intrinsics=None):
        L = lib()
        self.C, self.P = n_cameras, n_points
        self.cam_idx = np.ascontiguousarray(cam_idx, np.int32)
        self.pt_idx = np.ascontiguousarray(pt_idx, np.int32)
        self.obs = np.ascontiguousarray(obs, np.float64)
        self.E = len(self.cam_idx)
        if cam_fixed is None:
            cam_fixed = np.zeros(n_cameras, np.uint8)
            cam_fixed[0] = 1
        if pt_fixed is None:
            pt_fixed = np.zeros(n_points, np.uint8)
            pt_fixed[0] = 1
        self.cam_fixed = np.ascontiguousarray(cam_fixed, np.uint8)
        self.pt_fixed = np.ascontiguousarray(pt_fixed, np.uint8)
        self.weight = None if weight is None else np.ascontiguousarray(weight, np.float64)
        cams0 = np.ascontiguousarray(cams0, np.float64)
        pts0 = np.ascontiguousarray(pts0, np.float64)
        self._h = L.or_ba_create(n_cameras, n_points, self.E, ip(self.cam_idx), ip(self.pt_idx), dp(self.obs),
                                 dp(self.weight), self.cam_fixed.ctypes.data_as(_U8),
                                 self.pt_fixed.ctypes.data_as(_U8), focal, cx, cy, huber_delta, dp(cams0),
                                 dp(pts0))
        self.NU = L.or_ba_num_var_cams(self._h)
        self.NV = L.or_ba_num_var_pts(self._h)
        if intrinsics is not None:  # the BAL camera model (or_ba_set_intrinsics)
            L.or_ba_set_intrinsics(self._h, dp(np.ascontiguousarray(intrinsics, np.float64).ravel()))

    @classmethod
    def from_generated(cls, g, **kw):
        return cls(g["n_cameras"], g["n_points"], g["cam_idx"], g["pt_idx"], g["obs"], g["focal"], g["cx"],
                   g["cy"], g["cams0"], g["pts0"], huber_delta=kw.pop("huber_delta", g.get("huber_delta", 0.0)),
                   intrinsics=kw.pop("intrinsics", g.get("intrinsics")), **kw)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_ba_destroy(self._h)
            self._h = None

    def update(self, hessian=True):
        lib().or_ba_update(self._h, 1 if hessian else 0)
        return self.hessian_struct().chi2

    def hessian_struct(self):
        return lib().or_ba_hessian(self._h).contents

    def hessian(self):
        h = self.hessian_struct()
        v = hessian_views(h)
        out = {k: np.array(a, copy=True) for k, a in v.items()}
        out["chi2"] = h.chi2
        return out

    def schur_step(self, rel, abs_, want_s=False):
        h = lib().or_ba_hessian(self._h)
        nu = 6 * self.NU
        xU = np.zeros(max(nu, 1))
        xV = np.zeros(max(3 * self.NV, 1))
        S = np.zeros((nu, nu)) if want_s else None
        bS = np.zeros(nu) if want_s else None
        ok = lib().or_schur_step(h, rel, abs_, dp(xU), dp(xV), dp(S), dp(bS))
        out = dict(ok=bool(ok), xU=xU[:nu], xV=xV[:3 * self.NV])
        if want_s:
            out["S"] = S
            out["bS"] = bS
        return out

    def schur_assemble(self, rel, abs_):
        """The damped reduced system (S, bS) without factorising it."""
        h = lib().or_ba_hessian(self._h)
        nu = 6 * self.NU
        S = np.zeros((nu, nu))
        bS = np.zeros(max(nu, 1))
        ok = lib().or_schur_assemble(h, rel, abs_, dp(S), dp(bS))
        return dict(ok=bool(ok), S=S, bS=bS[:nu])

    def set_priors(self, cam_ids=(), cam_p0=(), cam_w=(), pt_ids=(), pt_p0=(), pt_w=()):
        """Unary priors (or_ba_set_priors): per camera 6 values (omega0, t0)
        and 6 weights, per point 3 and 3."""
        self._pri = [np.ascontiguousarray(cam_ids, np.int32), np.ascontiguousarray(cam_p0, np.float64).ravel(),
                     np.ascontiguousarray(cam_w, np.float64).ravel(), np.ascontiguousarray(pt_ids, np.int32),
                     np.ascontiguousarray(pt_p0, np.float64).ravel(), np.ascontiguousarray(pt_w, np.float64).ravel()]
        ci, cp, cw, pi, pp, pw = self._pri
        lib().or_ba_set_priors(self._h, len(ci), ip(ci), dp(cp), dp(cw), len(pi), ip(pi), dp(pp), dp(pw))

    def oplus(self, xU, xV):
        lib().or_ba_oplus(self._h, dp(np.ascontiguousarray(xU)), dp(np.ascontiguousarray(xV)))

    def get_state_rt(self):
        cams = np.zeros(12 * self.C)
        pts = np.zeros(3 * self.P)
        lib().or_ba_get_state(self._h, dp(cams), dp(pts))
        return cams, pts

    def set_state_rt(self, cams, pts):
        lib().or_ba_set_state_rt(self._h, dp(np.ascontiguousarray(cams)), dp(np.ascontiguousarray(pts)))

    def edge_jacobians(self):
        r = np.zeros(2 * self.E)
        Jc = np.zeros(12 * self.E)
        Jp = np.zeros(6 * self.E)
        lib().or_ba_edge_jacobians(self._h, dp(r), dp(Jc), dp(Jp))
        return r.reshape(self.E, 2), Jc.reshape(self.E, 2, 6), Jp.reshape(self.E, 2, 3)

    def lm_solve(self, settings=None, dx_cap=0, **kw):
        s = settings or lm_defaults(**kw)
        ops = lib().or_ba_ops(self._h)
        return _run_driver(lib().or_lm_solve, s, ops, dx_cap=dx_cap)

    def gn_solve(self, settings=None, **kw):
        s = settings or gn_defaults(**kw)
        ops = lib().or_ba_ops(self._h)
        return _run_driver(lib().or_gn_solve, s, ops)


class PyProblem:
    """A problem written in Python (callbacks) over an oracle-allocated
    generic block Hessian; used for the linear toy problem of
    nelson/tests/testDoubleSectionCommon.h."""

    def __init__(self, BU, BV, NU, NV, Wi, Wj, Vpairs=()):
        self.h = Hessian()
        lib().or_hessian_alloc(C.byref(self.h), BU, BV, NU, NV, len(Wi))
        if len(Vpairs):  # off-diagonal V blocks (i, j), i < j, of V-V edges
            lib().or_hessian_alloc_voff(C.byref(self.h), len(Vpairs))
        v = hessian_views(self.h)
        v["Wi"][:] = Wi
        v["Wj"][:] = Wj
        if len(Vpairs):
            v["Voi"][:] = [p[0] for p in Vpairs]
            v["Voj"][:] = [p[1] for p in Vpairs]
        self._cbs = None

    def __del__(self):
        try:
            lib().or_hessian_free(C.byref(self.h))
        except Exception:
            pass

    def views(self):
        return hessian_views(self.h)

    # subclasses implement: update(hessian), oplus(xU, xV), backup(), rollback()
    def _ops(self):
        nu = self.h.BU * self.h.NU
        nv = self.h.BV * self.h.NV

        def upd(_, hess):
            self.update(bool(hess))

        def opl(_, xu, xv):
            self.oplus(np.ctypeslib.as_array(xu, shape=(nu,)).copy(), np.ctypeslib.as_array(xv, shape=(nv,)).copy())

        def bak(_):
            self.backup()

        def rol(_):
            self.rollback()

        def hes(_):
            return C.addressof(self.h)

        self._cbs = (UPDATE_CB(upd), OPLUS_CB(opl), VOID_CB(bak), VOID_CB(rol), HESS_CB(hes))
        o = Ops()
        o.ctx = None
        o.update, o.oplus, o.backup, o.rollback, o.hessian = self._cbs
        return o

    def lm_solve(self, **kw):
        return _run_driver(lib().or_lm_solve, lm_defaults(**kw), self._ops())

    def gn_solve(self, **kw):
        return _run_driver(lib().or_gn_solve, gn_defaults(**kw), self._ops())
