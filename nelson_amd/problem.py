"""Python mirror of nelson's problem / solver / driver interface over the C ABI.

Names follow the reference: BAProblem plays a DoubleSection subclass with BA
edges (nelson/DoubleSection.h:181-545), SolverCholeskySchur the Schur solver
(nelson/SolverCholeskySchur.h:53-117), LevenbergMarquardt / GaussNewton the
drivers (nelson/LevenbergMarquardt.h, nelson/GaussNewton.h).
"""
import ctypes as C

import numpy as np

from . import capi
from .capi import check

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_I) if a is not None else None


# ---------------------------------------------------------------------------
# synthetic workloads
# ---------------------------------------------------------------------------
_BASE = dict(obs_sigma=1.0, outlier_frac=0.0, outlier_sigma=0.0, rot_sigma=0.01, trans_sigma=0.05,
             point_sigma=0.1, focal=500.0, cx=320.0, cy=240.0)

CONFIGS = {
    # 10 cameras, 500 points, 4 observations per point, 20 LM iterations
    "cfg1": dict(_BASE, n_cameras=10, n_points=500, k_min=4, k_max=4, layout=0, arc=0, seed=1, iterations=20),
    # 1000 cameras on a ring, 200k points seen by 4-8 cameras of a 150-camera arc, 30 LM iterations
    "cfg2": dict(_BASE, n_cameras=1000, n_points=200000, k_min=4, k_max=8, layout=1, arc=150, seed=2,
                 iterations=30),
    # 4000 cameras along an open path (a ring whose camera windows do not
    # wrap), 400k points seen by 4-8 cameras of a 40-camera window: a banded S
    # (n = 23 994) for the sparse-S path
    "seq": dict(_BASE, n_cameras=4000, n_points=400000, k_min=4, k_max=8, layout=2, arc=40, seed=7,
                iterations=20),
    # 200 cameras, 20k points, 6 observations per point (differentiable BA)
    "cfg5": dict(_BASE, n_cameras=200, n_points=20000, k_min=6, k_max=6, layout=1, arc=40, seed=5,
                 iterations=50),
}


class _SplitMix64:
    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform_int(self, lo, hi):
        u = (self.next() >> 11) * 2.0 ** -53
        return min(hi, lo + int(u * (hi - lo + 1)))


def window_config(w):
    """Batched local-BA window w: cameras U{5..40}, points U{200..3000}, 3-6
    observations per point, 5% outliers at 20 px, Huber delta 2 px."""
    r = _SplitMix64(4000 + w)
    nc = r.uniform_int(5, 40)
    npt = r.uniform_int(200, 3000)
    return dict(_BASE, n_cameras=nc, n_points=npt, k_min=3, k_max=6, layout=0, arc=0, seed=4000 + w,
                outlier_frac=0.05, outlier_sigma=20.0, huber_delta=2.0, iterations=10)


def generate_ba(cfg):
    """Runs the library's seeded generator; returns a dict of numpy arrays."""
    lib = capi.lib()
    p = capi.GenParams()
    for name, _ in capi.GenParams._fields_:
        setattr(p, name, cfg[name])
    nc, npt = cfg["n_cameras"], cfg["n_points"]
    cap = npt * cfg["k_max"]
    out = dict(cams_gt=np.zeros(6 * nc), pts_gt=np.zeros(3 * npt), cams0=np.zeros(6 * nc),
               pts0=np.zeros(3 * npt), cam_idx=np.zeros(cap, np.int32), pt_idx=np.zeros(cap, np.int32),
               obs=np.zeros(2 * cap))
    ne = C.c_int(0)
    check(lib.nsn_generate_ba(C.byref(p), cap, _dp(out["cams_gt"]), _dp(out["pts_gt"]), _dp(out["cams0"]),
                              _dp(out["pts0"]), _ip(out["cam_idx"]), _ip(out["pt_idx"]), _dp(out["obs"]),
                              C.byref(ne)))
    e = ne.value
    out["cam_idx"] = out["cam_idx"][:e].copy()
    out["pt_idx"] = out["pt_idx"][:e].copy()
    out["obs"] = out["obs"][:2 * e].copy()
    out["n_cameras"], out["n_points"], out["n_edges"] = nc, npt, e
    for k in ("focal", "cx", "cy"):
        out[k] = cfg[k]
    out["huber_delta"] = cfg.get("huber_delta", 0.0)
    return out


def load_visibility(path):
    """Problem::load (nelson/samples/Problems.cpp:8-30): a visibility-graph
    text file -> dict(n_cameras, n_points, cam_idx, pt_idx)."""
    lib = capi.lib()
    nc, npt, ne = C.c_int(0), C.c_int(0), C.c_int(0)
    bpath = str(path).encode()
    check(lib.nsn_visibility_load(bpath, 0, C.byref(nc), C.byref(npt), C.byref(ne), None, None))
    ci = np.zeros(max(ne.value, 1), np.int32)
    pi = np.zeros(max(ne.value, 1), np.int32)
    check(lib.nsn_visibility_load(bpath, ne.value, C.byref(nc), C.byref(npt), C.byref(ne), _ip(ci), _ip(pi)))
    return dict(n_cameras=nc.value, n_points=npt.value, cam_idx=ci[:ne.value], pt_idx=pi[:ne.value])


def load_bal(path):
    """nsn_bal_load: a Bundle Adjustment in the Large problem file -> the dict
    of generate_ba (cams0 / pts0 the file's state, intrinsics (f, k1, k2) per
    camera for the BAL camera model), ready for BAProblem.from_generated."""
    lib = capi.lib()
    nc, npt, ne = C.c_int(0), C.c_int(0), C.c_int(0)
    bpath = str(path).encode()
    z = [None] * 6
    check(lib.nsn_bal_load(bpath, 0, C.byref(nc), C.byref(npt), C.byref(ne), *z))
    e = ne.value
    ci, pi = np.zeros(max(e, 1), np.int32), np.zeros(max(e, 1), np.int32)
    obs = np.zeros(max(2 * e, 1))
    cams, intr, pts = np.zeros(6 * nc.value), np.zeros(3 * nc.value), np.zeros(3 * npt.value)
    check(lib.nsn_bal_load(bpath, e, C.byref(nc), C.byref(npt), C.byref(ne), _ip(ci), _ip(pi), _dp(obs), _dp(cams),
                           _dp(intr), _dp(pts)))
    return dict(n_cameras=nc.value, n_points=npt.value, n_edges=e, cam_idx=ci[:e], pt_idx=pi[:e], obs=obs[:2 * e],
                cams0=cams, pts0=pts, intrinsics=intr, focal=0.0, cx=0.0, cy=0.0, huber_delta=0.0)


# the reference's BA sample on a visibility graph
# (nelson/samples/BASchurPerfomances.cpp:40-75: camera 100 x 100 px, focal 100,
# points in the unit cube, cameras at focal * 1 / (100 / 2) = 2 from the origin)
LADYBUG_SAMPLE = dict(focal=100.0, cx=50.0, cy=50.0, point_radius=1.0, cam_radius=2.0, obs_sigma=0.5,
                      rot_sigma=0.002, trans_sigma=0.002, point_sigma=0.002, seed=49)


def generate_on_graph(graph, **params):
    """nsn_generate_on_graph on a visibility graph (load_visibility's dict);
    params default to LADYBUG_SAMPLE.  Returns the dict of generate_ba."""
    lib = capi.lib()
    cfg = dict(LADYBUG_SAMPLE, **params)
    p = capi.GraphGenParams()
    for name, _ in capi.GraphGenParams._fields_:
        setattr(p, name, cfg[name])
    nc, npt = int(graph["n_cameras"]), int(graph["n_points"])
    ci = np.ascontiguousarray(graph["cam_idx"], np.int32)
    pi = np.ascontiguousarray(graph["pt_idx"], np.int32)
    e = len(ci)
    out = dict(cams_gt=np.zeros(6 * nc), pts_gt=np.zeros(3 * npt), cams0=np.zeros(6 * nc), pts0=np.zeros(3 * npt),
               obs=np.zeros(max(2 * e, 1)))
    check(lib.nsn_generate_on_graph(C.byref(p), nc, npt, e, _ip(ci), _ip(pi), _dp(out["cams_gt"]),
                                    _dp(out["pts_gt"]), _dp(out["cams0"]), _dp(out["pts0"]), _dp(out["obs"])))
    out["obs"] = out["obs"][:2 * e]
    out.update(cam_idx=ci.copy(), pt_idx=pi.copy(), n_cameras=nc, n_points=npt, n_edges=e, focal=cfg["focal"],
               cx=cfg["cx"], cy=cfg["cy"], huber_delta=0.0)
    return out


# ---------------------------------------------------------------------------
# problem (DoubleSection + BA edges)
# ---------------------------------------------------------------------------
class BAProblem:
    """A BA problem on one GPU: cameras (angle-axis, translation) and points.

    Camera 0 and point 0 are fixed by default, the gauge of the reference's
    BA sample (nelson/samples/BASchurPerfomances.cpp:245-246).
    """

    def __init__(self, n_cameras, n_points, cam_idx, pt_idx, obs, focal, cx, cy, weight=None, cam_fixed=None,
                 pt_fixed=None, huber_delta=0.0, precision="fp64", device=0, intrinsics=None, camera_order="given"):
        self._lib = capi.lib()
        self.n_cameras, self.n_points = int(n_cameras), int(n_points)
        self.cam_idx = np.ascontiguousarray(cam_idx, np.int32)
        self.pt_idx = np.ascontiguousarray(pt_idx, np.int32)
        self.obs = np.ascontiguousarray(obs, np.float64).reshape(-1)
        self.n_edges = len(self.cam_idx)
        if cam_fixed is None:
            cam_fixed = np.zeros(self.n_cameras, np.uint8)
            cam_fixed[0] = 1
        if pt_fixed is None:
            pt_fixed = np.zeros(self.n_points, np.uint8)
            pt_fixed[0] = 1
        self.cam_fixed = np.ascontiguousarray(cam_fixed, np.uint8)
        self.pt_fixed = np.ascontiguousarray(pt_fixed, np.uint8)
        self.weight = None if weight is None else np.ascontiguousarray(weight, np.float64)
        # per-camera (f, k1, k2) of the BAL camera model, or None: the shared pinhole
        self.intrinsics = None if intrinsics is None else np.ascontiguousarray(intrinsics, np.float64).reshape(-1)
        if self.intrinsics is not None and self.intrinsics.size != 3 * self.n_cameras:
            raise ValueError("intrinsics must hold (f, k1, k2) per camera")
        d = capi.ProblemDesc()
        d.n_cameras, d.n_points, d.n_edges = self.n_cameras, self.n_points, self.n_edges
        d.cam_idx, d.pt_idx, d.obs = _ip(self.cam_idx), _ip(self.pt_idx), _dp(self.obs)
        d.weight = _dp(self.weight)
        d.cam_fixed = self.cam_fixed.ctypes.data_as(C.POINTER(C.c_ubyte))
        d.pt_fixed = self.pt_fixed.ctypes.data_as(C.POINTER(C.c_ubyte))
        d.focal, d.cx, d.cy, d.huber_delta = focal, cx, cy, huber_delta
        # "rcm": reverse Cuthill-McKee order of the cameras inside the solver
        # (a narrow envelope of S); "nd": nested dissection (independent
        # diagonal chains of the factorisation); arrays per camera at the
        # interface keep camera order either way
        orders = {"given": capi.NSN_ORDER_AS_GIVEN, "rcm": capi.NSN_ORDER_RCM, "nd": capi.NSN_ORDER_ND}
        if camera_order not in orders:
            raise ValueError("camera_order must be 'given', 'rcm' or 'nd'")
        d.camera_order = orders[camera_order]
        d.precision = capi.NSN_MIXED if precision == "mixed" else capi.NSN_FP64
        d.device = device
        d.intrinsics = _dp(self.intrinsics)
        h = C.c_void_p()
        check(self._lib.nsn_problem_create(C.byref(d), C.byref(h)))
        self._h = h
        info = capi.ProblemInfo()
        check(self._lib.nsn_problem_info_get(self._h, C.byref(info)))
        self.NU, self.NV, self.nW = info.n_var_cameras, info.n_var_points, info.n_wblocks
        self.n_s, self.n_s_blocks, self.n_s_pairs = info.n_s, info.n_s_blocks, info.n_s_pairs
        self.n_tiles, self.n_tiles_dense = info.n_tiles, info.n_tiles_dense
        self.factor_flops = info.factor_flops
        self.n_chains, self.chain_steps, self.lookahead = info.n_chains, info.chain_steps, info.lookahead
        self._chi2 = None

    @staticmethod
    def plan(g, camera_order="given", cam_fixed=None, pt_fixed=None):
        """The structure stage of a generated problem on the host alone
        (nsn_problem_plan, no device): the problem figures (tile envelope,
        diagonal chains, critical path in steps, operations) and the order
        of the variable cameras inside the solver."""
        lib = capi.lib()
        C_, P_ = int(g["n_cameras"]), int(g["n_points"])
        cam_idx = np.ascontiguousarray(g["cam_idx"], np.int32)
        pt_idx = np.ascontiguousarray(g["pt_idx"], np.int32)
        obs = np.ascontiguousarray(g["obs"], np.float64).reshape(-1)
        if cam_fixed is None:
            cam_fixed = np.zeros(C_, np.uint8)
            cam_fixed[0] = 1
        if pt_fixed is None:
            pt_fixed = np.zeros(P_, np.uint8)
            pt_fixed[0] = 1
        cam_fixed = np.ascontiguousarray(cam_fixed, np.uint8)
        pt_fixed = np.ascontiguousarray(pt_fixed, np.uint8)
        orders = {"given": capi.NSN_ORDER_AS_GIVEN, "rcm": capi.NSN_ORDER_RCM, "nd": capi.NSN_ORDER_ND}
        d = capi.ProblemDesc()
        d.n_cameras, d.n_points, d.n_edges = C_, P_, len(cam_idx)
        d.cam_idx, d.pt_idx, d.obs = _ip(cam_idx), _ip(pt_idx), _dp(obs)
        d.weight = None
        d.cam_fixed = cam_fixed.ctypes.data_as(C.POINTER(C.c_ubyte))
        d.pt_fixed = pt_fixed.ctypes.data_as(C.POINTER(C.c_ubyte))
        d.focal, d.cx, d.cy = g["focal"], g["cx"], g["cy"]
        d.camera_order = orders[camera_order]
        info = capi.ProblemInfo()
        order = np.zeros(int((cam_fixed == 0).sum()), np.int32)
        check(lib.nsn_problem_plan(C.byref(d), C.byref(info), _ip(order)))
        return {k: getattr(info, k) for k, _ in capi.ProblemInfo._fields_}, order

    @staticmethod
    def structure_check(g, camera_order="given", cam_fixed=None, pt_fixed=None, weight=None, device=0):
        """The structure stage of a generated problem built on the host and on
        the GPU (nsn_structure_check): (number of differing fields, host ms,
        device ms)."""
        lib = capi.lib()
        C_, P_ = int(g["n_cameras"]), int(g["n_points"])
        cam_idx = np.ascontiguousarray(g["cam_idx"], np.int32)
        pt_idx = np.ascontiguousarray(g["pt_idx"], np.int32)
        obs = np.ascontiguousarray(g["obs"], np.float64).reshape(-1)
        if cam_fixed is None:
            cam_fixed = np.zeros(C_, np.uint8)
            cam_fixed[0] = 1
        if pt_fixed is None:
            pt_fixed = np.zeros(P_, np.uint8)
            pt_fixed[0] = 1
        cam_fixed = np.ascontiguousarray(cam_fixed, np.uint8)
        pt_fixed = np.ascontiguousarray(pt_fixed, np.uint8)
        weight = None if weight is None else np.ascontiguousarray(weight, np.float64)
        orders = {"given": capi.NSN_ORDER_AS_GIVEN, "rcm": capi.NSN_ORDER_RCM, "nd": capi.NSN_ORDER_ND}
        d = capi.ProblemDesc()
        d.n_cameras, d.n_points, d.n_edges = C_, P_, len(cam_idx)
        d.cam_idx, d.pt_idx, d.obs = _ip(cam_idx), _ip(pt_idx), _dp(obs)
        d.weight = _dp(weight)
        d.cam_fixed = cam_fixed.ctypes.data_as(C.POINTER(C.c_ubyte))
        d.pt_fixed = pt_fixed.ctypes.data_as(C.POINTER(C.c_ubyte))
        d.focal, d.cx, d.cy = g["focal"], g["cx"], g["cy"]
        d.camera_order = orders[camera_order]
        d.device = device
        nd = C.c_longlong(-1)
        th, td = C.c_double(0), C.c_double(0)
        check(lib.nsn_structure_check(C.byref(d), C.byref(nd), C.byref(th), C.byref(td)))
        return nd.value, th.value, td.value

    @classmethod
    def from_generated(cls, g, **kw):
        p = cls(g["n_cameras"], g["n_points"], g["cam_idx"], g["pt_idx"], g["obs"], g["focal"], g["cx"],
                g["cy"], huber_delta=kw.pop("huber_delta", g.get("huber_delta", 0.0)),
                intrinsics=kw.pop("intrinsics", g.get("intrinsics")), **kw)
        p.set_state(g["cams0"], g["pts0"])
        return p

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nsn_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_refinement(self, steps, tol=None):
        """Iterative refinement of every mixed-precision increment (ignored
        in FP64): exactly `steps` steps (0..8), or with `tol` at most `steps`
        steps, stopping once the FP64 residual's largest entry is at most
        tol times the gradient's.  Default: at most 2 steps, tol 1e-10."""
        if tol is None:
            check(self._lib.nsn_set_refinement(self._h, int(steps)))
        else:
            check(self._lib.nsn_set_refinement_tol(self._h, int(steps), float(tol)))

    def refinement_count(self):
        """Refinement solves run since the problem was created."""
        n = C.c_int(0)
        check(self._lib.nsn_refinement_count(self._h, C.byref(n)))
        return n.value

    # parameters ------------------------------------------------------------
    def set_state(self, cams, pts):
        cams = np.ascontiguousarray(cams, np.float64)
        pts = np.ascontiguousarray(pts, np.float64)
        check(self._lib.nsn_set_state(self._h, _dp(cams), _dp(pts)))

    def set_state_rt(self, cams_rt, pts):
        cams_rt = np.ascontiguousarray(cams_rt, np.float64)
        pts = np.ascontiguousarray(pts, np.float64)
        check(self._lib.nsn_set_state_rt(self._h, _dp(cams_rt), _dp(pts)))

    def get_state(self):
        cams = np.zeros(6 * self.n_cameras)
        pts = np.zeros(3 * self.n_points)
        check(self._lib.nsn_get_state(self._h, _dp(cams), _dp(pts)))
        return cams, pts

    def get_state_rt(self):
        cams = np.zeros(12 * self.n_cameras)
        pts = np.zeros(3 * self.n_points)
        check(self._lib.nsn_get_state_rt(self._h, _dp(cams), _dp(pts)))
        return cams, pts

    def set_priors(self, cam_ids=(), cam_p0=(), cam_w=(), pt_ids=(), pt_p0=(), pt_w=()):
        """Unary priors (nsn_set_priors): per camera the angle-axis and
        translation it is pulled to (6 values) and 6 weights, per point 3 and
        3; replaces earlier priors, used from the next linearisation."""
        ci = np.ascontiguousarray(cam_ids, np.int32).reshape(-1)
        cp = np.ascontiguousarray(cam_p0, np.float64).reshape(-1)
        cw = np.ascontiguousarray(cam_w, np.float64).reshape(-1)
        pi = np.ascontiguousarray(pt_ids, np.int32).reshape(-1)
        pp = np.ascontiguousarray(pt_p0, np.float64).reshape(-1)
        pw = np.ascontiguousarray(pt_w, np.float64).reshape(-1)
        if cp.size != 6 * ci.size or cw.size != 6 * ci.size or pp.size != 3 * pi.size or pw.size != 3 * pi.size:
            raise ValueError("priors: 6 values and 6 weights per camera, 3 and 3 per point")
        check(self._lib.nsn_set_priors(self._h, ci.size, _ip(ci), _dp(cp), _dp(cw), pi.size, _ip(pi), _dp(pp),
                                       _dp(pw)))

    def set_observations(self, obs=None, weight=None):
        """Replaces the observations [E, 2] and / or per-edge weights [E]
        (user edge order) for the next linearisation."""
        o = None if obs is None else np.ascontiguousarray(obs, np.float64).reshape(-1)
        w = None if weight is None else np.ascontiguousarray(weight, np.float64).reshape(-1)
        if o is not None and o.size != 2 * self.n_edges:
            raise ValueError("obs must hold 2 values per edge")
        if w is not None and w.size != self.n_edges:
            raise ValueError("weight must hold one value per edge")
        check(self._lib.nsn_set_observations(self._h, _dp(o), _dp(w)))
        if o is not None:
            self.obs = o.copy()
        if w is not None:
            self.weight = w.copy()

    def adjoint(self, g_cams=None, g_pts=None, hessian_steps=2):
        """Implicit-function-theorem adjoint at the current (converged) state.

        g_cams [C, 6]: dL/d(camera) in the left-perturbation coordinates
        (omega, upsilon) of oplus; g_pts [P, 3]: dL/d(point).  Returns
        (dL/d(obs) [E, 2], dL/d(weight) [E]).  hessian_steps refines the
        multipliers against the exact Hessian (0: Gauss-Newton Hessian)."""
        gc = None if g_cams is None else np.ascontiguousarray(g_cams, np.float64).reshape(-1)
        gp = None if g_pts is None else np.ascontiguousarray(g_pts, np.float64).reshape(-1)
        if gc is not None and gc.size != 6 * self.n_cameras:
            raise ValueError("g_cams must hold 6 values per camera")
        if gp is not None and gp.size != 3 * self.n_points:
            raise ValueError("g_pts must hold 3 values per point")
        dobs = np.zeros(2 * self.n_edges)
        dw = np.zeros(self.n_edges)
        check(self._lib.nsn_adjoint(self._h, _dp(gc), _dp(gp), int(hessian_steps), _dp(dobs), _dp(dw)))
        return dobs.reshape(-1, 2), dw

    def camera_translations(self):
        """[C, 3] translations t of the current state (X_cam = R X + t)."""
        cams, _ = self.get_state_rt()
        return cams.reshape(-1, 12)[:, 9:12].copy()

    def oplus(self):
        """The BA sample's oplus (nelson/samples/BASchurPerfomances.cpp:163-170)
        with the last Schur increment, which stays on the device."""
        check(self._lib.nsn_oplus(self._h))

    def backupSolution(self):
        check(self._lib.nsn_backup(self._h))

    def rollbackSolution(self):
        check(self._lib.nsn_rollback(self._h))

    # DoubleSection::update(bool) ---------------------------------------------
    def update(self, hessian=True):
        c = C.c_double(0)
        check(self._lib.nsn_linearize(self._h, 1 if hessian else 0, C.byref(c)))
        self._chi2 = c.value
        return c.value

    def chi2(self):
        return self._chi2

    def max_abs_b(self):
        v = C.c_double(0)
        check(self._lib.nsn_max_abs_b(self._h, C.byref(v)))
        return v.value

    def export_hessian(self):
        NU, NV, nW = self.NU, self.NV, self.nW
        U = np.zeros((NU, 6, 6)); V = np.zeros((NV, 3, 3)); W = np.zeros((nW, 6, 3))
        Wi = np.zeros(nW, np.int32); Wj = np.zeros(nW, np.int32)
        bU = np.zeros(6 * NU); bV = np.zeros(3 * NV); chi2 = C.c_double(0)
        check(self._lib.nsn_export_hessian(self._h, _dp(U), _dp(V), _dp(W), _ip(Wi), _ip(Wj), _dp(bU), _dp(bV),
                                           C.byref(chi2)))
        return dict(U=U, V=V, W=W, Wi=Wi, Wj=Wj, bU=bU, bV=bV, chi2=chi2.value)

    def import_hessian(self, U, V, W, bU, bV, chi2):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (U, V, W, bU, bV)]
        check(self._lib.nsn_import_hessian(self._h, *[_dp(a) for a in arrs], float(chi2)))


class SolverCholeskySchur:
    """SolverCholeskySchur (nelson/SolverCholeskySchur.h:53-117) on the GPU."""

    def __init__(self):
        self._p = None
        self._inc = None
        self._ok = False

    def init(self, problem):
        self._p = problem
        check(capi.lib().nsn_schur_init(problem.handle))

    def maxAbsHDiag(self):
        v = C.c_double(0)
        check(capi.lib().nsn_max_abs_h_diag(self._p.handle, C.byref(v)))
        return v.value

    def computeIncrement(self, relLambda, absLambda):
        p = self._p
        dxu = np.zeros(6 * p.NU)
        dxv = np.zeros(3 * p.NV)
        sq = C.c_double(0)
        rc = capi.lib().nsn_schur_step(p.handle, relLambda, absLambda, _dp(dxu), _dp(dxv), C.byref(sq))
        check(rc, allow=(capi.NSN_OK, capi.NSN_NOT_PD))
        self._inc = (dxu, dxv)
        self._sq = sq.value
        self._ok = rc == capi.NSN_OK
        return self._ok

    def incrementVector(self):
        return self._inc

    def incrementVectorSquaredNorm(self):
        return self._sq

    def phase_times_ms(self):
        ms = np.zeros(6)
        check(capi.lib().nsn_phase_times(self._p.handle, _dp(ms), 6))
        return dict(zip(["vinv_wvinv", "s_assembly", "factorize", "solves", "backsub", "refine"], ms))


class LevenbergMarquardtSettings(dict):
    """LevenbergMarquardtSettings (nelson/LevenbergMarquardt.h:18-34) with the
    defaults of nelson/LevenbergMarquardt.cpp:6-21."""

    def __init__(self, **kw):
        s = capi.LMSettings()
        capi.lib().nsn_lm_default_settings(C.byref(s))
        super().__init__({n: getattr(s, n) for n, _ in capi.LMSettings._fields_})
        self.update(kw)

    def to_c(self):
        s = capi.LMSettings()
        for n, _ in capi.LMSettings._fields_:
            setattr(s, n, self[n])
        return s


class GaussNewtonSettings(dict):
    """GaussNewtonSettings (nelson/GaussNewton.h:18-30), defaults GaussNewton.cpp:12-22."""

    def __init__(self, **kw):
        s = capi.GNSettings()
        capi.lib().nsn_gn_default_settings(C.byref(s))
        super().__init__({n: getattr(s, n) for n, _ in capi.GNSettings._fields_})
        self.update(kw)

    def to_c(self):
        s = capi.GNSettings()
        for n, _ in capi.GNSettings._fields_:
            setattr(s, n, self[n])
        return s


class _Driver:
    _settings_cls = None
    _fn = None

    def __init__(self, settings=None):
        self._settings = settings or self._settings_cls()
        self._res = None
        self._log = []

    def settings(self):
        return self._settings

    def solve(self, problem):
        s = self._settings.to_c()
        res = capi.LMResult()
        cap = 4096
        log = (capi.LogEntry * cap)()
        check(getattr(capi.lib(), self._fn)(problem.handle, C.byref(s), C.byref(res), log, cap))
        self._res = res
        self._log = [(log[i].it, log[i].subit, log[i].chi2, log[i].mu) for i in range(min(res.n_log, cap))]
        problem._chi2 = res.final_chi2
        return capi.TERMINATION_REASONS[res.reason]

    def stats(self):
        """[(it, subit, chi2, mu)], chi2 == -1 on a solver failure (LevenbergMarquardtStats)."""
        return list(self._log)

    def numIterations(self):
        return self._res.iterations

    def result(self):
        r = self._res
        return dict(reason=capi.TERMINATION_REASONS[r.reason], iterations=r.iterations,
                    sub_iterations=r.sub_iterations, final_chi2=r.final_chi2, elapsed_ms=r.elapsed_ms)


class LevenbergMarquardt(_Driver):
    """LevenbergMarquardt<SolverCholeskySchur>::solve with the loop on the device."""
    _settings_cls = LevenbergMarquardtSettings
    _fn = "nsn_lm_solve"


class GaussNewton(_Driver):
    """GaussNewton<SolverCholeskySchur>::solve."""
    _settings_cls = GaussNewtonSettings
    _fn = "nsn_gn_solve"
