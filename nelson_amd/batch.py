"""Batched independent BA windows over the C ABI (nsn_batch_*): many small
problems of uneven size solved by Levenberg-Marquardt in one persistent launch.
"""
import ctypes as C

import numpy as np

from . import capi
from .capi import check
from .problem import LevenbergMarquardtSettings


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else None


class BABatch:
    """windows: generated problems (dicts of generate_ba); camera 0 and point 0
    of every window are fixed (gauge), like BAProblem."""

    def __init__(self, windows, precision="mixed", device=0):
        self._lib = capi.lib()
        self.n = len(windows)
        self._keep = []
        descs = (capi.ProblemDesc * self.n)()
        self.C, self.P, self.E = [], [], []
        for d, g in zip(descs, windows):
            cam_idx = np.ascontiguousarray(g["cam_idx"], np.int32)
            pt_idx = np.ascontiguousarray(g["pt_idx"], np.int32)
            obs = np.ascontiguousarray(g["obs"], np.float64).reshape(-1)
            cf = np.zeros(g["n_cameras"], np.uint8)
            cf[0] = 1
            pf = np.zeros(g["n_points"], np.uint8)
            pf[0] = 1
            w = g.get("weight")
            w = None if w is None else np.ascontiguousarray(w, np.float64)
            k = g.get("intrinsics")  # the BAL camera model: passed on so that the library can refuse it
            k = None if k is None else np.ascontiguousarray(k, np.float64).reshape(-1)
            self._keep += [cam_idx, pt_idx, obs, cf, pf, w, k]
            d.n_cameras, d.n_points, d.n_edges = g["n_cameras"], g["n_points"], len(cam_idx)
            d.cam_idx, d.pt_idx, d.obs, d.weight = _ip(cam_idx), _ip(pt_idx), _dp(obs), _dp(w)
            d.cam_fixed = cf.ctypes.data_as(C.POINTER(C.c_ubyte))
            d.pt_fixed = pf.ctypes.data_as(C.POINTER(C.c_ubyte))
            d.focal, d.cx, d.cy = g["focal"], g["cx"], g["cy"]
            d.huber_delta = g.get("huber_delta", 0.0)
            d.precision = capi.NSN_MIXED if precision == "mixed" else capi.NSN_FP64
            d.device = device
            d.intrinsics = _dp(k)
            self.C.append(g["n_cameras"])
            self.P.append(g["n_points"])
            self.E.append(len(cam_idx))
        h = C.c_void_p()
        check(self._lib.nsn_batch_create(descs, self.n, C.byref(h)))
        self._h = h
        self.set_state(np.concatenate([g["cams0"] for g in windows]), np.concatenate([g["pts0"] for g in windows]))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nsn_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, cams, pts):
        cams = np.ascontiguousarray(cams, np.float64)
        pts = np.ascontiguousarray(pts, np.float64)
        check(self._lib.nsn_batch_set_state(self._h, _dp(cams), _dp(pts)))

    def get_state(self):
        cams = np.zeros(6 * sum(self.C))
        pts = np.zeros(3 * sum(self.P))
        check(self._lib.nsn_batch_get_state(self._h, _dp(cams), _dp(pts)))
        return cams, pts

    def tune(self, settings=None):
        s = (settings or LevenbergMarquardtSettings()).to_c()
        nb, grid = C.c_int(0), C.c_int(0)
        check(self._lib.nsn_batch_tune(self._h, C.byref(s), C.byref(nb), C.byref(grid)))
        return nb.value, grid.value

    def set_config(self, variant, grid):
        """Factorisation variant (8 or 16: panels of that width, 64: 64x64
        tiles) and number of workgroups of the persistent launch."""
        check(self._lib.nsn_batch_set_config(self._h, int(variant), int(grid)))

    def config(self):
        nb, grid = C.c_int(0), C.c_int(0)
        ms = np.zeros(9)
        check(self._lib.nsn_batch_config(self._h, C.byref(nb), C.byref(grid), _dp(ms), 9))
        return dict(variant=nb.value, grid=grid.value, tune_ms=ms.tolist())

    def lm_solve(self, settings=None):
        """LM on every window; returns (final chi2 per window, device ms)."""
        s = (settings or LevenbergMarquardtSettings()).to_c()
        chi2 = np.zeros(self.n)
        ms = C.c_double(0)
        check(self._lib.nsn_batch_lm_solve(self._h, C.byref(s), _dp(chi2), C.byref(ms)))
        return chi2, ms.value

    def results(self):
        r = np.zeros(self.n, np.int32)
        it = np.zeros(self.n, np.int32)
        sub = np.zeros(self.n, np.int32)
        check(self._lib.nsn_batch_results(self._h, _ip(r), _ip(it), _ip(sub)))
        return [capi.TERMINATION_REASONS[int(x)] for x in r], it, sub

    def set_refinement(self, max_steps, tol=1e-10):
        """Mixed precision: at most max_steps iterative-refinement steps per
        increment against each window's FP64 residual, stopping once
        max|r| <= tol * max|b| (tol 0: every step runs); 0 (the default)
        turns it off.  FP64 batches ignore it."""
        check(self._lib.nsn_batch_set_refinement(self._h, int(max_steps), float(tol)))

    def refinement_counts(self):
        """Refinement solves per window in the last lm_solve."""
        n = np.zeros(self.n, np.int32)
        check(self._lib.nsn_batch_refinement_counts(self._h, _ip(n)))
        return n
