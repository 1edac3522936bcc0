"""nsn_adjoint_noalloc (the adjoint with its work vectors in the library's
static device array instead of allocated buffers) against nsn_adjoint,
through the C ABI: two calls in a row, the second on a smaller problem (the
array keeps the first call's contents), each equal to the allocating entry
point's result; a problem beyond the array's size is refused."""
import ctypes as C

import numpy as np
import pytest

from nelson_amd import BAProblem, CONFIGS, GaussNewton, GaussNewtonSettings, LevenbergMarquardt, \
    LevenbergMarquardtSettings, generate_ba
from nelson_amd.autograd import translation_loss_grad
from nelson_amd.capi import NSN_EINVAL, NSN_OK

pytestmark = pytest.mark.gpu

_D = C.POINTER(C.c_double)


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _solved(nc, npt, seed, precision="fp64"):
    g = generate_ba(dict(CONFIGS["cfg1"], n_cameras=nc, n_points=npt, k_min=3, k_max=6, seed=seed))
    p = BAProblem.from_generated(g, precision=precision)
    LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=200)).solve(p)
    GaussNewton(GaussNewtonSettings(maxNumIt=8, epsIncSquare=0.0, epsBVector=0.0, epsChi2=0.0)).solve(p)
    t = p.camera_translations()
    t_gt = g["cams_gt"].reshape(-1, 6)[:, 3:6]
    gc = np.ascontiguousarray(translation_loss_grad(t, 2.0 * (t - t_gt)).reshape(-1))
    return g, p, gc


def _call(p, fn, gc, steps):
    dobs = np.zeros(2 * p.n_edges)
    dw = np.zeros(p.n_edges)
    rc = getattr(p._lib, fn)(p._h, _dp(gc), None, steps, _dp(dobs), _dp(dw))
    return rc, dobs, dw


def test_noalloc_adjoint_equals_allocating_adjoint():
    cases = [(14, 400, 41, "fp64", 2), (6, 90, 42, "fp64", 0), (9, 150, 43, "mixed", 2)]
    for nc, npt, seed, prec, steps in cases:  # each smaller than the first
        g, p, gc = _solved(nc, npt, seed, prec)
        rc, dn, wn = _call(p, "nsn_adjoint_noalloc", gc, steps)
        assert rc == NSN_OK, p._lib.nsn_last_error()
        rc, da, wa = _call(p, "nsn_adjoint", gc, steps)
        assert rc == NSN_OK, p._lib.nsn_last_error()
        scale = max(np.abs(da).max(), np.abs(wa).max())
        err = max(np.abs(dn - da).max(), np.abs(wn - wa).max()) / scale
        print(f"\n{nc} cameras / {npt} points ({prec}, {steps} steps): no-alloc vs allocating adjoint {err:.1e}")
        assert err <= 1e-12, err
        # a second no-alloc call on the same problem after the allocating one
        rc, dn2, wn2 = _call(p, "nsn_adjoint_noalloc", gc, steps)
        assert rc == NSN_OK and np.array_equal(dn2, dn) and np.array_equal(wn2, wn)


def test_noalloc_adjoint_refuses_a_problem_beyond_the_array():
    g = generate_ba(CONFIGS["cfg2"])  # 7 x 6 NU + 7 x 3 NV + 3 E > 4 Mi doubles
    p = BAProblem.from_generated(g)
    gc = np.zeros(6 * g["n_cameras"])
    rc, _, _ = _call(p, "nsn_adjoint_noalloc", gc, 0)
    assert rc == NSN_EINVAL
    assert b"nsn_adjoint" in p._lib.nsn_last_error()
