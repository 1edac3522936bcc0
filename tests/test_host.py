"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol its header declares, the Python bindings cover them, the
settings defaults are the reference's, and the seeded generator."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import nelson_amd as nsn
from nelson_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"\b(nsn_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(capi.library_path)
    declared = _declared()
    assert len(declared) > 25
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert _declared() == set(capi.SIGNATURES)


def test_lm_gn_defaults_match_reference():
    """nelson/LevenbergMarquardt.cpp:6-21 and nelson/GaussNewton.cpp:12-22."""
    s = nsn.LevenbergMarquardtSettings()
    assert (s["maxNumIt"], s["maxNumSubIt"], s["minNumIt"]) == (100, 10, 0)
    assert (s["epsBVector"], s["epsIncSquare"], s["epsChi2"]) == (1e-15, 1e-30, 1e-15)
    assert (s["tau"], s["penaltyV"], s["initV"]) == (1e-3, 2.0, 2.0)
    assert (s["maxScaleFactor"], s["minScaleFactor"]) == (2.0 / 3.0, 1.0 / 3.0)
    g = nsn.GaussNewtonSettings()
    assert (g["maxNumIt"], g["minNumIt"], g["absLambda"], g["relLambda"]) == (100, 0, 0.0, 0.0)


def test_generator_is_deterministic_and_sized(cfg1):
    g2 = nsn.generate_ba(nsn.CONFIGS["cfg1"])
    for k in ("cams0", "pts0", "obs", "cam_idx", "pt_idx"):
        assert np.array_equal(cfg1[k], g2[k])
    assert cfg1["n_edges"] == 2000
    # each point seen by 4 distinct cameras
    for q in range(cfg1["n_points"]):
        cams = cfg1["cam_idx"][cfg1["pt_idx"] == q]
        assert len(cams) == 4 and len(set(cams)) == 4
    # gauge: camera 0 and point 0 unperturbed
    assert np.array_equal(cfg1["cams0"][:6], cfg1["cams_gt"][:6])
    assert np.array_equal(cfg1["pts0"][:3], cfg1["pts_gt"][:3])
    # noise-free projections of the ground truth are within ~5 sigma of the observations
    from scipy.spatial.transform import Rotation
    e = np.arange(cfg1["n_edges"])
    R = Rotation.from_rotvec(cfg1["cams_gt"].reshape(-1, 6)[cfg1["cam_idx"], :3]).as_matrix()
    t = cfg1["cams_gt"].reshape(-1, 6)[cfg1["cam_idx"], 3:]
    X = cfg1["pts_gt"].reshape(-1, 3)[cfg1["pt_idx"]]
    Xc = np.einsum("eij,ej->ei", R, X) + t
    assert np.all(Xc[:, 2] > 4.0) and np.all(Xc[:, 2] < 11.0)
    uv = 500.0 * Xc[:, :2] / Xc[:, 2:3] + np.array([320.0, 240.0])
    res = uv - cfg1["obs"].reshape(-1, 2)[e]
    assert 0.8 < res.std() < 1.2 and np.abs(res).max() < 6.0


def test_ring_generator_covisibility():
    """Layout 1: every point is observed from a contiguous arc of `arc`
    cameras, so cameras i, j share a point only if their circular distance is
    below the arc (S is banded with wrap-around, ~2*arc/C dense)."""
    cfg = dict(nsn.CONFIGS["cfg2"], n_cameras=200, n_points=5000, arc=30)
    g = nsn.generate_ba(cfg)
    C_ = cfg["n_cameras"]
    by_pt = {}
    for c, q in zip(g["cam_idx"], g["pt_idx"]):
        by_pt.setdefault(q, []).append(c)
    worst = 0
    for cams in by_pt.values():
        for a in cams:
            for b in cams:
                d = abs(int(a) - int(b))
                worst = max(worst, min(d, C_ - d))
    assert worst < 30
    ks = np.bincount(g["pt_idx"])
    assert ks.min() >= 4 and ks.max() <= 8


def test_generator_rejects_small_capacity():
    p = capi.GenParams()
    for k, v in nsn.CONFIGS["cfg1"].items():
        if hasattr(p, k):
            setattr(p, k, v)
    buf = np.zeros(100)
    ib = np.zeros(10, np.int32)
    ne = C.c_int(0)
    dp = buf.ctypes.data_as(C.POINTER(C.c_double))
    ip = ib.ctypes.data_as(C.POINTER(C.c_int))
    big = np.zeros(6000)
    bp = big.ctypes.data_as(C.POINTER(C.c_double))
    rc = capi.lib().nsn_generate_ba(C.byref(p), 10, bp, bp, bp, bp, ip, ip, bp, C.byref(ne))
    assert rc == capi.NSN_EINVAL
    assert b"edge_cap" in capi.lib().nsn_last_error()


def test_problem_create_without_gpu_fails_loudly():
    n = C.c_int(0)
    capi.lib().nsn_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a GPU is present")
    g = nsn.generate_ba(dict(nsn.CONFIGS["cfg1"], n_points=20))
    with pytest.raises(capi.NelsonError) as ei:
        nsn.BAProblem.from_generated(g)
    assert ei.value.code == capi.NSN_EHIP


def test_cpp_interface_compiles_and_links():
    """include/nelson_amd.hpp compiles as C++17 and the facade test program
    is built next to the library (it runs in the GPU suite)."""
    import subprocess
    src = os.path.join(ROOT, "tests", "cpp", "test_facade.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        src], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    assert os.path.exists(os.path.join(ROOT, "tests", "cpp", "test_facade"))


def _block_desc(n_u, n_v, w, uo):
    lib = capi.lib()
    d = capi.BlockDesc()
    wu = np.array([a for a, _ in w] or [0], np.int32)
    wv = np.array([b for _, b in w] or [0], np.int32)
    oi = np.array([a for a, _ in uo] or [0], np.int32)
    oj = np.array([b for _, b in uo] or [0], np.int32)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    d.n_u, d.n_v, d.n_w, d.w_u, d.w_v = n_u, n_v, len(w), ip(wu), ip(wv)
    d.n_uoff, d.uoff_i, d.uoff_j = len(uo), ip(oi), ip(oj)
    d.precision, d.device = capi.NSN_FP64, 0
    h = C.c_void_p()
    rc = lib.nsn_problem_create_blocks(C.byref(d), C.byref(h))
    return rc, lib.nsn_last_error().decode()


@pytest.mark.parametrize("n_u,n_v,w,uo,what", [
    (0, 2, [], [], "bad sizes"),
    (3, 2, [(0, 0), (3, 1)], [], "out of range"),
    (3, 2, [(0, 1), (0, 1)], [], "listed twice"),
    (3, 2, [(0, 1)], [(1, 1)], "strictly upper"),
    (3, 2, [(0, 1)], [(2, 1)], "strictly upper"),
    (3, 2, [(0, 1)], [(0, 2), (0, 2)], "listed twice"),
])
def test_block_problem_structure_is_validated(n_u, n_v, w, uo, what):
    """nsn_problem_create_blocks checks the block structure on the host
    before touching a device: repeated or out-of-range W blocks, off-diagonal
    U blocks that are not strictly upper or repeated."""
    rc, msg = _block_desc(n_u, n_v, w, uo)
    assert rc == capi.NSN_EINVAL and what in msg, (rc, msg)


def _chains_from_order(g, order):
    """Independent restatement: the tile rows of S (64 rows, 6 per camera in
    solver order) whose first non-zero block lies in the row's own tile
    column, the last row (the right-hand side) dense."""
    pos = np.full(g["n_cameras"], -1)
    pos[order] = np.arange(len(order))
    NT = (6 * len(order) + 1 + 63) // 64
    nz = np.arange(NT)
    cam, pt = np.asarray(g["cam_idx"]), np.asarray(g["pt_idx"])
    keep = (pos[cam] >= 0) & (pt != 0)  # variable cameras, variable points (point 0 is fixed)
    cam, pt = pos[cam[keep]], pt[keep]
    srt = np.lexsort((cam, pt))
    cam, pt = cam[srt], pt[srt]
    starts = np.r_[0, np.flatnonzero(np.diff(pt)) + 1, len(pt)]
    for a, b in zip(starts[:-1], starts[1:]):
        c = np.unique(cam[a:b])
        if len(c) < 2:
            continue
        lo = c.min()
        for h in c[1:]:
            for ti in range(6 * h // 64, (6 * h + 5) // 64 + 1):
                nz[ti] = min(nz[ti], 6 * lo // 64)
    nz[NT - 1] = 0
    return int(np.sum(nz == np.arange(NT))), NT


@pytest.mark.parametrize("name,cfg,min_chains", [
    ("ring", dict(nsn.CONFIGS["cfg2"], n_cameras=300, arc=40, n_points=30000, seed=31), 2),
    ("path", dict(nsn.CONFIGS["seq"], n_cameras=600, arc=30, n_points=60000, seed=33), 4),
])
def test_nested_dissection_plan(name, cfg, min_chains):
    """nsn_problem_plan on the host: the nested-dissection order is a
    permutation of the variable cameras whose independent parts start on
    tile boundaries (chain count equal to a numpy restatement of the tile
    envelope), with a shorter critical path than the given order; the given
    order is one chain of NT steps."""
    g = nsn.generate_ba(cfg)
    gi, go = nsn.BAProblem.plan(g, "given")
    ni, no = nsn.BAProblem.plan(g, "nd")
    assert sorted(go.tolist()) == list(range(1, g["n_cameras"])) == go.tolist()
    assert sorted(no.tolist()) == list(range(1, g["n_cameras"]))
    nc, NT = _chains_from_order(g, go)
    assert (gi["n_chains"], gi["chain_steps"], nc) == (1, NT, 1)
    nc, NT = _chains_from_order(g, no)
    assert ni["n_chains"] == nc >= min_chains, (ni["n_chains"], nc)
    assert ni["chain_steps"] < 0.8 * NT, (ni["chain_steps"], NT)
    assert ni["factor_flops"] > 0


@pytest.mark.skipif(shutil.which("nm") is None, reason="nm not found")
def test_library_marks_its_calls_with_roctx_ranges():
    """The C ABI's long calls push and pop roctx ranges (rocprofiler-sdk-roctx),
    so a rocprofv3 --marker-trace run shows them beside their kernels."""
    out = subprocess.run(["nm", "-D", "--undefined-only", capi.library_path], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    assert "roctxRangePushA" in out and "roctxRangePop" in out
