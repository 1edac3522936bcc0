"""A single section (pose graph) through the C++ interface: the reference's
2-D scan alignment test (nelson/tests/testGaussNewton2DScansAlign.cpp) as the
program tests/cpp/test_scan2d writes it against nelson_amd::SingleSection,
on problems written here (tests/scan2d.py): SE(2) poses, point-to-line edges
between scans, full and sparse edge sets.  Pure and damped Gauss-Newton and
Levenberg-Marquardt must follow the oracle's drivers on the oracle's copy of
the problem (1e-9 per logged cost above the rounding floor), end with
chi2 < 1e-12 as the reference's test requires, and give the same poses in
the natural, minimum-degree (permuteAMD) and reversed internal orders; the
same edges written as n-ary edges (EdgeNary, nelson/EdgeNary.h:125-150) give
the same runs."""
import json
import os
import subprocess

import numpy as np
import pytest

from scan2d import ScanProblem, make, write

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_scan2d")
SUCCESS = {"SuccessSmallIncrement", "SuccessEpsB", "SuccessSmallResiduals"}


def _run(d, tmp_path):
    path = tmp_path / "scans.txt"
    write(d, str(path))
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (nelson_amd.build builds it with the library)")
    r = subprocess.run([EXE, str(path)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return {x["solve"]: x for x in (json.loads(line) for line in r.stdout.splitlines() if line.strip())}


def _same_logs(a, b, what):
    for x, y in zip(a, b):
        assert x[0] == y[0] and x[1] == y[1], (what, a, b)
        assert abs(x[2] - y[2]) <= 1e-9 * abs(y[2]) + 1e-18, (what, a, b)
        assert abs(x[3] - y[3]) <= 1e-9 * abs(y[3]) + 1e-300, (what, a, b)
        if y[2] < 1e-12:  # at the rounding floor the stopping tests may fire one entry apart
            return
    assert len(a) == len(b), (what, a, b)


@pytest.mark.parametrize("full,seed", [(True, 1), (False, 2), (True, 3)])
def test_scan_alignment_single_section(full, seed, tmp_path):
    d = make(full=full, seed=seed)
    res = _run(d, tmp_path)
    runs = {"gn": dict(solve="gn_solve"), "gn_damped": dict(solve="gn_solve", relLambda=0.01),
            "lm": dict(solve="lm_solve")}
    for name, kw in runs.items():
        kw = dict(kw)
        p = ScanProblem(d)
        ro = getattr(p, kw.pop("solve"))(**kw)
        for suffix in ("", "_amd", "_reversed"):
            r = res[name + suffix]
            assert r["reason"] == ro["reason"] or {r["reason"], ro["reason"]} <= SUCCESS, (name + suffix, r, ro)
            _same_logs(r["log"], ro["log"], name + suffix)
            assert r["chi2_after"] < 1e-12, (name + suffix, r["chi2_after"])
            assert np.allclose(np.array(r["poses"]).reshape(-1, 3), p.poses, rtol=0, atol=1e-9), name + suffix
        # the same edges as n-ary edges (nelson/tests/testGaussNewton2DScansAlignNary.cpp), N = 2
        # fixed and set at addEdge, natural and minimum-degree order
        if name == "gn_damped":
            continue
        for suffix in ("_nary", "_nary_amd", "_nary_dynamic", "_nary_dynamic_amd"):
            r = res[name + suffix]
            assert r["reason"] == ro["reason"] or {r["reason"], ro["reason"]} <= SUCCESS, (name + suffix, r, ro)
            _same_logs(r["log"], ro["log"], name + suffix)
            assert r["chi2_after"] < 1e-12, (name + suffix, r["chi2_after"])
            assert np.allclose(np.array(r["poses"]).reshape(-1, 3), p.poses, rtol=0, atol=1e-9), name + suffix
