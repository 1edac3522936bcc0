"""Error paths of the C ABI on the GPU: every device allocation of
nsn_problem_create is checked (NSN_ENOMEM, no handle, no fault), using the
library's allocation-failure injection (NSN_TEST_FAIL_ALLOC = index of the
allocation that fails)."""
import os

import pytest

from nelson_amd import BAProblem, LevenbergMarquardt, LevenbergMarquardtSettings, generate_ba
from nelson_amd.capi import NSN_ENOMEM, NelsonError

from helpers import small_cfg

pytestmark = pytest.mark.gpu


# 65 allocations when the problem has no per-camera intrinsics (indices 0..64)
@pytest.mark.parametrize("k", [0, 1, 7, 20, 33, 45, 60, 63, 64])
def test_each_allocation_failure_returns_enomem(k):
    g = generate_ba(small_cfg(n_cameras=6, n_points=80, k_min=3, k_max=4, seed=31))
    os.environ["NSN_TEST_FAIL_ALLOC"] = str(k)
    try:
        with pytest.raises(NelsonError) as e:
            BAProblem.from_generated(g)
        assert e.value.code == NSN_ENOMEM, e.value
    finally:
        del os.environ["NSN_TEST_FAIL_ALLOC"]
    # the library is still usable afterwards
    p = BAProblem.from_generated(g)
    assert LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=3)).solve(p) != "SolverFailure"


def test_injection_past_the_last_allocation_is_harmless():
    g = generate_ba(small_cfg(n_cameras=6, n_points=80, k_min=3, k_max=4, seed=31))
    os.environ["NSN_TEST_FAIL_ALLOC"] = "100000"
    try:
        p = BAProblem.from_generated(g)
    finally:
        del os.environ["NSN_TEST_FAIL_ALLOC"]
    assert LevenbergMarquardt(LevenbergMarquardtSettings(maxNumIt=3)).solve(p) != "SolverFailure"
