"""nelson_amd: MI355X (gfx950) Schur-complement Gauss-Newton / Levenberg-Marquardt
step for bundle adjustment, behind nelson's DoubleSection / SolverCholeskySchur /
LevenbergMarquardt interface.

The compute runs in libnelson_amd.so (hand-written HIP kernels + a C ABI,
include/nelson_amd.h).  This module is a thin ctypes layer over that ABI; it
has no CPU fallback and raises if the library is missing.
"""
from .capi import (  # noqa: F401
    NSN_OK, NSN_NOT_PD, NSN_FP64, NSN_MIXED, TERMINATION_REASONS, NelsonError, lib, library_path,
    load_library,
)
from .problem import (  # noqa: F401
    BAProblem, SolverCholeskySchur, LevenbergMarquardt, LevenbergMarquardtSettings, GaussNewton,
    GaussNewtonSettings, generate_ba, window_config, CONFIGS, load_visibility, generate_on_graph, LADYBUG_SAMPLE,
    load_bal,
)
from .batch import BABatch  # noqa: F401
