"""ctypes declarations of the C ABI in include/nelson_amd.h."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(HERE, "libnelson_amd.so")

NSN_OK = 0
NSN_NOT_PD = 1
NSN_EINVAL = -1
NSN_EHIP = -2
NSN_ENOMEM = -3
NSN_ESTATE = -4
NSN_EDEVICE = -5
NSN_FP64 = 0
NSN_MIXED = 1

# nelson/LevenbergMarquardt.cpp:26-34
TERMINATION_REASONS = [
    "SuccessAtStart", "SuccessSmallIncrement", "SuccessEpsB", "SuccessSmallResiduals",
    "SolverFailure", "MaxIterationReached", "MaxSubIterationReached",
]


class NelsonError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nelson_amd error {code}: {msg}")
        self.code = code


class BlockDesc(C.Structure):
    _fields_ = [
        ("n_u", C.c_int), ("n_v", C.c_int), ("n_w", C.c_int), ("w_u", C.POINTER(C.c_int)),
        ("w_v", C.POINTER(C.c_int)), ("n_uoff", C.c_int), ("uoff_i", C.POINTER(C.c_int)),
        ("uoff_j", C.POINTER(C.c_int)), ("precision", C.c_int), ("device", C.c_int),
    ]


class ProblemDesc(C.Structure):
    _fields_ = [
        ("n_cameras", C.c_int), ("n_points", C.c_int), ("n_edges", C.c_int),
        ("cam_idx", C.POINTER(C.c_int)), ("pt_idx", C.POINTER(C.c_int)),
        ("obs", C.POINTER(C.c_double)), ("weight", C.POINTER(C.c_double)),
        ("cam_fixed", C.POINTER(C.c_ubyte)), ("pt_fixed", C.POINTER(C.c_ubyte)),
        ("focal", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
        ("huber_delta", C.c_double), ("precision", C.c_int), ("device", C.c_int),
        ("intrinsics", C.POINTER(C.c_double)), ("camera_order", C.c_int),
    ]


NSN_ORDER_AS_GIVEN = 0
NSN_ORDER_RCM = 1
NSN_ORDER_ND = 2


class ProblemInfo(C.Structure):
    _fields_ = [("n_var_cameras", C.c_int), ("n_var_points", C.c_int), ("n_wblocks", C.c_int),
                ("n_s", C.c_int), ("n_s_blocks", C.c_int), ("n_s_pairs", C.c_longlong), ("n_tiles", C.c_int),
                ("n_tiles_dense", C.c_int), ("factor_flops", C.c_double),
                ("n_chains", C.c_int), ("chain_steps", C.c_int), ("lookahead", C.c_int),
                ("chunk_tiles", C.c_int)]


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
    _fields_ = [("reason", C.c_int), ("iterations", C.c_int), ("sub_iterations", C.c_int),
                ("n_log", C.c_int), ("final_chi2", C.c_double), ("elapsed_ms", C.c_double)]


class GenParams(C.Structure):
    _fields_ = [("n_cameras", C.c_int), ("n_points", C.c_int), ("k_min", C.c_int), ("k_max", C.c_int),
                ("layout", C.c_int), ("arc", C.c_int), ("obs_sigma", C.c_double),
                ("outlier_frac", C.c_double), ("outlier_sigma", C.c_double),
                ("rot_sigma", C.c_double), ("trans_sigma", C.c_double), ("point_sigma", C.c_double),
                ("focal", C.c_double), ("cx", C.c_double), ("cy", C.c_double), ("seed", C.c_ulonglong)]


class GraphGenParams(C.Structure):
    _fields_ = [("focal", C.c_double), ("cx", C.c_double), ("cy", C.c_double), ("point_radius", C.c_double),
                ("cam_radius", C.c_double), ("obs_sigma", C.c_double), ("rot_sigma", C.c_double),
                ("trans_sigma", C.c_double), ("point_sigma", C.c_double), ("seed", C.c_ulonglong)]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)

# name -> (restype, argtypes)
SIGNATURES = {
    "nsn_last_error": (C.c_char_p, []),
    "nsn_version": (C.c_int, [_I, _I]),
    "nsn_device_count": (C.c_int, [_I]),
    "nsn_lm_default_settings": (None, [C.POINTER(LMSettings)]),
    "nsn_gn_default_settings": (None, [C.POINTER(GNSettings)]),
    "nsn_generate_ba": (C.c_int, [C.POINTER(GenParams), C.c_int, _D, _D, _D, _D, _I, _I, _D, _I]),
    "nsn_generate_on_graph": (C.c_int, [C.POINTER(GraphGenParams), C.c_int, C.c_int, C.c_int, _I, _I, _D, _D, _D,
                                        _D, _D]),
    "nsn_visibility_load": (C.c_int, [C.c_char_p, C.c_int, _I, _I, _I, _I, _I]),
    "nsn_bal_load": (C.c_int, [C.c_char_p, C.c_int, _I, _I, _I, _I, _I, _D, _D, _D, _D]),
    "nsn_problem_create": (C.c_int, [C.POINTER(ProblemDesc), C.POINTER(_P)]),
    "nsn_structure_check": (C.c_int, [C.POINTER(ProblemDesc), C.POINTER(C.c_longlong), _D, _D]),
    "nsn_problem_destroy": (C.c_int, [_P]),
    "nsn_problem_info_get": (C.c_int, [_P, C.POINTER(ProblemInfo)]),
    "nsn_problem_plan": (C.c_int, [C.POINTER(ProblemDesc), C.POINTER(ProblemInfo), _I]),
    "nsn_set_state": (C.c_int, [_P, _D, _D]),
    "nsn_get_state": (C.c_int, [_P, _D, _D]),
    "nsn_get_state_rt": (C.c_int, [_P, _D, _D]),
    "nsn_set_state_rt": (C.c_int, [_P, _D, _D]),
    "nsn_backup": (C.c_int, [_P]),
    "nsn_rollback": (C.c_int, [_P]),
    "nsn_linearize": (C.c_int, [_P, C.c_int, _D]),
    "nsn_max_abs_b": (C.c_int, [_P, _D]),
    "nsn_export_hessian": (C.c_int, [_P, _D, _D, _D, _I, _I, _D, _D, _D]),
    "nsn_import_hessian": (C.c_int, [_P, _D, _D, _D, _D, _D, C.c_double]),
    "nsn_schur_init": (C.c_int, [_P]),
    "nsn_max_abs_h_diag": (C.c_int, [_P, _D]),
    "nsn_schur_step": (C.c_int, [_P, C.c_double, C.c_double, _D, _D, _D]),
    "nsn_oplus": (C.c_int, [_P]),
    "nsn_rho": (C.c_int, [_P, C.c_double, C.c_double, _D]),
    "nsn_lm_solve": (C.c_int, [_P, C.POINTER(LMSettings), C.POINTER(LMResult), C.POINTER(LogEntry), C.c_int]),
    "nsn_gn_solve": (C.c_int, [_P, C.POINTER(GNSettings), C.POINTER(LMResult), C.POINTER(LogEntry), C.c_int]),
    "nsn_phase_times": (C.c_int, [_P, _D, C.c_int]),
    "nsn_set_refinement": (C.c_int, [_P, C.c_int]),
    "nsn_set_refinement_tol": (C.c_int, [_P, C.c_int, C.c_double]),
    "nsn_refinement_count": (C.c_int, [_P, _I]),
    "nsn_problem_create_blocks": (C.c_int, [C.POINTER(BlockDesc), C.POINTER(_P)]),
    "nsn_import_hessian_blocks": (C.c_int, [_P, _D, _D, _D, _D, _D, _D, C.c_double]),
    "nsn_adjoint": (C.c_int, [_P, _D, _D, C.c_int, _D, _D]),
    "nsn_adjoint_noalloc": (C.c_int, [_P, _D, _D, C.c_int, _D, _D]),
    "nsn_set_observations": (C.c_int, [_P, _D, _D]),
    "nsn_batch_create": (C.c_int, [C.POINTER(ProblemDesc), C.c_int, C.POINTER(_P)]),
    "nsn_batch_destroy": (C.c_int, [_P]),
    "nsn_batch_set_state": (C.c_int, [_P, _D, _D]),
    "nsn_batch_get_state": (C.c_int, [_P, _D, _D]),
    "nsn_batch_lm_solve": (C.c_int, [_P, C.POINTER(LMSettings), _D, _D]),
    "nsn_batch_results": (C.c_int, [_P, _I, _I, _I]),
    "nsn_batch_tune": (C.c_int, [_P, C.POINTER(LMSettings), _I, _I]),
    "nsn_batch_set_config": (C.c_int, [_P, C.c_int, C.c_int]),
    "nsn_batch_config": (C.c_int, [_P, _I, _I, _D, C.c_int]),
    "nsn_batch_set_refinement": (C.c_int, [_P, C.c_int, C.c_double]),
    "nsn_batch_refinement_counts": (C.c_int, [_P, _I]),
    "nsn_measure_peaks": (C.c_int, [C.c_int, _D, _D, _D]),
    "nsn_set_priors": (C.c_int, [_P, C.c_int, _I, _D, _D, C.c_int, _I, _D, _D]),
}

_lib = None


def load_library(path=None):
    """Loads libnelson_amd.so (raises FileNotFoundError when it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or library_path
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} is not built; run `python -m nelson_amd.build` (hipcc, gfx950) first")
    lib_ = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib_, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib_
    return lib_


def lib():
    return load_library()


def check(rc, allow=(NSN_OK,)):
    if rc in allow:
        return rc
    msg = load_library().nsn_last_error()
    raise NelsonError(rc, msg.decode() if msg else "")
