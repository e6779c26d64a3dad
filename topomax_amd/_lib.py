"""ctypes binding of ``libtmx.so`` (C ABI declared in include/tmx.h).

The library is loaded from the package directory; there is no fallback: if it is missing
or cannot be loaded the import of any GPU entry point raises.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtmx.so")

TMX_OK = 0
TMX_ERR_INVALID = -1
TMX_ERR_HIP = -2
TMX_ERR_NOT_CONVERGED = -3
TMX_ERR_BREAKDOWN = -4
TMX_ERR_NONFINITE = -5
TMX_ERR_UNSUPPORTED = -6

TMX_LEFT, TMX_RIGHT, TMX_TOP, TMX_BOTTOM = 0, 1, 2, 3
TMX_FIX_LEFT, TMX_FIX_RIGHT, TMX_FIX_TOP, TMX_FIX_BOTTOM = 1, 2, 4, 8
TMX_FP64, TMX_MIXED, TMX_FP32, TMX_MIXED_IR, TMX_MIXED_RU = 0, 1, 2, 3, 4
TMX_MAX_TRACTIONS = 16
TMX_OPT_FUSED, TMX_OPT_TAIL, TMX_OPT_NU_PRE, TMX_OPT_NU_POST = 0, 1, 2, 3
TMX_OPT_OMEGA_ELAST, TMX_OPT_OMEGA_FILTER = 4, 5
TMX_OPT_NU_COARSE, TMX_OPT_FILTER_RTOL = 6, 7
TMX_OPT_CHEB_ELAST, TMX_OPT_CHEB_FILTER, TMX_OPT_CHEB_RATIO, TMX_OPT_BAND = 10, 11, 12, 13
TMX_OPT_BAND_FINE_FILTER, TMX_OPT_FINE_CELL_WEIGHTS, TMX_OPT_GRAPHS = 14, 15, 16
TMX_OPT_SWEEP_PAIRS, TMX_OPT_SWEEP_PAIRS_POST, TMX_OPT_SWEEP_PAIR_RESTRICT = 17, 18, 19
TMX_OPT_FILTER_PIPE, TMX_OPT_FILTER_DCT = 20, 21
TMX_OPT_PAIR_LEVELS, TMX_OPT_PAIR_ROWS_COARSE, TMX_OPT_PAIR_NU_COARSE = 22, 23, 24
TMX_OPT_BAND_WIDE = 25
TMX_OPT_BENCH_COLD = 26
TMX_OPT_RU_EVERY = 27
TMX_OPT_SETUP_OVERLAP = 28
TMX_OPT_NU_FINE_FILTER, TMX_OPT_NU_COARSE_FILTER = 8, 9
TMX_BATCH_MERGED, TMX_BATCH_ALONE = 0, 1

SIDE_CODES = {"Left": TMX_LEFT, "Right": TMX_RIGHT, "Top": TMX_TOP, "Bottom": TMX_BOTTOM}
FIX_BITS = {"Left": TMX_FIX_LEFT, "Right": TMX_FIX_RIGHT, "Top": TMX_FIX_TOP,
            "Bottom": TMX_FIX_BOTTOM}


class Traction(ctypes.Structure):
    _fields_ = [("side", ctypes.c_int), ("center", ctypes.c_double),
                ("length", ctypes.c_double), ("value", ctypes.c_double * 2)]


class BodyForce(ctypes.Structure):
    _fields_ = [("center", ctypes.c_double * 2), ("radius", ctypes.c_double),
                ("value", ctypes.c_double * 2)]


class ProblemDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("width", ctypes.c_double), ("height", ctypes.c_double),
        ("young_modulus", ctypes.c_double), ("poisson_ratio", ctypes.c_double),
        ("penalty_minimum", ctypes.c_double), ("filter_radius", ctypes.c_double),
        ("fixed_sides", ctypes.c_uint),
        ("n_tractions", ctypes.c_int), ("tractions", Traction * TMX_MAX_TRACTIONS),
        ("has_body_force", ctypes.c_int), ("body_force", BodyForce),
        ("precision", ctypes.c_int), ("degree", ctypes.c_int),
    ]


class SolveOpts(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("precision", ctypes.c_int), ("inner_rtol", ctypes.c_double),
                ("max_outer", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("outer_iterations", ctypes.c_int),
                ("rel_residual", ctypes.c_double), ("setup_ms", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("levels", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchReport(ctypes.Structure):
    _fields_ = [("merged_designs", ctypes.c_int), ("solo_designs", ctypes.c_int),
                ("launches", ctypes.c_int), ("iterations", ctypes.c_int * 3),
                ("record_ms", ctypes.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["iterations"] = list(self.iterations)
        return d


_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int

# name -> (restype, argtypes); every symbol declared in include/tmx.h
SIGNATURES = {
    "tmx_last_error": (ctypes.c_char_p, []),
    "tmx_abi_version": (_I, []),
    "tmx_problem_create": (_I, [ctypes.POINTER(ProblemDesc), ctypes.POINTER(_P)]),
    "tmx_problem_destroy": (None, [_P]),
    "tmx_set_stream": (_I, [_P, _P]),
    "tmx_problem_info": (_I, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                              ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(_I)]),
    "tmx_set_penalization": (_I, [_P, _D]),
    "tmx_filter_apply": (_I, [_P, _P, _P]),
    "tmx_filter_solve": (_I, [_P, _P, _P]),
    "tmx_mass_apply": (_I, [_P, _P, _P]),
    "tmx_elasticity_solve": (_I, [_P, _P, _P, ctypes.POINTER(SolveOpts),
                                  ctypes.POINTER(SolveStats)]),
    "tmx_compliance": (_I, [_P, _P, ctypes.POINTER(_D)]),
    "tmx_sensitivity": (_I, [_P, _P, _P, _P]),
    "tmx_objective_and_gradient": (_I, [_P, _P, ctypes.POINTER(_D), _P, _P, _P, _P,
                                        ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats)]),
    "tmx_matvec": (_I, [_P, _P, _P, _P, _I]),
    "tmx_load_vector": (_I, [_P, _P]),
    "tmx_integrate": (_I, [_P, _P, ctypes.POINTER(_D)]),
    "tmx_integrate_noalloc": (_I, [_I, _I, _D, _D, _P, _P, _P]),
    "tmx_autotune": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_D)]),
    "tmx_kernel_timing": (_I, [_P, _I, _I]),
    "tmx_kernel_timing_read": (_I, [_P, _I, ctypes.POINTER(_D), ctypes.POINTER(_I)]),
    "tmx_set_tiles": (_I, [_P, _I, _I]),
    "tmx_set_option": (_I, [_P, _I, _D]),
    "tmx_precondition": (_I, [_P, _I, _I, _P, _P, _P]),
    "tmx_bench_kernel": (_I, [_P, _I, _I, _I, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "tmx_batch_create": (_I, [ctypes.POINTER(_P), _I, ctypes.POINTER(_P)]),
    "tmx_batch_destroy": (None, [_P]),
    "tmx_batch_set_workers": (_I, [_P, _I]),
    "tmx_batch_set_mode": (_I, [_P, _I]),
    "tmx_batch_last_report": (_I, [_P, ctypes.POINTER(BatchReport)]),
    "tmx_batch_objective_and_gradient": (_I, [_P, ctypes.POINTER(_P), ctypes.POINTER(_D),
                                              ctypes.POINTER(_P), ctypes.POINTER(SolveOpts),
                                              ctypes.POINTER(SolveStats)]),
    "tmx_emd_half_step": (_I, [_P, _P, _P, _D, _P]),
    "tmx_emd_volume": (_I, [_P, _P, _D, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "tmx_emd_apply": (_I, [_P, _P, _D, _P, _P]),
    "tmx_emd_change": (_I, [_P, _P, _P, ctypes.POINTER(_D)]),
    "tmx_logit": (_I, [_P, _P, _P]),
    "tmx_interpolate_design": (_I, [_P, _I, _I, _P, _P]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load libtmx.so and bind every exported entry point (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m topomax_amd.build` "
            "(topomax_amd has no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class TmxError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class NotConvergedError(TmxError):
    pass


class NonFiniteError(TmxError):
    pass


def check(code: int):
    """Map a C status to the reference's exception classes: misuse -> ValueError,
    numerical / HIP failure -> RuntimeError subclasses."""
    if code == TMX_OK:
        return
    msg = load().tmx_last_error().decode(errors="replace")
    if code in (TMX_ERR_INVALID, TMX_ERR_UNSUPPORTED):
        raise ValueError(msg)
    if code == TMX_ERR_NOT_CONVERGED:
        raise NotConvergedError(code, msg)
    if code == TMX_ERR_NONFINITE:
        raise NonFiniteError(code, msg)
    raise TmxError(code, msg)
