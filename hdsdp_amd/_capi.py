"""ctypes signatures of the solver C API (include/hdsdp.h) and the device ABI
(include/hdsdp_gpu.h).  ``bind(cdll)`` declares them on a loaded library."""
from __future__ import annotations

import ctypes as C

P_INT = C.POINTER(C.c_int)
P_DBL = C.POINTER(C.c_double)
P_VOID = C.c_void_p

_API = {
    "HDSDPCreate": (C.c_int, [C.POINTER(P_VOID)]),
    "HDSDPInit": (C.c_int, [P_VOID, C.c_int, C.c_int]),
    "HDSDPSetCone": (C.c_int, [P_VOID, C.c_int, P_VOID]),
    "HDSDPSetDualObjective": (None, [P_VOID, P_DBL]),
    "HDSDPOptimize": (C.c_int, [P_VOID, C.c_int]),
    "HDSDPSetIntParam": (None, [P_VOID, C.c_int, C.c_int]),
    "HDSDPSetDblParam": (None, [P_VOID, C.c_int, C.c_double]),
    "HDSDPSetIntParamAny": (None, [P_VOID, C.c_int, C.c_int]),
    "HDSDPSetDblParamAny": (None, [P_VOID, C.c_int, C.c_double]),
    "HDSDPSetPrecision": (None, [P_VOID, C.c_int]),
    "HDSDPSetVerbose": (None, [P_VOID, C.c_int]),
    "HDSDPSetProfiling": (None, [P_VOID, C.c_int]),
    "HDSDPGetProfile": (C.c_int, [P_VOID, P_DBL, P_DBL, C.c_int]),
    "HDSDPSetup": (C.c_int, [P_VOID]),
    "HDSDPGetBarrierPoint": (C.c_int, [P_VOID, P_DBL, P_DBL]),
    "HDSDPCenter": (C.c_int, [P_VOID, C.c_double, P_DBL, C.c_int, C.c_double, P_INT, P_DBL]),
    "HDSDPAdjoint": (C.c_int, [P_VOID, C.c_double, P_DBL, P_DBL, P_DBL, P_DBL]),
    "HDSDPGetRowDual": (C.c_int, [P_VOID, P_DBL, P_DBL, P_DBL]),
    "HDSDPGetConeValues": (None, [P_VOID, C.c_int, P_DBL, P_DBL, P_DBL]),
    "HDSDPCheckSolution": (C.c_int, [P_VOID, P_DBL]),
    "HDSDPGetDimacsErrors": (None, [P_VOID, P_DBL]),
    "HDSDPGetIterLog": (C.c_int, [P_VOID, C.c_int, P_DBL]),
    "HDSDPGetSummary": (None, [P_VOID, C.c_double, P_INT, P_INT, P_DBL, P_DBL, P_DBL]),
    "HDSDPClear": (None, [P_VOID]),
    "HDSDPDestroy": (None, [C.POINTER(P_VOID)]),
    "HUserDataCreate": (C.c_int, [C.POINTER(P_VOID)]),
    "HUserDataSetConeData": (None, [P_VOID, C.c_int, C.c_int, C.c_int, P_INT, P_INT, P_DBL]),
    "HUserDataSetRankOneConeData": (None, [P_VOID, C.c_int, C.c_int, C.c_int, P_INT, P_DBL, P_DBL, P_DBL]),
    "HUserDataChooseCone": (C.c_int, [P_VOID]),
    "HUserDataClear": (None, [P_VOID]),
    "HUserDataDestroy": (None, [C.POINTER(P_VOID)]),
    # SDPA input (include/hdsdp_file_io.h)
    "HReadSDPA": (C.c_int, [C.c_char_p, P_INT, P_INT, C.POINTER(P_INT), C.POINTER(P_DBL),
                            C.POINTER(C.POINTER(P_INT)), C.POINTER(C.POINTER(P_INT)), C.POINTER(C.POINTER(P_DBL)),
                            P_INT, P_INT, C.POINTER(P_INT), C.POINTER(P_INT), C.POINTER(P_DBL), P_INT]),
    "HFreeSDPA": (None, [C.c_int, P_INT, P_DBL, C.POINTER(P_INT), C.POINTER(P_INT), C.POINTER(P_DBL), P_INT, P_INT,
                         P_DBL]),
    # device ABI
    "hdsdp_dev_create": (C.c_int, [C.POINTER(P_VOID), C.c_int, C.c_int, C.c_int]),
    "hdsdp_dev_load": (C.c_int, [P_VOID, P_VOID]),
    "hdsdp_dev_scale_obj": (C.c_int, [P_VOID, C.c_double]),
    "hdsdp_dev_destroy": (None, [C.POINTER(P_VOID)]),
    "hdsdp_dev_update_buffer": (C.c_int, [P_VOID, C.c_double, C.c_double, P_DBL, C.c_double, C.c_int]),
    "hdsdp_dev_factorize": (C.c_int, [P_VOID, C.c_int, P_INT]),
    "hdsdp_dev_axpy_check": (C.c_int, [P_VOID, C.c_double, C.c_int, P_INT]),
    "hdsdp_dev_logdet": (C.c_int, [P_VOID, C.c_int, P_DBL]),
    "hdsdp_dev_ratio_test": (C.c_int, [P_VOID, C.c_int, P_DBL]),
    "hdsdp_dev_schur": (C.c_int, [P_VOID, C.c_int, C.c_double, P_DBL, P_DBL, P_DBL]),
    "hdsdp_dev_hsd": (C.c_int, [P_VOID, C.c_double, P_DBL, P_DBL, P_DBL, P_DBL]),
    "hdsdp_dev_kkt_uses_ldl": (C.c_int, [P_VOID]),
    "hdsdp_dev_kkt_params": (C.c_int, [P_VOID, C.c_double, C.c_double, C.c_int, C.c_int]),
    "hdsdp_dev_kkt_add_diag": (C.c_int, [P_VOID, P_DBL]),
    "hdsdp_dev_kkt_regularize": (C.c_int, [P_VOID, C.c_double]),
    "hdsdp_dev_kkt_factorize": (C.c_int, [P_VOID]),
    "hdsdp_dev_kkt_solve": (C.c_int, [P_VOID, P_DBL, P_DBL]),
    "hdsdp_dev_kkt_solve_exact": (C.c_int, [P_VOID, P_DBL, P_DBL]),
    "hdsdp_dev_get_schur": (C.c_int, [P_VOID, P_DBL]),
    "hdsdp_dev_kkt_solve_noalloc": (C.c_int, [P_VOID, C.c_int, P_VOID, P_VOID, C.c_double, C.c_double, C.c_int,
                                              C.c_int, P_VOID, P_INT]),
    "hdsdp_dev_primal": (C.c_int, [P_VOID, C.c_double, P_DBL, P_DBL, P_INT]),
    "hdsdp_dev_primal_stats": (C.c_int, [P_VOID, P_DBL, P_DBL, P_DBL, P_DBL]),
    "hdsdp_dev_get_primal": (C.c_int, [P_VOID, P_DBL]),
    "hdsdp_dev_get_dual": (C.c_int, [P_VOID, P_DBL]),
    "hdsdp_dev_adjoint": (C.c_int, [P_VOID, C.c_double, P_DBL, P_DBL, P_DBL]),
    "hdsdp_dev_psdp_init": (C.c_int, [P_VOID]),
    "hdsdp_dev_psdp_rhs": (C.c_int, [P_VOID, C.c_double, P_DBL, P_DBL]),
    "hdsdp_dev_psdp_direction": (C.c_int, [P_VOID, C.c_double, C.c_int, P_DBL]),
    "hdsdp_dev_psdp_step": (C.c_int, [P_VOID, C.c_double, P_INT, P_DBL, P_DBL]),
    "hdsdp_dev_stats": (C.c_int, [P_VOID, P_DBL, C.c_int]),
    "hdsdp_dev_set_profiling": (C.c_int, [P_VOID, C.c_int]),
    "hdsdp_dev_get_profile": (C.c_int, [P_VOID, P_DBL, P_DBL, C.c_int]),
    "hdsdp_dev_backend_name": (C.c_char_p, []),
}

# Device-time profile slots (HDSDP_PROF_* in include/hdsdp_gpu.h)
PROF_SLOTS = ("schur", "pairs", "mfactor", "cg", "sfactor", "lanczos", "buffer", "primal", "gram", "gram16")

# Entry points every build of the library exports (include/hdsdp.h, include/hdsdp_gpu.h).
EXPORTED = tuple(_API)


class CoeffView(C.Structure):
    """hdsdp_coeff_view (include/hdsdp_gpu.h)."""

    _fields_ = [
        ("type", C.c_int),
        ("nnz", C.c_int),
        ("row", P_INT),
        ("col", P_INT),
        ("val", P_DBL),
        ("sign", C.c_double),
    ]


def bind(lib: C.CDLL) -> C.CDLL:
    for name, (res, args) in _API.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def dptr(a):
    return a.ctypes.data_as(P_DBL) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(P_INT) if a is not None else None
