"""Python driver for the solver C API.

``solve(inst, ...)`` runs the MI355X library (``hdsdp_amd/lib/libhdsdp_amd.so``);
it raises if the library is missing or was not built for the GPU.
``run_solver(lib, inst, ...)`` drives any library that exports the API in
include/hdsdp.h.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

from . import _capi
from .instances import SDPInstance

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libhdsdp_amd.so")

STATUS_NAMES = [
    "UNKNOWN", "DUAL_FEASIBLE", "DUAL_OPTIMAL", "PRIMAL_DUAL_OPTIMAL", "MAXITER",
    "SUSPECT_INFEAS_OR_UNBOUNDED", "INFEAS_OR_UNBOUNDED", "TIMELIMIT", "USER_INTERRUPT",
    "INTERNAL_ERROR", "NUMERICAL",
]

INT_PARAM = {"maxiter": 0, "corra": 1, "corrb": 2, "threads": 3, "psdp": 4}
DBL_PARAM = {"relfeas": 0, "relopt": 1, "absfeas": 2, "absopt": 3, "timelimit": 4}

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the product library; fail loudly when it is absent."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library {path} is not built (run __graft_entry__.build())")
    lib = _capi.bind(C.CDLL(path, mode=C.RTLD_LOCAL))
    name = lib.hdsdp_dev_backend_name().decode()
    if not name.startswith("hip"):
        raise RuntimeError(f"{path} is not the HIP build (backend {name})")
    if path == LIB_PATH:
        _lib = lib
    return lib


@dataclasses.dataclass
class SolveResult:
    retcode: int
    status: int
    iters: int
    y: np.ndarray
    pobj: float
    dobj: float
    mu: float
    seconds: float
    seconds_to_gap: float
    dimacs: np.ndarray
    log: np.ndarray  # records x 9: phase, iter, pObj, dObj, inf, mu, step, prox, t
    profile: dict | None = None  # slot -> (device ms, timed regions) when profiling

    @property
    def status_name(self) -> str:
        return STATUS_NAMES[self.status] if 0 <= self.status < len(STATUS_NAMES) else str(self.status)

    @property
    def rel_gap(self) -> float:
        return (self.pobj - self.dobj) / (abs(self.pobj) + abs(self.dobj) + 1.0)


def run_solver(lib: C.CDLL, inst: SDPInstance, *, psdp: int = 0, relopt: float | None = None,
               absopt: float | None = None, maxiter: int | None = None,
               timelimit: float | None = None, precision: int = 0, verbose: bool = False,
               gap_target: float = 1e-6, profile: bool = False, keep=None) -> SolveResult:
    """Solve ``inst`` through the C API of ``lib``.  ``keep`` (a list) receives the
    solver handle instead of destroying it, for callers that query it further."""
    m, n = inst.m, inst.n
    solver = C.c_void_p()
    ud = C.c_void_p()
    beg32 = np.ascontiguousarray(inst.beg, dtype=np.int32)
    idx = np.ascontiguousarray(inst.idx, dtype=np.int32)
    elem = np.ascontiguousarray(inst.elem, dtype=np.float64)
    b = np.ascontiguousarray(inst.b, dtype=np.float64)
    arrays = [beg32, idx, elem, b]
    if lib.HUserDataCreate(C.byref(ud)):
        raise RuntimeError("HUserDataCreate failed")
    if inst.factor is not None:
        fac = np.asfortranarray(inst.factor, dtype=np.float64)
        sgn = np.ascontiguousarray(inst.sign, dtype=np.float64)
        arrays += [fac, sgn]
        nnz_c = int(inst.beg[1])
        lib.HUserDataSetRankOneConeData(ud, m, n, nnz_c, _capi.iptr(idx), _capi.dptr(elem),
                                        fac.ctypes.data_as(_capi.P_DBL), _capi.dptr(sgn))
    else:
        if inst.nnz > np.iinfo(np.int32).max:
            raise ValueError("instance exceeds the int32 CSC interface")
        lib.HUserDataSetConeData(ud, 4, m, n, _capi.iptr(beg32), _capi.iptr(idx), _capi.dptr(elem))
    if lib.HDSDPCreate(C.byref(solver)):
        raise RuntimeError("HDSDPCreate failed")
    try:
        if lib.HDSDPInit(solver, m, 1):
            raise RuntimeError("HDSDPInit failed")
        lib.HDSDPSetVerbose(solver, 1 if verbose else 0)
        if lib.HDSDPSetCone(solver, 0, ud):
            raise RuntimeError("HDSDPSetCone failed")
        lib.HDSDPSetDualObjective(solver, _capi.dptr(b))
        lib.HDSDPSetIntParamAny(solver, INT_PARAM["psdp"], int(psdp))
        if relopt is not None:
            lib.HDSDPSetDblParam(solver, DBL_PARAM["relopt"], float(relopt))
        if absopt is not None:
            lib.HDSDPSetDblParam(solver, DBL_PARAM["absopt"], float(absopt))
        if maxiter is not None:
            lib.HDSDPSetIntParam(solver, INT_PARAM["maxiter"], int(maxiter))
        if timelimit is not None:
            lib.HDSDPSetDblParam(solver, DBL_PARAM["timelimit"], float(timelimit))
        lib.HDSDPSetPrecision(solver, int(precision))
        lib.HDSDPSetProfiling(solver, 1 if profile else 0)
        rc = lib.HDSDPOptimize(solver, 1)
        y = np.zeros(m)
        pobj, dobj = C.c_double(), C.c_double()
        lib.HDSDPGetRowDual(solver, C.byref(pobj), C.byref(dobj), _capi.dptr(y))
        st, it = C.c_int(), C.c_int()
        mu, secs, secs_gap = C.c_double(), C.c_double(), C.c_double()
        lib.HDSDPGetSummary(solver, gap_target, C.byref(st), C.byref(it), C.byref(mu),
                            C.byref(secs), C.byref(secs_gap))
        nlog = lib.HDSDPGetIterLog(solver, 0, None)
        log = np.zeros((max(nlog, 0), 9))
        if nlog > 0:
            lib.HDSDPGetIterLog(solver, nlog, _capi.dptr(log))
        dimacs = np.zeros(6)
        lib.HDSDPGetDimacsErrors(solver, _capi.dptr(dimacs))
        prof = None
        if profile:
            ns = len(_capi.PROF_SLOTS)
            ms, cnt = np.zeros(ns), np.zeros(ns)
            lib.HDSDPGetProfile(solver, _capi.dptr(ms), _capi.dptr(cnt), ns)
            prof = {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(_capi.PROF_SLOTS)}
        res = SolveResult(int(rc), int(st.value), int(it.value), y, pobj.value, dobj.value, mu.value,
                          secs.value, secs_gap.value, dimacs, log, prof)
        if keep is not None:
            keep.append((solver, ud, arrays))
            solver = None
        return res
    finally:
        if solver is not None:
            lib.HDSDPDestroy(C.byref(solver))
            lib.HUserDataDestroy(C.byref(ud))


def solve(inst: SDPInstance, **kw) -> SolveResult:
    """Solve on the MI355X through the HIP library."""
    return run_solver(load_library(), inst, **kw)
