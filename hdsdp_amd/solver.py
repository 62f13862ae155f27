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


class Solver:
    """One solver handle on ``inst`` through the C API of ``lib`` (create, set the
    cone, parameters; ``close`` destroys it).  The arrays handed to the C side are
    kept alive here: the user data stores the caller's pointers without copying
    (``interface/hdsdp_user_data.c:60-71``)."""

    def __init__(self, lib: C.CDLL, inst: SDPInstance, *, psdp: int = 0, relopt: float | None = None,
                 absopt: float | None = None, maxiter: int | None = None, timelimit: float | None = None,
                 precision: int = 0, verbose: bool = False, profile: bool = False):
        self.lib, self.inst = lib, inst
        self.m, self.n = m, n = inst.m, inst.n
        self.profile = profile
        self.h = C.c_void_p()
        self.ud = C.c_void_p()
        beg32 = np.ascontiguousarray(inst.beg, dtype=np.int32)
        idx = np.ascontiguousarray(inst.idx, dtype=np.int32)
        elem = np.ascontiguousarray(inst.elem, dtype=np.float64)
        b = np.ascontiguousarray(inst.b, dtype=np.float64)
        self._arrays = [beg32, idx, elem, b]
        if lib.HUserDataCreate(C.byref(self.ud)):
            raise RuntimeError("HUserDataCreate failed")
        if inst.factor is not None:
            fac = np.asfortranarray(inst.factor, dtype=np.float64)
            sgn = np.ascontiguousarray(inst.sign, dtype=np.float64)
            self._arrays += [fac, sgn]
            nnz_c = int(inst.beg[1])
            lib.HUserDataSetRankOneConeData(self.ud, m, n, nnz_c, _capi.iptr(idx), _capi.dptr(elem),
                                            fac.ctypes.data_as(_capi.P_DBL), _capi.dptr(sgn))
        else:
            if inst.nnz > np.iinfo(np.int32).max:
                lib.HUserDataDestroy(C.byref(self.ud))
                raise ValueError("instance exceeds the int32 CSC interface")
            lib.HUserDataSetConeData(self.ud, 4, m, n, _capi.iptr(beg32), _capi.iptr(idx), _capi.dptr(elem))
        if lib.HDSDPCreate(C.byref(self.h)):
            lib.HUserDataDestroy(C.byref(self.ud))
            raise RuntimeError("HDSDPCreate failed")
        try:
            if lib.HDSDPInit(self.h, m, 1):
                raise RuntimeError("HDSDPInit failed")
            lib.HDSDPSetVerbose(self.h, 1 if verbose else 0)
            if lib.HDSDPSetCone(self.h, 0, self.ud):
                raise RuntimeError("HDSDPSetCone failed")
            lib.HDSDPSetDualObjective(self.h, _capi.dptr(b))
            lib.HDSDPSetIntParamAny(self.h, INT_PARAM["psdp"], int(psdp))
            if relopt is not None:
                lib.HDSDPSetDblParam(self.h, DBL_PARAM["relopt"], float(relopt))
            if absopt is not None:
                lib.HDSDPSetDblParam(self.h, DBL_PARAM["absopt"], float(absopt))
            if maxiter is not None:
                lib.HDSDPSetIntParam(self.h, INT_PARAM["maxiter"], int(maxiter))
            if timelimit is not None:
                lib.HDSDPSetDblParam(self.h, DBL_PARAM["timelimit"], float(timelimit))
            lib.HDSDPSetPrecision(self.h, int(precision))
            lib.HDSDPSetProfiling(self.h, 1 if profile else 0)
        except BaseException:
            self.close()
            raise

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if self.h:
            self.lib.HDSDPDestroy(C.byref(self.h))
            self.h = C.c_void_p()
        if self.ud:
            self.lib.HUserDataDestroy(C.byref(self.ud))
            self.ud = C.c_void_p()

    def optimize(self, gap_target: float = 1e-6) -> SolveResult:
        lib, m = self.lib, self.m
        rc = lib.HDSDPOptimize(self.h, 1)
        y = np.zeros(m)
        pobj, dobj = C.c_double(), C.c_double()
        lib.HDSDPGetRowDual(self.h, C.byref(pobj), C.byref(dobj), _capi.dptr(y))
        st, it = C.c_int(), C.c_int()
        mu, secs, secs_gap = C.c_double(), C.c_double(), C.c_double()
        lib.HDSDPGetSummary(self.h, gap_target, C.byref(st), C.byref(it), C.byref(mu),
                            C.byref(secs), C.byref(secs_gap))
        nlog = lib.HDSDPGetIterLog(self.h, 0, None)
        log = np.zeros((max(nlog, 0), 9))
        if nlog > 0:
            lib.HDSDPGetIterLog(self.h, nlog, _capi.dptr(log))
        dimacs = np.zeros(6)
        lib.HDSDPGetDimacsErrors(self.h, _capi.dptr(dimacs))
        prof = None
        if self.profile:
            ns = len(_capi.PROF_SLOTS)
            ms, cnt = np.zeros(ns), np.zeros(ns)
            lib.HDSDPGetProfile(self.h, _capi.dptr(ms), _capi.dptr(cnt), ns)
            prof = {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(_capi.PROF_SLOTS)}
        return SolveResult(int(rc), int(st.value), int(it.value), y, pobj.value, dobj.value, mu.value,
                           secs.value, secs_gap.value, dimacs, log, prof)

    def profile_now(self) -> dict:
        """Device-time profile so far: slot -> (ms, timed regions); needs profile=True."""
        ns = len(_capi.PROF_SLOTS)
        ms, cnt = np.zeros(ns), np.zeros(ns)
        self.lib.HDSDPGetProfile(self.h, _capi.dptr(ms), _capi.dptr(cnt), ns)
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(_capi.PROF_SLOTS)}

    def setup(self) -> None:
        """Pre-solve and device set-up without the interior-point iterations."""
        if self.lib.HDSDPSetup(self.h):
            raise RuntimeError("HDSDPSetup failed")

    def barrier_point(self) -> tuple[float, np.ndarray]:
        """(mu, y) of the last optimize, in the units of the data."""
        mu, y = C.c_double(), np.zeros(self.m)
        if self.lib.HDSDPGetBarrierPoint(self.h, C.byref(mu), _capi.dptr(y)):
            raise RuntimeError("no barrier point (optimize first)")
        return mu.value, y

    def center(self, mu: float, y0: np.ndarray, maxiter: int = 60, tol: float = 1e-13):
        """Newton-centre y0 at mu: returns (y_mu, Newton steps, relative residual)."""
        y = np.ascontiguousarray(y0, dtype=np.float64).copy()
        nit, res = C.c_int(), C.c_double()
        rc = self.lib.HDSDPCenter(self.h, float(mu), _capi.dptr(y), int(maxiter), float(tol), C.byref(nit),
                                  C.byref(res))
        if rc:
            raise RuntimeError(f"HDSDPCenter failed after {nit.value} steps (residual {res.value:.3e})")
        return y, nit.value, res.value

    def adjoint(self, mu: float, y: np.ndarray, dLdy: np.ndarray):
        """(dL/db, dL/dC) of a loss of the centred y_mu from dL/dy."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        g = np.ascontiguousarray(dLdy, dtype=np.float64)
        db = np.zeros(self.m)
        dC = np.zeros((self.n, self.n), order="F")
        if self.lib.HDSDPAdjoint(self.h, float(mu), _capi.dptr(y), _capi.dptr(g), _capi.dptr(db),
                                 dC.ctypes.data_as(_capi.P_DBL)):
            raise RuntimeError("HDSDPAdjoint failed")
        return db, np.ascontiguousarray(dC)


def run_solver(lib: C.CDLL, inst: SDPInstance, *, gap_target: float = 1e-6, **opts) -> SolveResult:
    """Solve ``inst`` through the C API of ``lib`` (options as for ``Solver``)."""
    with Solver(lib, inst, **opts) as sv:
        return sv.optimize(gap_target)


def solve(inst: SDPInstance, **kw) -> SolveResult:
    """Solve on the MI355X through the HIP library."""
    return run_solver(load_library(), inst, **kw)
