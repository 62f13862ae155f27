"""Drive the hdsdp_dev_* ABI (include/hdsdp_gpu.h) from numpy for tests.

Builds random presolved coefficient sets of every kind, feeds the same inputs
to any library exporting the ABI (the HIP product or the CPU oracle build), and
provides dense numpy restatements for closed-form checks.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from hdsdp_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "libhdsdp_oracle.so")
REF_RUNNER = os.path.join(ROOT, "oracle", "_ref", "hdsdp_ref_run")
GOLDEN = os.path.join(ROOT, "tests", "golden")

ZERO, SPARSE, DENSE, SPR1, DSR1 = 0, 1, 2, 3, 4

_libs = {}


def oracle_lib():
    if "oracle" not in _libs:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle library not built: make -f oracle/Makefile")
        _libs["oracle"] = _capi.bind(C.CDLL(ORACLE_LIB, mode=C.RTLD_LOCAL))
    return _libs["oracle"]


class Coeff:
    """One presolved coefficient matrix (numpy side)."""

    def __init__(self, kind, n, row=None, col=None, val=None, sign=0.0):
        self.kind, self.n = kind, n
        self.row = None if row is None else np.ascontiguousarray(row, dtype=np.int32)
        self.col = None if col is None else np.ascontiguousarray(col, dtype=np.int32)
        self.val = None if val is None else np.ascontiguousarray(val, dtype=np.float64)
        self.sign = float(sign)

    @property
    def nnz(self):
        if self.kind in (SPARSE, SPR1):
            return len(self.val)
        if self.kind == DSR1:
            return self.n
        if self.kind == DENSE:
            return self.n * (self.n + 1) // 2
        return 0

    def dense(self) -> np.ndarray:
        n = self.n
        A = np.zeros((n, n))
        if self.kind == SPARSE:
            A[self.row, self.col] += self.val
            off = self.row != self.col
            A[self.col[off], self.row[off]] += self.val[off]
        elif self.kind == DENSE:
            p = 0
            for j in range(n):
                A[j:, j] = self.val[p:p + n - j]
                A[j, j:] = self.val[p:p + n - j]
                p += n - j
        elif self.kind == SPR1:
            a = np.zeros(n)
            a[self.row] = self.val
            A = self.sign * np.outer(a, a)
        elif self.kind == DSR1:
            A = self.sign * np.outer(self.val, self.val)
        return A

    def view(self) -> _capi.CoeffView:
        v = _capi.CoeffView()
        v.type = self.kind
        v.nnz = self.nnz if self.kind in (SPARSE, SPR1) else (self.n if self.kind == DSR1 else 0)
        v.row = _capi.iptr(self.row)
        v.col = _capi.iptr(self.col)
        v.val = _capi.dptr(self.val)
        v.sign = self.sign
        return v


def rand_sparse(rng, n, nnz):
    pos = rng.choice(n * (n + 1) // 2, size=nnz, replace=False)
    iu, ju = np.tril_indices(n)  # row-major enumeration of the lower triangle
    rows, cols = iu[pos], ju[pos]
    order = np.lexsort((rows, cols))  # column-major order
    return Coeff(SPARSE, n, rows[order], cols[order], rng.standard_normal(nnz)[order])


def rand_dense(rng, n):
    return Coeff(DENSE, n, val=rng.standard_normal(n * (n + 1) // 2) / np.sqrt(n))


def rand_spr1(rng, n, nnz):
    idx = np.sort(rng.choice(n, size=nnz, replace=False))
    a = rng.standard_normal(nnz)
    a /= np.linalg.norm(a)
    return Coeff(SPR1, n, row=idx, val=a, sign=rng.uniform(0.5, 2.0) * rng.choice([-1, 1]))


def rand_dsr1(rng, n):
    a = rng.standard_normal(n)
    a /= np.linalg.norm(a)
    return Coeff(DSR1, n, val=a, sign=rng.uniform(0.5, 2.0) * rng.choice([-1, 1]))


def rand_coeffs(rng, n, m, kinds):
    """C (sparse) plus m constraints cycling through `kinds`."""
    out = [rand_sparse(rng, n, max(1, n))]
    for i in range(m):
        k = kinds[i % len(kinds)]
        if k == SPARSE:
            out.append(rand_sparse(rng, n, int(rng.integers(1, 4))))
        elif k == "SPARSE_ONE":
            out.append(rand_sparse(rng, n, 1))
        elif k == "SPARSE_BIG":
            out.append(rand_sparse(rng, n, int(rng.integers(6, 3 * n))))
        elif k == DENSE:
            out.append(rand_dense(rng, n))
        elif k == SPR1:
            out.append(rand_spr1(rng, n, int(rng.integers(1, 4))))
        elif k == "SPR1_BIG":
            out.append(rand_spr1(rng, n, min(n, 6)))
        elif k == DSR1:
            out.append(rand_dsr1(rng, n))
        elif k == ZERO:
            out.append(Coeff(ZERO, n))
    return out


class Dev:
    """One hdsdp_dev object on a given library."""

    def __init__(self, lib, coeffs, precision=0):
        self.lib = lib
        self.coeffs = coeffs
        self.n = coeffs[0].n
        self.m = len(coeffs) - 1
        self.h = C.c_void_p()
        rc = lib.hdsdp_dev_create(C.byref(self.h), self.m, self.n, precision)
        if rc:
            raise RuntimeError(f"hdsdp_dev_create failed ({rc})")
        views = (_capi.CoeffView * (self.m + 1))(*[c.view() for c in coeffs])
        if lib.hdsdp_dev_load(self.h, C.cast(views, C.c_void_p)):
            raise RuntimeError("hdsdp_dev_load failed")

    def close(self):
        if self.h:
            self.lib.hdsdp_dev_destroy(C.byref(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ok(self, rc, what):
        if rc:
            raise RuntimeError(f"{what} failed ({rc})")

    def update(self, cCoef, aScal, aCoef, eye, which):
        a = np.ascontiguousarray(aCoef, dtype=np.float64)
        self._ok(self.lib.hdsdp_dev_update_buffer(self.h, cCoef, aScal, _capi.dptr(a), eye, which), "update")

    def factorize(self, which):
        ok = C.c_int()
        self._ok(self.lib.hdsdp_dev_factorize(self.h, which, C.byref(ok)), "factorize")
        return ok.value

    def axpy_check(self, step, which):
        ok = C.c_int()
        self._ok(self.lib.hdsdp_dev_axpy_check(self.h, step, which, C.byref(ok)), "axpy_check")
        return ok.value

    def logdet(self, which):
        v = C.c_double()
        self._ok(self.lib.hdsdp_dev_logdet(self.h, which, C.byref(v)), "logdet")
        return v.value

    def ratio(self, which):
        v = C.c_double(-1.0)
        self._ok(self.lib.hdsdp_dev_ratio_test(self.h, which, C.byref(v)), "ratio_test")
        return v.value

    def schur(self, typeKKT, Rd):
        a = np.zeros(self.m)
        b = np.zeros(self.m)
        t = C.c_double()
        self._ok(self.lib.hdsdp_dev_schur(self.h, typeKKT, Rd, _capi.dptr(a), _capi.dptr(b), C.byref(t)), "schur")
        return a, b, t.value

    def hsd(self, Rd):
        a = np.zeros(self.m)
        cs, csc, csrd = C.c_double(), C.c_double(), C.c_double()
        self._ok(self.lib.hdsdp_dev_hsd(self.h, Rd, _capi.dptr(a), C.byref(cs), C.byref(csc), C.byref(csrd)), "hsd")
        return a, cs.value, csc.value, csrd.value

    def get_schur(self):
        M = np.zeros((self.m, self.m), order="F")
        self._ok(self.lib.hdsdp_dev_get_schur(self.h, M.ctypes.data_as(_capi.P_DBL)), "get_schur")
        return M

    def kkt_params(self, rel, absol, maxit=-1, restart=-1):
        self.lib.hdsdp_dev_kkt_params(self.h, rel, absol, maxit, restart)

    def add_diag(self, d):
        d = np.ascontiguousarray(d, dtype=np.float64)
        self._ok(self.lib.hdsdp_dev_kkt_add_diag(self.h, _capi.dptr(d)), "add_diag")

    def regularize(self, r):
        self._ok(self.lib.hdsdp_dev_kkt_regularize(self.h, r), "regularize")

    def kkt_factorize(self):
        return self.lib.hdsdp_dev_kkt_factorize(self.h)

    def kkt_solve(self, rhs):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        sol = np.zeros(self.m)
        rc = self.lib.hdsdp_dev_kkt_solve(self.h, _capi.dptr(rhs), _capi.dptr(sol))
        self._ok(rc, "kkt_solve")
        return sol

    def uses_ldl(self):
        return self.lib.hdsdp_dev_kkt_uses_ldl(self.h)

    def primal(self, mu, y, dy):
        y = np.ascontiguousarray(y, dtype=np.float64)
        dy = np.ascontiguousarray(dy, dtype=np.float64)
        ok = C.c_int()
        self._ok(self.lib.hdsdp_dev_primal(self.h, mu, _capi.dptr(y), _capi.dptr(dy), C.byref(ok)), "primal")
        return ok.value

    def primal_stats(self):
        AX = np.zeros(self.m)
        cx, xs, tx = C.c_double(), C.c_double(), C.c_double()
        self._ok(self.lib.hdsdp_dev_primal_stats(self.h, _capi.dptr(AX), C.byref(cx), C.byref(xs), C.byref(tx)),
                 "primal_stats")
        return AX, cx.value, xs.value, tx.value

    def get_primal(self):
        X = np.zeros((self.n, self.n), order="F")
        self._ok(self.lib.hdsdp_dev_get_primal(self.h, X.ctypes.data_as(_capi.P_DBL)), "get_primal")
        return X

    def get_dual(self):
        S = np.zeros((self.n, self.n), order="F")
        self._ok(self.lib.hdsdp_dev_get_dual(self.h, S.ctypes.data_as(_capi.P_DBL)), "get_dual")
        return S

    def adjoint(self, mu, dLdy):
        dLdy = np.ascontiguousarray(dLdy, dtype=np.float64)
        db = np.zeros(self.m)
        dC = np.zeros((self.n, self.n), order="F")
        self._ok(self.lib.hdsdp_dev_adjoint(self.h, mu, _capi.dptr(dLdy), _capi.dptr(db),
                                            dC.ctypes.data_as(_capi.P_DBL)), "adjoint")
        return db, dC


def dense_buffer(coeffs, cCoef, aScal, aCoef, eye):
    n = coeffs[0].n
    S = cCoef * coeffs[0].dense() + eye * np.eye(n)
    for i, c in enumerate(coeffs[1:]):
        if aCoef[i] != 0.0:
            S = S + aScal * aCoef[i] * c.dense()
    return S


def interior_point(coeffs, rng, shift=None):
    """(y, S) with S = C - A^T y + shift I comfortably positive definite."""
    n, m = coeffs[0].n, len(coeffs) - 1
    y = 0.1 * rng.standard_normal(m)
    S = dense_buffer(coeffs, 1.0, -1.0, y, 0.0)
    lam = np.linalg.eigvalsh(S).min()
    eye = (1.0 - lam) if shift is None else shift
    return y, eye


# ---- golden fixtures ------------------------------------------------------------

def load_golden(name):
    import json

    import golden.make_golden as mg

    with open(os.path.join(GOLDEN, name + ".json")) as f:
        fix = json.load(f)
    make, opts, _ = mg.CASES[name]
    inst = make()
    assert mg.instance_digest(inst) == fix["digest"], "instance generator drifted from the fixture"
    fix["y"] = np.load(os.path.join(GOLDEN, name + "_y.npy"))
    if "dense" in fix:
        fix["dense"]["y"] = np.load(os.path.join(GOLDEN, name + "_dense_y.npy"))
    return inst, fix


def solve_like_golden(lib, name, **kw):
    from hdsdp_amd.solver import run_solver

    inst, fix = load_golden(name)
    opts = dict(fix["options"])
    opts.setdefault("psdp", 0)
    opts.update(kw)
    res = run_solver(lib, inst, **opts)
    return inst, fix, res


def parity_report(fix, res):
    y0 = fix["y"]
    return {
        "status": (res.status, fix["status"]),
        "iters": (res.iters, fix["iters"]),
        "dobj_rel": abs(res.dobj - fix["dobj"]) / max(1.0, abs(fix["dobj"])),
        "pobj_rel": abs(res.pobj - fix["pobj"]) / max(1.0, abs(fix["pobj"])),
        "y_rel": float(np.linalg.norm(res.y - y0) / max(np.linalg.norm(y0), 1e-300)),
    }


# The parity bar: 1e-8 relative, or the reference's own spread on the instance where that
# is wider (the largest distance between reference runs that differ only in the order of
# their arithmetic: thread count, MKL instruction-set path, PARDISO or LAPACK for the
# dual matrix -- recorded in each fixture by make_golden.py).
PARITY_FLOOR = 1e-8


def golden_tolerances(fix):
    sp = fix["spread"]
    return {k: max(PARITY_FLOOR, sp[k]) for k in ("y", "pobj", "dobj")}


def check_dimacs(fix, res):
    """The six DIMACS errors of a solve against the reference's on the same instance
    (interface/hdsdp.c:753-906): errors 2-4 (lambda_max(X) sign, dual infeasibility, the
    fixed zero) equal, error 1 (||AX - b||) at rounding level, and errors 5 and 6 (gap and
    <X, S> over |pObj| + |dObj| + 1) within the objectives' parity tolerances, which bound
    how far the final iterate may sit from the reference's."""
    tol = golden_tolerances(fix)
    ref, got = np.asarray(fix["dimacs"]), np.asarray(res.dimacs)
    print(fix["name"], "dimacs ref", ref, "got", got, flush=True)
    assert got[1] == ref[1] and got[2] == ref[2] and got[3] == ref[3]
    assert got[0] <= max(1e-12, 10.0 * ref[0])
    for k in (4, 5):
        assert abs(got[k] - ref[k]) <= tol["pobj"] + tol["dobj"], (k, got[k], ref[k])


def check_parity(fix, res, label=""):
    """Status and iteration count equal to the reference's default build (and to its
    dense-dual build where the fixture has one, unless the reference's own spread runs
    reach this count as well); y*, pObj and dObj within golden_tolerances of both.
    Returns the reports."""
    tol = golden_tolerances(fix)
    reps = {"default": parity_report(fix, res)}
    if "dense" in fix:
        reps["dense"] = parity_report(fix["dense"], res)
    print(label or fix["name"], {"tol": tol, **reps}, flush=True)
    # iteration counts the reference itself reaches on this input (its own spread runs)
    ref_iters = {fix["iters"]} | {r["iters"] for r in fix["spread"]["runs"]}
    for which, rep in reps.items():
        ref = fix if which == "default" else fix["dense"]
        assert res.status == ref["status"], (which, rep)
        if which == "default":
            assert res.iters == ref["iters"], (which, rep)
        else:
            assert res.iters == ref["iters"] or res.iters in ref_iters, (which, rep, sorted(ref_iters))
        assert rep["dobj_rel"] <= tol["dobj"], (which, rep, tol)
        assert rep["pobj_rel"] <= tol["pobj"], (which, rep, tol)
        assert rep["y_rel"] <= tol["y"], (which, rep, tol)
    ref = [(int(r[0]), int(r[1])) for r in fix["log"]]
    got = [(int(r[0]), int(r[1])) for r in res.log]
    assert got == ref
    return reps


def adjoint_fd_check(lib, inst, n_dirs, seed=0, h=1e-6, relopt=1e-8, noise=False, **solve_opts):
    """hdsdp_amd.adjoint.fd_check, returned as (relative errors, mu, dL/db, dL/dC, y_mu)
    and, with noise, the finite-difference error estimates per direction as well."""
    from hdsdp_amd.adjoint import fd_check

    r = fd_check(lib, inst, n_dirs, seed=seed, h=h, relopt=relopt, noise=noise, **solve_opts)
    if noise:
        return r["errs"], r["mu"], r["dLdb"], r["dLdC"], r["y"], r["fd_noise"]
    return r["errs"], r["mu"], r["dLdb"], r["dLdC"], r["y"]


def indefinite_fallback_run(d, coeffs, y, eye, rng, mode):
    """Drive one device's Schur system into the indefinite solver and back out of it:
    M is made indefinite through the diagonal (mode "shift": M - s I with s between two
    middle eigenvalues; mode "zero_diag": the diagonal removed, which leaves Bunch-Kaufman
    only 2 x 2 pivots at first), a Jacobi CG capped at three steps fails, its Cholesky
    switch fails, and the system switches to the direct solver for good.  Returns the
    solutions of three right-hand sides (the last after a second factorisation with a
    changed diagonal), the matrices they solve, the switch flag before and after, and the
    Schur matrix as the device built it."""
    m = len(coeffs) - 1
    d.update(1.0, -1.0, y, eye, 0)
    assert d.factorize(0) == 1
    d.schur(0, 0.0)
    M = d.get_schur()
    M = np.tril(M) + np.tril(M, -1).T
    if mode == "shift":
        lam = np.linalg.eigvalsh(M)
        k = m // 2
        delta = -0.5 * (lam[k] + lam[k + 1]) * np.ones(m)
    else:
        delta = -np.diag(M).copy()
    d.add_diag(delta)
    A1 = M + np.diag(delta)
    d.kkt_params(1e-14, 1e-14, 3)
    flag0 = d.uses_ldl()
    assert d.kkt_factorize() == 0
    sols, mats = [], []
    for k in range(2):
        rhs = rng.standard_normal(m)
        sols.append((d.kkt_solve(rhs), rhs))
        mats.append(A1)
    delta2 = rng.uniform(-0.01, 0.01, m) * np.abs(np.diag(A1)).max()
    d.add_diag(delta2)
    A2 = A1 + np.diag(delta2)
    assert d.kkt_factorize() == 0
    rhs = rng.standard_normal(m)
    sols.append((d.kkt_solve(rhs), rhs))
    mats.append(A2)
    return sols, mats, flag0, d.uses_ldl(), M
