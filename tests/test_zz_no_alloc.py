"""hdsdp_dev_kkt_solve_noalloc: the Jacobi-preconditioned Schur CG on a caller-held
Schur matrix, with its vectors in the library's static device array instead of
allocated memory.  It must return exactly what hdsdp_dev_kkt_solve returns for the
same matrix and right-hand side (the same kernels in the same order), on two calls in
a row on one GPU, the second with a smaller system (the shared array keeps the first
call's contents), and the same solution as the oracle's host restatement to rounding."""
import ctypes as C

import numpy as np
import pytest
import torch

import devtools as dt

pytestmark = pytest.mark.gpu

SIG = (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int,
                 C.c_void_p, C.POINTER(C.c_int)])


def bind_noalloc(lib):
    fn = lib.hdsdp_dev_kkt_solve_noalloc
    fn.restype, fn.argtypes = SIG
    return fn


def schur_system(lib, n, m, seed):
    """A Schur matrix built by the device path from random coefficients, made diagonally
    dominant through the bound-cone diagonal so that the Jacobi CG converges."""
    rng = np.random.default_rng(seed)
    coeffs = dt.rand_coeffs(rng, n, m, [dt.SPARSE, dt.SPR1, dt.DENSE, dt.DSR1])
    y, eye = dt.interior_point(coeffs, rng)
    d = dt.Dev(lib, coeffs)
    d.update(1.0, -1.0, y, eye, 0)
    assert d.factorize(0) == 1
    d.schur(0, 0.0)
    M0 = d.get_schur()
    d.add_diag(np.full(m, 2.0 * np.abs(M0).sum(axis=0).max()))
    rhs = rng.standard_normal(m)
    return d, rhs


@pytest.mark.parametrize("seed", [11])
def test_noalloc_matches_kkt_solve(gpu_lib, oracle, seed):
    fn = bind_noalloc(gpu_lib)
    ofn = bind_noalloc(oracle)
    rel, absol, maxit, restart = 1e-10, 1e-10, 200, 20
    for call, (n, m) in enumerate([(40, 300), (24, 90)]):
        print(f"[noalloc] call {call}: n={n} m={m} seed={seed + call}", flush=True)
        d, rhs = schur_system(gpu_lib, n, m, seed + call)
        try:
            d.kkt_params(rel, absol, maxit, restart)
            assert d.kkt_factorize() == 0
            ref = d.kkt_solve(rhs)
            M = d.get_schur()
        finally:
            d.close()
        # the device path kept the Jacobi preconditioner (no Cholesky of M was built)
        Md = torch.tensor(M, dtype=torch.float64, device="cuda")  # symmetric: either order
        bd = torch.tensor(rhs, dtype=torch.float64, device="cuda")
        xd = torch.full((m,), np.nan, dtype=torch.float64, device="cuda")
        it = C.c_int(-1)
        stream = torch.cuda.current_stream()
        torch.cuda.synchronize()
        rc = fn(Md.data_ptr(), m, bd.data_ptr(), xd.data_ptr(), rel, absol, maxit, restart,
                C.c_void_p(stream.cuda_stream), C.byref(it))
        torch.cuda.synchronize()
        assert rc == 0, rc
        x = xd.cpu().numpy()
        print(f"[noalloc] call {call}: {it.value} CG iterations, max |x - ref| {np.max(np.abs(x - ref)):.3e}",
              flush=True)
        assert np.array_equal(x, ref)
        # the oracle's restatement of the same iteration on the host
        xo = np.zeros(m)
        ito = C.c_int(-1)
        Mf = np.asfortranarray(M)
        assert ofn(Mf.ctypes.data, m, rhs.ctypes.data, xo.ctypes.data, rel, absol, maxit, restart, None,
                   C.byref(ito)) == 0
        assert ito.value == it.value
        assert np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo))
        # and it solves the system
        assert np.linalg.norm(M @ x - rhs) <= 1e-8 * np.linalg.norm(rhs)


def test_noalloc_rejects_oversized_system(gpu_lib):
    fn = bind_noalloc(gpu_lib)
    it = C.c_int(-1)
    # refused on the host before any device access: no pointer is dereferenced
    assert fn(None, 65537, None, None, 1e-8, 1e-8, 10, 20, None, C.byref(it)) != 0
    assert fn(None, 0, None, None, 1e-8, 1e-8, 10, 20, None, C.byref(it)) != 0
