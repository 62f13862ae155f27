"""Device ABI parity: the HIP library against the CPU oracle and numpy closed
forms on identical inputs (buffers, Cholesky/PSD check, log-det, Schur
assembly, CG, Lanczos step length, primal recovery).  Tolerances: FP64 results
within 1e-10 relative of the oracle; Lanczos step lengths within 1e-8."""
import numpy as np
import pytest

import devtools as dt

pytestmark = pytest.mark.gpu

CASES = {
    # name: (n, m, kinds)
    "spr1_unit": (37, 45, [dt.SPR1]),
    "sparse_small": (41, 70, [dt.SPARSE]),
    "mixed_sparse": (66, 90, [dt.SPARSE, dt.SPR1, "SPARSE_BIG", "SPR1_BIG"]),
    "dense": (30, 40, [dt.DENSE]),
    "dsr1": (35, 50, [dt.DSR1]),
    "all_kinds_blocks": (130, 150, [dt.SPARSE, dt.SPR1, dt.DENSE, dt.DSR1, "SPARSE_BIG", dt.ZERO]),
    # single-entry constraints (edges, diagonal entries): the LDS pair kernel, with a
    # ragged last block of Schur columns (m not a multiple of 8) and, in the second
    # case, heavy rows and empty constraints among them
    "single_entry": (45, 121, ["SPARSE_ONE"]),
    "single_entry_heavy": (50, 150, ["SPARSE_ONE", "SPARSE_ONE", "SPARSE_BIG", dt.ZERO]),
}


def make_pair(gpu_lib, oracle, name, seed=0):
    n, m, kinds = CASES[name]
    rng = np.random.default_rng(seed)
    coeffs = dt.rand_coeffs(rng, n, m, kinds)
    return coeffs, dt.Dev(gpu_lib, coeffs), dt.Dev(oracle, coeffs), rng


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("name", list(CASES))
def test_buffer_factor_logdet(gpu_lib, oracle, name):
    coeffs, g, o, rng = make_pair(gpu_lib, oracle, name)
    y, eye = dt.interior_point(coeffs, rng)
    for d in (g, o):
        d.update(1.0, -1.0, y, eye, 0)
        assert d.factorize(0) == 1
    S = dt.dense_buffer(coeffs, 1.0, -1.0, y, eye)
    ld = np.linalg.slogdet(S)[1]
    assert abs(g.logdet(0) - ld) <= 1e-10 * max(1.0, abs(ld))
    assert abs(g.logdet(0) - o.logdet(0)) <= 1e-11 * max(1.0, abs(ld))
    # not positive definite: shift far below the smallest eigenvalue
    lam = np.linalg.eigvalsh(S).min()
    for d in (g, o):
        d.update(1.0, -1.0, y, eye - lam - 1.0, 1)
        assert d.factorize(1) == 0


@pytest.mark.parametrize("name", list(CASES))
def test_schur_matches_oracle_and_closed_form(gpu_lib, oracle, name):
    coeffs, g, o, rng = make_pair(gpu_lib, oracle, name, seed=1)
    y, eye = dt.interior_point(coeffs, rng)
    Rd = -0.37
    res = []
    for d in (g, o):
        d.update(1.0, -1.0, y, eye, 0)
        assert d.factorize(0) == 1
        a, b, t = d.schur(0, Rd)
        res.append((a, b, t, d.get_schur()))
    (ga, gb, gt, gM), (oa, ob, ot, oM) = res
    S = dt.dense_buffer(coeffs, 1.0, -1.0, y, eye)
    W = np.linalg.inv(S)
    A = [c.dense() for c in coeffs[1:]]
    M = np.array([[np.trace(Ai @ W @ Aj @ W) for Aj in A] for Ai in A])
    asinv = np.array([np.sum(Ai * W) for Ai in A])
    asrd = Rd * np.array([np.sum(Ai * (W @ W)) for Ai in A])
    assert rel(gM, M) < 1e-9
    assert rel(gM, oM) < 1e-10
    assert rel(ga, asinv) < 1e-9 and rel(ga, oa) < 1e-10
    assert rel(gb, asrd) < 1e-9 and rel(gb, ob) < 1e-10
    assert abs(gt - np.trace(W)) <= 1e-10 * abs(np.trace(W))
    # corrector: vectors only, M untouched
    for d in (g, o):
        a2, b2, t2 = d.schur(1, Rd)
        assert t2 == 0.0
    assert rel(g.get_schur(), gM) == 0.0


@pytest.mark.parametrize("name", list(CASES))
def test_hsd_terms_match_oracle_and_closed_form(gpu_lib, oracle, name):
    """The self-dual embedding's terms of a HOMOGENEOUS build (B = W C W):
    <A_i, B>, <C, W>, <C, B> and Rd tr(B) against numpy and the oracle."""
    coeffs, g, o, rng = make_pair(gpu_lib, oracle, name, seed=5)
    y, eye = dt.interior_point(coeffs, rng)
    Rd = -0.37
    res = []
    for d in (g, o):
        d.update(1.0, -1.0, y, eye, 0)
        assert d.factorize(0) == 1
        d.schur(2, Rd)
        res.append(d.hsd(Rd))
    (ga, gcs, gcsc, grd), (oa, ocs, ocsc, ord_) = res
    W = np.linalg.inv(dt.dense_buffer(coeffs, 1.0, -1.0, y, eye))
    Cm = coeffs[0].dense()
    B = W @ Cm @ W
    A = [c.dense() for c in coeffs[1:]]
    assert rel(ga, [np.sum(Ai * B) for Ai in A]) < 1e-9 and rel(ga, oa) < 1e-10
    for got, orc, ref in ((gcs, ocs, np.sum(Cm * W)), (gcsc, ocsc, np.sum(Cm * B)), (grd, ord_, Rd * np.trace(B))):
        scale = max(1.0, abs(ref), np.abs(Cm).max() * np.abs(W).sum())
        assert abs(got - ref) <= 1e-10 * scale and abs(got - orc) <= 1e-11 * scale


# (a zero constraint leaves M a zero row: "zero_diag" runs on the dense case only)
@pytest.mark.parametrize("name,mode", [("dense", "shift"), ("dense", "zero_diag"), ("all_kinds_blocks", "shift")])
def test_indefinite_schur_switches_to_lu(gpu_lib, oracle, name, mode):
    """The indefinite fallback: after the Cholesky switch fails on an indefinite M the
    device solves by LU with partial pivoting, the oracle by Bunch-Kaufman LDL^T; both
    match numpy's solve of that M, and each other."""
    coeffs, g, o, rng = make_pair(gpu_lib, oracle, name, seed=7)
    y, eye = dt.interior_point(coeffs, rng)
    res = []
    for d in (g, o):
        r = np.random.default_rng(11)
        res.append(dt.indefinite_fallback_run(d, coeffs, y, eye, r, mode))
    (gs, gm, gf0, gf1, _), (os_, om, of0, of1, _) = res
    assert (gf0, gf1) == (0, 1) and (of0, of1) == (0, 1)
    for (gx, rhs), (ox, _), A in zip(gs, os_, gm):
        ref = np.linalg.solve(A, rhs)
        bound = 1e-15 * np.linalg.cond(A) * 100
        assert np.linalg.norm(gx - ref) / np.linalg.norm(ref) <= bound
        assert np.linalg.norm(gx - ox) / np.linalg.norm(ox) <= 2 * bound


def test_indefinite_schur_switch_in_mixed_precision(gpu_lib):
    """The same switch on a mixed-precision device (dense constraints, whose Schur
    blocks come from bf16 split products): the LU solve runs on the FP64 M."""
    n, m, kinds = CASES["dense"]
    rng = np.random.default_rng(7)
    coeffs = dt.rand_coeffs(rng, n, m, kinds)
    d = dt.Dev(gpu_lib, coeffs, precision=1)
    y, eye = dt.interior_point(coeffs, rng)
    sols, mats, f0, f1, Mread = dt.indefinite_fallback_run(d, coeffs, y, eye, np.random.default_rng(11), "shift")
    assert (f0, f1) == (0, 1)
    d64 = dt.Dev(gpu_lib, coeffs, precision=0)
    d64.update(1.0, -1.0, y, eye, 0)
    assert d64.factorize(0) == 1
    d64.schur(0, 0.0)
    M64 = d64.get_schur()
    M64 = np.tril(M64) + np.tril(M64, -1).T
    for (x, rhs), A in zip(sols, mats):
        A = A - Mread + M64  # the same diagonal changes on the FP64 Schur matrix
        ref = np.linalg.solve(A, rhs)
        assert np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 1e-15 * np.linalg.cond(A) * 100


@pytest.mark.parametrize("name", ["spr1_unit", "mixed_sparse", "all_kinds_blocks"])
def test_cg_solve(gpu_lib, oracle, name):
    coeffs, g, o, rng = make_pair(gpu_lib, oracle, name, seed=2)
    y, eye = dt.interior_point(coeffs, rng)
    m = len(coeffs) - 1
    diag = rng.uniform(0.1, 1.0, m)
    rhs = rng.standard_normal(m)
    sols = []
    for d in (g, o):
        d.update(1.0, -1.0, y, eye, 0)
        assert d.factorize(0) == 1
        d.schur(0, 0.0)
        d.add_diag(diag)
        d.regularize(1e-6)
        d.kkt_params(5e-12, 1e-12)
        assert d.kkt_factorize() == 0
        sols.append(d.kkt_solve(rhs))
    M = g.get_schur()
    x = np.linalg.solve(M, rhs)
    assert rel(sols[0], x) < 1e-7
    assert rel(sols[0], sols[1]) < 1e-7


@pytest.mark.parametrize("name", list(CASES))
def test_ratio_test_lanczos(gpu_lib, oracle, name):
    coeffs, g, o, rng = make_pair(gpu_lib, oracle, name, seed=3)
    y, eye = dt.interior_point(coeffs, rng)
    dy = rng.standard_normal(len(coeffs) - 1)
    steps = []
    for d in (g, o):
        d.update(1.0, -1.0, y, eye, 0)
        assert d.factorize(0) == 1
        d.update(0.0, -1.0, dy, 0.3, 2)
        steps.append([d.ratio(0), d.ratio(0)])  # second call warm-starts
    assert rel(steps[0], steps[1]) < 1e-8
    S = dt.dense_buffer(coeffs, 1.0, -1.0, y, eye)
    dS = dt.dense_buffer(coeffs, 0.0, -1.0, dy, 0.3)
    L = np.linalg.cholesky(S)
    Li = np.linalg.inv(L)
    lam = np.linalg.eigvalsh(Li @ (-dS) @ Li.T).max()
    if lam > 0:
        # the safeguarded Lanczos bound never oversteps the boundary S + a dS = 0
        assert steps[0][0] <= 1.0 / lam * (1 + 1e-8)
        assert np.linalg.eigvalsh(S + 0.999 * steps[0][0] * dS).min() > 0


@pytest.mark.parametrize("name", ["spr1_unit", "all_kinds_blocks"])
def test_axpy_and_primal_recovery(gpu_lib, oracle, name):
    coeffs, g, o, rng = make_pair(gpu_lib, oracle, name, seed=4)
    y, eye = dt.interior_point(coeffs, rng)
    m = len(coeffs) - 1
    dy = 0.01 * rng.standard_normal(m)
    out = []
    for d in (g, o):
        d.update(1.0, -1.0, y, eye, 0)
        assert d.factorize(0) == 1
        d.update(0.0, -1.0, dy, 0.0, 2)
        ok = d.axpy_check(0.5, 1)
        ld = d.logdet(1)
        d.update(1.0, -1.0, y, eye, 1)  # checker = S for the recovery
        assert d.primal(0.7, y, dy) == 1
        out.append((ok, ld, d.get_primal(), d.primal_stats()))
    (gok, gld, gX, gst), (ook, old, oX, ost) = out
    assert gok == ook == 1
    assert abs(gld - old) <= 1e-11 * max(1.0, abs(old))
    S = dt.dense_buffer(coeffs, 1.0, -1.0, y, eye)
    W = np.linalg.inv(S)
    dS = dt.dense_buffer(coeffs, 0.0, 1.0, dy, 0.0)
    X = 0.7 * (W + W @ dS @ W)
    assert rel(gX, X) < 1e-9 and rel(gX, oX) < 1e-10
    AX = np.array([np.sum(c.dense() * X) for c in coeffs[1:]])
    assert rel(gst[0], AX) < 1e-9
    assert abs(gst[1] - np.sum(coeffs[0].dense() * X)) <= 1e-9 * max(1.0, np.abs(X).sum())
    assert abs(gst[3] - np.trace(X)) <= 1e-10 * max(1.0, abs(np.trace(X)))


@pytest.mark.parametrize("name", ["dense", "dsr1"])
def test_schur_mixed_precision_gram(gpu_lib, name, monkeypatch):
    """Mixed precision builds the dense and rank-one Schur matrices from bf16 split
    products (three terms per operand, FP32-equivalent): within 1e-6 relative of the
    FP64 closed form (the two-term split reaches only ~3e-6); the right-hand sides
    stay FP64.  (Rank-one Gram blocks take the bf16 path only for r >= 32 n unless
    forced, as here.)"""
    monkeypatch.setenv("HDSDP_GRAM_BF16", "1")
    n, m, kinds = CASES[name]
    rng = np.random.default_rng(5)
    coeffs = dt.rand_coeffs(rng, n, m, kinds)
    y, eye = dt.interior_point(coeffs, rng)
    d = dt.Dev(gpu_lib, coeffs, precision=1)
    d.update(1.0, -1.0, y, eye, 0)
    assert d.factorize(0) == 1
    a, b, t = d.schur(0, -0.5)
    M = d.get_schur()
    W = np.linalg.inv(dt.dense_buffer(coeffs, 1.0, -1.0, y, eye))
    A = [c.dense() for c in coeffs[1:]]
    Mref = np.array([[np.sum(Ai * (W @ Aj @ W)) for Aj in A] for Ai in A])
    assert rel(M, Mref) < 1e-6
    assert rel(a, [np.sum(Ai * W) for Ai in A]) < 1e-10


@pytest.mark.parametrize("planes", ["1", "0"])
@pytest.mark.parametrize("n,m", [(260, 300), (130, 700), (37, 90)])
def test_dense_schur_mixed_planes(gpu_lib, n, m, planes, monkeypatch):
    """The mixed-precision dense Schur matrix from bf16 planes (L^-T and A_p split by the
    producing kernels, V_p = A_p L^-T written as planes by the product's epilogue, the
    packed T_p padded to a multiple of 32 with zeros): several 128-wide tiles with ragged
    edges (n = 260), the split-K Gram in parts plus a ragged last part (m = 700), and n not
    a multiple of 8 (n = 37, padded planes).  Within 1e-6 relative of the FP64 closed form
    and within 1e-6 of the FP64 device build.  Both mixed kernels: the planes (split once by
    the producers) and the products that split their FP32 operands in the loop."""
    monkeypatch.setenv("HDSDP_BF16_PLANES", planes)
    rng = np.random.default_rng(19)
    coeffs = dt.rand_coeffs(rng, n, m, [dt.DENSE])
    y, eye = dt.interior_point(coeffs, rng)
    out = []
    for prec in (1, 0):
        d = dt.Dev(gpu_lib, coeffs, precision=prec)
        d.update(1.0, -1.0, y, eye, 0)
        assert d.factorize(0) == 1
        d.schur(0, 0.0)
        out.append(d.get_schur())
        d.close()
    Mx, M64 = out
    W = np.linalg.inv(dt.dense_buffer(coeffs, 1.0, -1.0, y, eye))
    L = np.linalg.cholesky(W)
    T = np.stack([(L.T @ c.dense() @ L).ravel() for c in coeffs[1:]])
    Mref = T @ T.T
    print(n, m, "mixed vs closed form", rel(Mx, Mref), "vs fp64", rel(Mx, M64), flush=True)
    assert rel(Mx, Mref) < 1e-6
    assert rel(Mx, M64) < 1e-6


@pytest.mark.parametrize("n,m", [(260, 300), (130, 700)])
def test_dense_schur_large_tiles(gpu_lib, n, m):
    """Dense Schur matrices large enough for the 128-wide product kernels: the
    triangular T = L^-1 A L^-T products (16-byte loads, skipped zero k ranges, edge
    tiles) at n = 260, and the split-K Gram over lower tiles only (XCD-balanced
    launch) at m = 700.  Against the closed form within 1e-9 relative."""
    rng = np.random.default_rng(9)
    coeffs = dt.rand_coeffs(rng, n, m, [dt.DENSE])
    y, eye = dt.interior_point(coeffs, rng)
    g = dt.Dev(gpu_lib, coeffs)
    g.update(1.0, -1.0, y, eye, 0)
    assert g.factorize(0) == 1
    a, b, t = g.schur(0, 0.0)
    M = g.get_schur()
    W = np.linalg.inv(dt.dense_buffer(coeffs, 1.0, -1.0, y, eye))
    L = np.linalg.cholesky(W)
    T = np.stack([(L.T @ c.dense() @ L).ravel() for c in coeffs[1:]])
    Mref = T @ T.T
    assert rel(M, Mref) < 1e-9
    assert rel(a, [np.sum(c.dense() * W) for c in coeffs[1:]]) < 1e-9


def test_rank_one_buffer_split_k(gpu_lib):
    """The rank-m buffer update F diag(w) F^T of m = 4000 dense rank-one constraints at
    n = 300 runs as a k-split product (three k parts of 1344, 1344 and 1312 in one
    launch) and an in-order reduction of the partials: S = C - A^T y + eye I against
    the dense sum within 1e-11 relative."""
    rng = np.random.default_rng(12)
    coeffs = dt.rand_coeffs(rng, 300, 4000, [dt.DSR1])
    y = 0.1 * rng.standard_normal(4000)
    g = dt.Dev(gpu_lib, coeffs)
    try:
        g.update(1.0, -1.0, y, 0.7, 0)
        S = g.get_dual()
    finally:
        g.close()
    Sref = dt.dense_buffer(coeffs, 1.0, -1.0, y, 0.7)
    assert rel(S, Sref) < 1e-11


@pytest.mark.parametrize("m", [150, 700])
def test_cg_solve_lower_triangle_product(gpu_lib, oracle, m, monkeypatch):
    """CG on the one-pass product with M that reads only its lower triangle (the path
    of m >= 16384, forced here for every size): 256-row by 128-column tiles, ragged at
    both edges at m = 700, their partials added in a fixed order.  The solution against
    numpy's solve of the same M and against the oracle's CG, within 1e-7."""
    monkeypatch.setenv("HDSDP_SYMV_LOWER_MIN", "1")
    rng = np.random.default_rng(13)
    coeffs = dt.rand_coeffs(rng, 50, m, [dt.SPARSE, dt.SPR1, dt.DSR1])
    y, eye = dt.interior_point(coeffs, rng)
    diag = rng.uniform(0.1, 1.0, m)
    rhs = rng.standard_normal(m)
    g, o = dt.Dev(gpu_lib, coeffs), dt.Dev(oracle, coeffs)
    sols = []
    try:
        for d in (g, o):
            d.update(1.0, -1.0, y, eye, 0)
            assert d.factorize(0) == 1
            d.schur(0, 0.0)
            d.add_diag(diag)
            d.regularize(1e-6)
            d.kkt_params(5e-12, 1e-12)
            assert d.kkt_factorize() == 0
            sols.append(d.kkt_solve(rhs))
        M = g.get_schur()
    finally:
        g.close()
        o.close()
    x = np.linalg.solve(M, rhs)
    assert rel(sols[0], x) < 1e-7
    assert rel(sols[0], sols[1]) < 1e-7


@pytest.mark.parametrize("n", [257, 300, 520, 777])
def test_blocked_cholesky_look_ahead(gpu_lib, n):
    """The blocked Cholesky with look-ahead (n > 4 panels of 64): trailing updates of
    two panels per pass, odd and even panel counts, a ragged last panel.  log det S
    against numpy's slogdet within 1e-10 relative, and a matrix shifted below its
    smallest eigenvalue reported not positive definite."""
    rng = np.random.default_rng(n)
    coeffs = dt.rand_coeffs(rng, n, 12, [dt.DENSE])
    y, eye = dt.interior_point(coeffs, rng)
    g = dt.Dev(gpu_lib, coeffs)
    try:
        g.update(1.0, -1.0, y, eye, 0)
        assert g.factorize(0) == 1
        S = dt.dense_buffer(coeffs, 1.0, -1.0, y, eye)
        ld = np.linalg.slogdet(S)[1]
        assert abs(g.logdet(0) - ld) <= 1e-10 * max(1.0, abs(ld))
        lam = np.linalg.eigvalsh(S).min()
        g.update(1.0, -1.0, y, eye - lam - 1.0, 1)
        assert g.factorize(1) == 0
    finally:
        g.close()
