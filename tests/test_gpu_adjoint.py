"""Adjoint of the centred dual map on the GPU.

The differentiated map is y_mu(b, C), the maximiser of b^T y + mu log det(C - A^T y)
at the final barrier parameter of a solve; HDSDPAdjoint solves
M lambda = dL/dy with the Schur matrix of the centred point and returns
dL/db = lambda / mu, dL/dC = S^-1 (sum lambda_i A_i) S^-1.  Checked against
  * central finite differences (h = 1e-6) of centred GPU solves of the perturbed
    data along seeded random directions, within 1e-5 relative, on solves to a
    1e-7 gap: at the final mu S is near-singular (cond ~ 1/mu) and the centred
    point carries an FP64 rounding floor of about eps cond(S), which the FD
    quotient divides by 2h; at a 1e-8 gap (mu ~ 3e-10) that floor alone reaches
    ~1e-5 on the dense instance (measured 2e-7 .. 1.1e-5 over directions);
  * the CPU oracle's adjoint at the same (mu, y), which pins the device kernels
    (Schur matrix, exact solve, the W B W products) against a plain C restatement;
  * the torch.autograd.Function wrapper, whose gradients must be those of
    HDSDPAdjoint.
MaxCut and dense instances are nondegenerate; Lovasz theta SDPs are not (y* is
not unique) and the centred map's derivatives there grow like 1/mu, beyond what a
1e-6 step resolves."""
import numpy as np
import pytest
import torch

import devtools as dt
from hdsdp_amd import instances as I
from hdsdp_amd.adjoint import SDPDual, centred_solve
from hdsdp_amd.solver import Solver

pytestmark = pytest.mark.gpu

FD_CASES = {
    # name: (instance, directions, gap of the solve, seed of the directions and y_t)
    "maxcut_n200": (lambda: I.maxcut(200, 0.05, 5, "unif"), 6, 1e-7, 0),
    "dense_n40_m60": (lambda: I.dense_random(40, 60, 24), 4, 1e-7, 0),
    # the adjoint workload at its own size: MaxCut on G(2000, 0.01), weights U(0, 1),
    # solved to a 1e-8 gap, 20 directions seeded with the workload's seed, 5 (the
    # benchmark's adjoint check draws the same ones)
    "maxcut_n2000_gap1e-8": (lambda: I.maxcut(2000, 0.01, 5, "unif"), 20, 1e-8, 5),
}


@pytest.mark.parametrize("name", list(FD_CASES))
def test_adjoint_matches_central_fd(gpu_lib, name):
    make, ndir, gap, seed = FD_CASES[name]
    errs, mu, db, dC, y = dt.adjoint_fd_check(gpu_lib, make(), ndir, seed=seed, relopt=gap)
    print(name, "mu", mu, "rel errs", errs, flush=True)
    assert max(errs) < 1e-5
    assert np.array_equal(dC, dC.T)


@pytest.mark.parametrize("name", ["maxcut_n60", "dense_n20_m30"])
def test_adjoint_matches_oracle_at_same_point(gpu_lib, oracle, name):
    inst = I.maxcut(60, 0.1, 7, "unif") if name == "maxcut_n60" else I.dense_random(20, 30, 22)
    sv, mu, y, _ = centred_solve(gpu_lib, inst, relopt=1e-8, absopt=1e30)
    g = np.random.default_rng(1).standard_normal(inst.m)
    try:
        db, dC = sv.adjoint(mu, y, g)
    finally:
        sv.close()
    with Solver(oracle, inst) as so:
        so.setup()
        db0, dC0 = so.adjoint(mu, y, g)
    e_b = np.linalg.norm(db - db0) / np.linalg.norm(db0)
    e_C = np.linalg.norm(dC - dC0) / np.linalg.norm(dC0)
    print(name, "mu", mu, "dLdb rel", e_b, "dLdC rel", e_C, flush=True)
    assert e_b < 1e-6 and e_C < 1e-6


def test_adjoint_reuses_the_centring_factor(gpu_lib, oracle):
    """HDSDPAdjoint right after the centring does not factor the Schur matrix again: the
    last Newton step's FP64 factor (of the Schur matrix one step before the centred point)
    preconditions iterative refinement on the centred point's M.  No new factorisation is
    counted, and the gradients match the oracle's (which factors afresh) within 1e-6."""
    inst = I.maxcut(60, 0.1, 7, "unif")
    g = np.random.default_rng(3).standard_normal(inst.m)
    sv = Solver(gpu_lib, inst, relopt=1e-8, absopt=1e30, profile=True)
    try:
        sv.optimize()
        mu, y0 = sv.barrier_point()
        y, _, _ = sv.center(mu, y0)
        before = sv.profile_now()["mfactor"][1]
        db, dC = sv.adjoint(mu, y, g)
        after = sv.profile_now()["mfactor"][1]
    finally:
        sv.close()
    with Solver(oracle, inst) as so:
        so.setup()
        db0, dC0 = so.adjoint(mu, y, g)
    e_b = np.linalg.norm(db - db0) / np.linalg.norm(db0)
    e_C = np.linalg.norm(dC - dC0) / np.linalg.norm(dC0)
    print("factorisations before/after the adjoint", before, after, "dLdb rel", e_b, "dLdC rel", e_C, flush=True)
    assert after == before
    assert e_b < 1e-6 and e_C < 1e-6


def test_autograd_function(gpu_lib):
    inst = I.maxcut(60, 0.1, 7, "unif")
    opts = dict(relopt=1e-8, absopt=1e30)
    b = torch.tensor(inst.b, dtype=torch.float64, requires_grad=True)
    C = torch.tensor(inst.dense_coeff(0), dtype=torch.float64, requires_grad=True)
    yt = torch.tensor(np.random.default_rng(2).standard_normal(inst.m))
    y = SDPDual.apply(b, C, inst, opts)
    loss = ((y - yt) ** 2).sum()
    loss.backward()
    sv, mu, y2, _ = centred_solve(gpu_lib, inst, **opts)
    try:
        db, dC = sv.adjoint(mu, y2, 2.0 * (y2 - yt.numpy()))
    finally:
        sv.close()
    assert np.allclose(y.detach().numpy(), y2, rtol=1e-10, atol=1e-12)
    assert np.allclose(b.grad.numpy(), db, rtol=1e-8, atol=1e-10 * np.abs(db).max())
    assert np.allclose(C.grad.numpy(), dC, rtol=1e-8, atol=1e-10 * np.abs(dC).max())
    assert torch.equal(C.grad, C.grad.T)
