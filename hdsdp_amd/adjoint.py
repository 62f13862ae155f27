"""Differentiation through the SDP solve.

The differentiated map is the centred dual point at the final barrier parameter:
for data ``(b, C)`` (the constraint matrices ``A_i`` fixed) and ``mu > 0``,
``y_mu(b, C)`` maximises ``b^T y + mu log det(C - sum_i y_i A_i)``.  The forward
pass solves the SDP on the GPU, takes ``mu`` and ``y`` of the last interior-point
iterate and Newton-centres ``y`` at that ``mu`` (``HDSDPCenter``), so the output
is a smooth function of ``(b, C)``.  The backward pass calls ``HDSDPAdjoint``:
with the Schur matrix ``M`` of the centred point, ``M lambda = dL/dy``,
``dL/db = lambda / mu`` and ``dL/dC = S^-1 (sum_i lambda_i A_i) S^-1``.

``C`` enters as a dense symmetric tensor; the forward pass uses ``(C + C^T) / 2``,
so its gradient is the symmetric ``dL/dC``.
"""
from __future__ import annotations

import numpy as np
import torch

from . import instances as I
from .solver import Solver, load_library


def centred_solve(lib, inst: I.SDPInstance, *, tol: float = 1e-13, maxiter: int = 60, **solve_opts):
    """Solve ``inst``, then centre the final iterate at its barrier parameter.
    Returns ``(solver, mu, y_mu, solve result)``; the caller closes ``solver``."""
    sv = Solver(lib, inst, **solve_opts)
    try:
        res = sv.optimize()
        mu, y0 = sv.barrier_point()
        y, _, _ = sv.center(mu, y0, maxiter=maxiter, tol=tol)
    except BaseException:
        sv.close()
        raise
    return sv, mu, y, res


def centre_at(lib, inst: I.SDPInstance, mu: float, y0: np.ndarray | None = None, *, tol: float = 1e-13,
              maxiter: int = 60, **solve_opts) -> np.ndarray:
    """``y_mu`` of ``inst`` at a given ``mu`` (for finite differences of the centred
    map).  From ``y0`` (S(y0) positive definite for this data) when given, else from
    the final iterate of an interior-point solve of ``inst``."""
    with Solver(lib, inst, **solve_opts) as sv:
        if y0 is None:
            sv.optimize()
            _, y0 = sv.barrier_point()
        else:
            sv.setup()
        y, _, _ = sv.center(mu, y0, maxiter=maxiter, tol=tol)
    return y


def fd_check(lib, inst: I.SDPInstance, n_dirs: int, *, seed: int = 0, h: float = 1e-6, relopt: float = 1e-8,
             log=None, noise: bool = False, **solve_opts):
    """Adjoint of the centred map against central finite differences.

    Solves ``inst`` to ``relopt``, centres at the final mu and takes the gradient of
    L = ||y_mu - y_t||^2 (y_t seeded normal) by HDSDPAdjoint.  Along ``n_dirs`` seeded
    random directions (db, dC), dC symmetric, it compares <dL/db, db> + <dL/dC, dC>
    with (L(+h) - L(-h)) / 2h, where L(+-h) centres a solve of the perturbed data at
    the same mu.  With ``noise`` the quotient is also taken with step 2h: its distance
    from the step-h quotient measures that direction's finite-difference error (the
    truncation term grows fourfold from h to 2h, the rounding term of the centred
    points halves).  ``log(k, err)`` is called after each direction.  Returns a dict:
    relative errors per direction, mu, the gradients, y_mu, the seconds of the adjoint
    call, and with ``noise`` the relative distances |fd(h) - fd(2h)| / |fd(h)|."""
    import time

    rng = np.random.default_rng(seed)
    opts = dict(relopt=relopt, absopt=1e30, **solve_opts)
    sv, mu, y, _ = centred_solve(lib, inst, **opts)
    try:
        yt = rng.standard_normal(inst.m)
        t0 = time.perf_counter()
        db, dC = sv.adjoint(mu, y, 2.0 * (y - yt))
        t_adj = time.perf_counter() - t0
    finally:
        sv.close()
    C0 = inst.dense_coeff(0)
    errs, noises = [], []

    def quotient(step, dbv, dCv):
        Lv = []
        for sgn in (1.0, -1.0):
            ip = I.with_data(inst, b=inst.b + sgn * step * dbv, C=C0 + sgn * step * dCv)
            yp = centre_at(lib, ip, mu, **opts)
            Lv.append(float(np.sum((yp - yt) ** 2)))
        return (Lv[0] - Lv[1]) / (2.0 * step)

    for _ in range(n_dirs):
        dbv = rng.standard_normal(inst.m)
        dCv = rng.standard_normal((inst.n, inst.n))
        dCv = 0.5 * (dCv + dCv.T)
        fd = quotient(h, dbv, dCv)
        an = float(db @ dbv + np.sum(dC * dCv))
        errs.append(abs(fd - an) / max(abs(fd), 1e-300))
        if noise:
            noises.append(abs(fd - quotient(2.0 * h, dbv, dCv)) / max(abs(fd), 1e-300))
        if log:
            log(len(errs), errs[-1])
    out = {"errs": errs, "mu": mu, "dLdb": db, "dLdC": dC, "y": y, "adjoint_s": t_adj}
    if noise:
        out["fd_noise"] = noises
    return out


class SDPDual(torch.autograd.Function):
    """``y = SDPDual.apply(b, C, inst, opts)``: the centred dual solution of
    ``inst`` with its right-hand side and objective replaced by ``b`` (m) and
    ``C`` (n x n), float64 tensors.  ``opts`` are solve options (``relopt``,
    ``precision``, ...)."""

    @staticmethod
    def forward(ctx, b, C, inst, opts=None):
        opts = dict(opts or {})
        lib = load_library()
        Cs = 0.5 * (C + C.transpose(0, 1))
        data = I.with_data(inst, b=b.detach().cpu().numpy(), C=Cs.detach().cpu().numpy())
        sv, mu, y, _ = centred_solve(lib, data, **opts)
        ctx.solver, ctx.mu, ctx.y, ctx.data = sv, mu, y, data
        return torch.as_tensor(y, dtype=b.dtype, device=b.device)

    @staticmethod
    def backward(ctx, gy):
        sv = ctx.solver
        try:
            db, dC = sv.adjoint(ctx.mu, ctx.y, gy.detach().cpu().numpy().astype(np.float64))
        finally:
            sv.close()
        return (torch.as_tensor(db, dtype=gy.dtype, device=gy.device),
                torch.as_tensor(dC, dtype=gy.dtype, device=gy.device), None, None)
