"""Mixed precision (FP32 Cholesky of the Schur matrix as the preconditioner of
the FP64 CG) against the plain FP64 GPU path on the same instances.  The CG
still stops on FP64 residuals at the reference's tolerances, so the iterates
agree to the solve tolerance: same status and iteration count, objectives
within 1e-8 relative and y* within 1e-7 relative at the default optimality
tolerance."""
import numpy as np
import pytest

import devtools as dt
from hdsdp_amd import instances as I
from hdsdp_amd.solver import run_solver

pytestmark = pytest.mark.gpu

CASES = {
    "maxcut_n200": lambda: I.maxcut(200, 0.05, 1, "pm1"),
    "theta_n50": lambda: I.lovasz_theta(50, 0.3, 21),
    "dense_n40_m60": lambda: I.dense_random(40, 60, 24),
    "rankone_n60_m500": lambda: I.rank_one(60, 500, 5, 45),
}


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_matches_fp64(gpu_lib, name):
    inst = CASES[name]()
    r64 = run_solver(gpu_lib, inst, precision=0)
    rmx = run_solver(gpu_lib, inst, precision=1)
    rel_y = np.linalg.norm(rmx.y - r64.y) / np.linalg.norm(r64.y)
    print(name, r64.status, r64.iters, rmx.status, rmx.iters, rel_y, flush=True)
    assert rmx.status == r64.status == 3
    assert rmx.iters == r64.iters
    assert abs(rmx.dobj - r64.dobj) <= 1e-8 * max(1.0, abs(r64.dobj))
    assert abs(rmx.pobj - r64.pobj) <= 1e-8 * max(1.0, abs(r64.pobj))
    assert rel_y <= 1e-7
