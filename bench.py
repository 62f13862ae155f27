"""Benchmark of the HDSDP dual-scaling interior-point path on one MI355X.

Workload: Lovasz theta SDP of the seeded Erdos-Renyi graph G(n=400, p=0.1),
seed 2 (n = 400, m = 8084 constraints: tr X = 1 and X_ij = 0 per edge), solved
by the plain FP64 GPU path to a 1e-6 relative gap.  One step is one complete
solve (HDSDPOptimize: presolve, Phase A, Phase B, primal check); ``value`` is
seconds per IPM iteration over the timed steps.

The JSON line also carries:
  * ``parity``: status, iterations and the error of y*, pObj and dObj against
    the reference HDSDP's FP64 result on the same instance (stored fixture);
  * ``roofline``: the Schur pair kernel (one launch per Schur build; the LDS
    variant for single-entry constraints, as on C2),
    HBM-bound: algorithmic bytes = 8 per Schur entry written (lower triangle
    of M over the sparse constraints) + the n x n FP64 S^-1 read once,
    divided by its average launch time from HIP events on the library's
    stream during the timed steps; ``traffic`` is the calibrated FETCH_SIZE +
    WRITE_SIZE of that kernel from profiles/theta400_schur_pairs_traffic.md;
  * ``phases``: device milliseconds per IPM iteration by phase;
  * ``cpu_baseline``: the reference HDSDP build (oracle/_ref, MKL) solving the
    same instance on the host cores, when that build is present;
  * ``adjoint`` (workload C5): the adjoint of the centred dual map after a
    1e-8-gap solve (HDSDPAdjoint time) and its relative error against central
    finite differences (h = 1e-6) along ``--adjoint-dirs`` seeded directions.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

REF_RUNNER = os.path.join(ROOT, "oracle", "_ref", "hdsdp_ref_run")
GOLDEN = os.path.join(ROOT, "tests", "golden")
PAIR_TRAFFIC = os.path.join(ROOT, "profiles", "theta400_schur_pairs_traffic.md")

# workload -> (description, instance constructor, reference fixture or None)
WORKLOADS = {
    "C1": ("MaxCut SDP, G(n=200, p=0.05) seed 1, weights +-1: n=200, m=200",
           lambda I: I.maxcut(200, 0.05, 1, "pm1"), "c1_gap1e-6"),
    "C2": ("Lovasz theta SDP, G(n=400, p=0.1) seed 2: n=400, m=8084",
           lambda I: I.lovasz_theta(400, 0.1, 2), "c2_gap1e-6"),
    "C3": ("dense random SDP, n=1000, m=2000, seed 3",
           lambda I: I.dense_random(1000, 2000, 3), None),
    "C4": ("rank-one SDP (factor form), n=2000, m=20000, r=20, seed 4",
           lambda I: I.rank_one(2000, 20000, 20, 4), None),
    "C5": ("MaxCut SDP, G(n=2000, p=0.01) seed 5, weights U(0,1): n=2000, m=2000",
           lambda I: I.maxcut(2000, 0.01, 5, "unif"), "c5_gap1e-6"),
}
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="C2")
    ap.add_argument("--maxiter", type=int, default=None, help="cap IPM iterations (timing samples)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--adjoint-dirs", type=int, default=20, help="C5: finite-difference directions (0: none)")
    return ap.parse_args()


def pair_traffic(workload):
    """Calibrated HBM bytes per launch of the pair kernel (C2 only), from the
    counter runs summarised in profiles/."""
    if workload != "C2" or not os.path.exists(PAIR_TRAFFIC):
        return None
    with open(PAIR_TRAFFIC) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def adjoint_report(lib, inst, ndirs):
    """Adjoint of the centred map after a 1e-8-gap FP64 solve, checked by central FD."""
    from hdsdp_amd.adjoint import fd_check

    t0 = time.perf_counter()
    r = fd_check(lib, inst, ndirs, seed=5, h=1e-6, relopt=1e-8,
                 log=lambda k, e: print(f"[bench] adjoint FD direction {k}: rel err {e:.2e}", file=sys.stderr,
                                        flush=True))
    errs = r["errs"]
    return {
        "map": "centred dual y_mu(b, C) at the final mu of a 1e-8-gap FP64 solve",
        "loss": "||y_mu - y_t||^2, y_t ~ N(0, 1) seeded",
        "mu": r["mu"],
        "adjoint_ms": r["adjoint_s"] * 1e3,
        "fd_h": 1e-6,
        "fd_directions": ndirs,
        "fd_max_rel_err": max(errs) if errs else None,
        "fd_median_rel_err": float(np.median(errs)) if errs else None,
        "check_s": time.perf_counter() - t0,
    }


def parity(res, fixture):
    if fixture is None:
        return None
    path = os.path.join(GOLDEN, fixture)
    with open(path + ".json") as f:
        fix = json.load(f)
    y0 = np.load(path + "_y.npy")
    return {
        "vs": "reference HDSDP FP64 (CPU) on the same instance",
        "status": [res.status, fix["status"]],
        "iters": [res.iters, fix["iters"]],
        "dobj_rel_err": abs(res.dobj - fix["dobj"]) / max(1.0, abs(fix["dobj"])),
        "pobj_rel_err": abs(res.pobj - fix["pobj"]) / max(1.0, abs(fix["pobj"])),
        "y_rel_err": float(np.linalg.norm(res.y - y0) / np.linalg.norm(y0)),
        "gap": res.rel_gap,
    }


def cpu_baseline(inst, threads):
    """Reference HDSDP (built from its sources into oracle/_ref) on the host:
    one full solve of the same instance to the same 1e-6 gap."""
    if not os.path.exists(REF_RUNNER):
        return None
    from hdsdp_amd import instances as I

    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "inst.bin")
        I.write_instance(inst, path)
        env = dict(os.environ, MKL_NUM_THREADS=str(threads), OMP_NUM_THREADS=str(threads))
        args = [REF_RUNNER, path, os.path.join(td, "y.bin"), "psdp=0", "relopt=1e-6", "absopt=1e30",
                f"threads={threads}"]
        r = subprocess.run(args, capture_output=True, text=True, timeout=600, env=env)
    if "RESULT " not in r.stdout:
        return {"error": (r.stdout + r.stderr)[-400:]}
    res = json.loads(r.stdout.split("RESULT ", 1)[1].splitlines()[0])
    if res.get("status", -1) < 0 or res["iters"] <= 0:
        return {"error": "reference solve failed"}
    return {
        "value": res["time"] / res["iters"],
        "unit": "s/iter",
        "cores": threads,
        "kind": "reference",
        "sample": f"one full solve of the same instance to 1e-6 gap ({res['iters']} iterations, "
                  f"{res['time']:.2f} s, MKL {threads} threads)",
        "solve_s": res["time"],
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the solver's "
              "Schur/Cholesky path runs on a single MI355X.", file=sys.stderr)
        sys.exit(2)

    import torch

    from hdsdp_amd import instances as I
    from hdsdp_amd.solver import Solver, load_library, run_solver

    lib = load_library()
    desc, make, fixture = WORKLOADS[args.workload]
    inst = make(I)
    prec = 1 if args.precision == "mixed" else 0
    opts = dict(relopt=1e-6, absopt=1e30, precision=prec, maxiter=args.maxiter)
    for _ in range(args.warmup):
        run_solver(lib, inst, **opts)
        print(f"[bench] warmup solve done", file=sys.stderr, flush=True)

    # Each step: the pre-solve (host CSC -> HBM upload, rank-one detection, scaling)
    # runs untimed, then HDSDPOptimize on the resident data is timed between two
    # device synchronisations; the upload-inclusive time is reported beside it.
    results = []
    elapsed = 0.0
    upload = 0.0
    for k in range(args.steps):
        with Solver(lib, inst, profile=True, **opts) as sv:
            torch.cuda.synchronize()
            tu = time.perf_counter()
            sv.setup()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            results.append(sv.optimize())
            torch.cuda.synchronize()
            t1 = time.perf_counter()
        upload += t0 - tu
        elapsed += t1 - t0
        print(f"[bench] step {k} done", file=sys.stderr, flush=True)

    iters = sum(r.iters for r in results)
    s_per_iter = elapsed / iters
    prof = {k: [sum(r.profile[k][0] for r in results), sum(r.profile[k][1] for r in results)]
            for k in results[0].profile}

    m, n = inst.m, inst.n
    # Schur pair kernel: 8 bytes per lower-triangle entry of M between the small
    # (at most 4 entries) sparse constraints, plus S^-1 read once
    if inst.factor is None:
        col_nnz = np.diff(inst.beg)[1:]
        n_small = int(np.count_nonzero(col_nnz <= 4))
    else:
        n_small = 0
    pair_bytes = 8.0 * n_small * (n_small + 1) / 2 + 8.0 * n * n
    # the library's kernel for this step: the LDS variant when every small constraint is
    # a single entry and n <= 480, else the gather kernel
    single = inst.factor is None and n_small > 0 and n <= 480 and bool(np.all(col_nnz[col_nnz <= 4] == 1))
    pair_kernel = "schur_pair_single_kernel" if single else "schur_pair_kernel"
    pair_ms, pair_cnt = prof["pairs"]
    pair_avg_s = pair_ms / 1e3 / max(pair_cnt, 1)
    achieved = pair_bytes / pair_avg_s / 1e9 if pair_cnt else 0.0
    mf_ms, mf_cnt = prof["mfactor"]
    mf_flops = 2.0 * m ** 3 / 3.0  # Cholesky (m^3/3) + triangular inverse (m^3/3)
    # algorithmic flops of one Schur build of the Gram formulations (dense: T_i = L^-1 A_i L^-T
    # by triangular products and the symmetric Gram of the T_i; rank-one: W F and F^T W F)
    sc_ms, sc_cnt = prof["schur"]
    if args.workload == "C3":
        schur_flops = 2.0 * m * n ** 3 + m * (m + 1.0) * n ** 2
    elif args.workload == "C4":
        schur_flops = 1.0 * n * n * m + m * (m + 1.0) * n
    else:
        schur_flops = None

    line = {
        "metric": "IPM s/iter + solve s to 1e-6 gap, mixed vs FP64 GPU; Schur TFLOP/s vs peak",
        "value": s_per_iter,
        "unit": "s/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if prec == 0 else "f64+f32",
        "data": "synthetic (seeded instance generator, numpy PCG64)",
        "config": {"workload": f"{desc}, {args.precision} GPU, gap 1e-6", "name": args.workload, "n": n, "m": m,
                   "precision": args.precision, "maxiter": args.maxiter},
        "iters_per_solve": results[0].iters,
        "solve_s": elapsed / args.steps,
        "presolve_upload_s": upload / args.steps,
        "s_per_iter_incl_upload": (elapsed + upload) / sum(r.iters for r in results),
        "solve_s_to_gap": float(np.mean([r.seconds_to_gap for r in results])),
        "status": results[-1].status,
        "dobj": results[-1].dobj,
        "pobj": results[-1].pobj,
        "parity": parity(results[-1], fixture if args.maxiter is None else None),
        "roofline": {
            "bound": "hbm",
            "kernel": pair_kernel,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": pair_traffic(args.workload),
            "bytes_per_launch": pair_bytes,
            "avg_launch_us": pair_avg_s * 1e6,
            "launches": pair_cnt,
        },
        "schur_factor": {
            "bound": "mfma",
            "achieved_tflops": (mf_flops * mf_cnt / (mf_ms / 1e3) / 1e12) if mf_ms > 0 else 0.0,
            "peak_tflops": FP64_PEAK_TFLOPS,
            "avg_ms": mf_ms / max(mf_cnt, 1),
            "count": mf_cnt,
        },
        "schur_build": None if schur_flops is None or sc_cnt == 0 else {
            "bound": "mfma",
            "algorithmic_flops": schur_flops,
            "avg_ms": sc_ms / sc_cnt,
            "achieved_tflops": schur_flops / (sc_ms / sc_cnt / 1e3) / 1e12,
            "peak_tflops": FP64_PEAK_TFLOPS,
            "count": sc_cnt,
        },
        "phases_ms_per_iter": {k: v[0] / iters for k, v in prof.items()},
    }
    if not args.no_cpu_baseline and inst.factor is None:
        line["cpu_baseline"] = cpu_baseline(inst, args.cpu_threads)
    if args.workload == "C5" and args.adjoint_dirs > 0 and prec == 0:
        print("[bench] adjoint check", file=sys.stderr, flush=True)
        line["adjoint"] = adjoint_report(lib, inst, args.adjoint_dirs)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
