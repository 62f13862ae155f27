"""Benchmark of the HDSDP dual-scaling interior-point path on one MI355X.

Default workload: the dense random SDP of n = 1000, m = 2000 (seed 3; every A_i a
dense symmetric N(0, 1/n) matrix, C = sum_i y0_i A_i + I, b = A(I)), solved to a
1e-6 relative gap by the plain FP64 GPU path and by the mixed-precision path
(bf16 split Schur products, FP32 Cholesky of M as the preconditioner of FP64
refinement on the exact Schur operator, FP64 everywhere else).  One step is one
complete solve (HDSDPOptimize on data already resident in HBM); with
``--precision both`` (the default) the steps alternate FP64, mixed, FP64, ...
``value`` is the mixed path's seconds per IPM iteration (the FP64 path's with
``--precision fp64``).

The JSON line also carries, per precision, s/iter, the seconds of a solve and the
seconds to the first iteration at a 1e-6 relative gap, the iteration count and
status; ``mixed_vs_fp64`` the FP64 / mixed ratios of those times and the error of
the mixed y*, pObj, dObj and gap against the FP64 run; ``schur_build`` the
algorithmic flops of one Schur assembly (dense: 2 m n^3 + m (m + 1) n^2) against
the FP64 peak (FP64 path) and the bf16 peak (mixed path, six bf16 products per
FP64 product); ``roofline`` the dominant kernel of the headline path;
``phases_ms_per_iter`` device milliseconds per IPM iteration by phase; ``parity``
(workloads with a stored reference result) the FP64 path against the reference
HDSDP; ``cpu_baseline`` the reference HDSDP build (oracle/_ref, MKL) on the host
cores, on a bounded sample of the workload's family.
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
PROFILES = os.path.join(ROOT, "profiles")

# workload -> (description, instance constructor, reference fixture or None,
#              CPU-baseline sample: (description, constructor, maxiter or None))
WORKLOADS = {
    "C1": ("MaxCut SDP, G(n=200, p=0.05) seed 1, weights +-1: n=200, m=200",
           lambda I: I.maxcut(200, 0.05, 1, "pm1"), "c1_gap1e-6",
           ("the same instance, full solve to 1e-6 gap", lambda I: I.maxcut(200, 0.05, 1, "pm1"), None)),
    "C2": ("Lovasz theta SDP, G(n=400, p=0.1) seed 2: n=400, m=8084",
           lambda I: I.lovasz_theta(400, 0.1, 2), "c2_gap1e-6",
           ("the same instance, full solve to 1e-6 gap", lambda I: I.lovasz_theta(400, 0.1, 2), None)),
    "C3": ("dense random SDP, n=1000, m=2000, seed 3",
           lambda I: I.dense_random(1000, 2000, 3), None,
           ("dense random SDP of the same family at n=120, m=240 (seed 33), full solve to 1e-6 gap "
            "(the n=1000 instance is 12 GB of CSC input and minutes per reference iteration; the "
            "reference aborts when an iteration cap ends its phase A, so no capped sample is taken)",
            lambda I: I.dense_random(120, 240, 33), None)),
    "C4": ("rank-one SDP (factor form), n=2000, m=20000, r=20, seed 4",
           lambda I: I.rank_one(2000, 20000, 20, 4), None,
           ("rank-one SDP of the same family at n=60, m=500 (seed 45, CSC input), full solve to 1e-6 gap "
            "(the n=2000, m=20000 instance overflows the reference's int32 CSC input)",
            lambda I: I.rank_one(60, 500, 5, 45, form="csc"), None)),
    "C5": ("MaxCut SDP, G(n=2000, p=0.01) seed 5, weights U(0,1): n=2000, m=2000",
           lambda I: I.maxcut(2000, 0.01, 5, "unif"), "c5_gap1e-6",
           ("the same instance, full solve to 1e-6 gap", lambda I: I.maxcut(2000, 0.01, 5, "unif"), None)),
}
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
BF16_PRODUCTS = 6  # bf16 MFMA products per FP64-equivalent product (three-term split)


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="C3")
    ap.add_argument("--maxiter", type=int, default=None, help="cap IPM iterations (timing samples)")
    ap.add_argument("--warmup-maxiter", type=int, default=3,
                    help="IPM iterations of each warm-up solve (compiles kernels, touches memory)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--adjoint-dirs", type=int, default=20, help="C5: finite-difference directions (0: none)")
    return ap.parse_args()


def schur_flops(workload, n, m):
    """Algorithmic flops of one Schur assembly of the Gram formulations."""
    if workload == "C3":
        return 2.0 * m * n ** 3 + m * (m + 1.0) * n ** 2
    if workload == "C4":
        return 1.0 * n * n * m + m * (m + 1.0) * n
    return None


def profile_value(name, key):
    """A figure recorded in a profiles/ summary (``key: value`` line), or None."""
    path = os.path.join(PROFILES, name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(key + ":"):
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
    """The FP64 result against the reference's stored result on the same instance, with the
    reference's own spread on it (runs that differ only in the order of their arithmetic)."""
    if fixture is None:
        return None
    path = os.path.join(GOLDEN, fixture)
    with open(path + ".json") as f:
        fix = json.load(f)
    y0 = np.load(path + "_y.npy")
    out = {
        "vs": "reference HDSDP FP64 (CPU, default build) on the same instance",
        "status": [res.status, fix["status"]],
        "iters": [res.iters, fix["iters"]],
        "dobj_rel_err": abs(res.dobj - fix["dobj"]) / max(1.0, abs(fix["dobj"])),
        "pobj_rel_err": abs(res.pobj - fix["pobj"]) / max(1.0, abs(fix["pobj"])),
        "y_rel_err": float(np.linalg.norm(res.y - y0) / np.linalg.norm(y0)),
        "gap": res.rel_gap,
    }
    if "spread" in fix:
        out["reference_spread"] = {k: fix["spread"][k] for k in ("y", "pobj", "dobj")}
    return out


def cpu_baseline(sample):
    """Reference HDSDP (built from its sources into oracle/_ref) on the host, on a bounded
    sample: the workload's instance or a labelled smaller one of the same family, for a
    full solve to the 1e-6 gap or its first few iterations.  The thread count is the one
    the reference reports using (it caps its MKL threads at min(12, MKL's maximum),
    interface/hdsdp.c:306-315)."""
    desc, make, maxiter = sample
    if not os.path.exists(REF_RUNNER):
        return None
    from hdsdp_amd import instances as I

    inst = make(I)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "inst.bin")
        I.write_instance(inst, path)
        args = [REF_RUNNER, path, os.path.join(td, "y.bin"), "psdp=0", "relopt=1e-6", "absopt=1e30"]
        if maxiter:
            args.append(f"maxiter={maxiter}")
        r = subprocess.run(args, capture_output=True, text=True, timeout=600)
    if "RESULT " not in r.stdout:
        return {"error": (r.stdout + r.stderr)[-400:]}
    res = json.loads(r.stdout.split("RESULT ", 1)[1].splitlines()[0])
    if res.get("status", -1) < 0 or res["iters"] <= 0:
        return {"error": "reference solve failed"}
    return {
        "value": res["time"] / res["iters"],
        "unit": "s/iter",
        "cores": res.get("threads"),
        "kind": "reference",
        "sample": f"{desc}: {res['iters']} iterations in {res['time']:.2f} s, MKL {res.get('threads')} threads",
        "solve_s": res["time"],
        "instance": inst.name,
        "maxiter": maxiter,
    }


def gpu_on_sample(lib, sample, precs):
    """The GPU paths on the CPU baseline's sample instance (same iteration cap)."""
    from hdsdp_amd import instances as I
    from hdsdp_amd.solver import run_solver

    _, make, maxiter = sample
    inst = make(I)
    out = {}
    for name, p in precs:
        res = run_solver(lib, inst, relopt=1e-6, absopt=1e30, precision=p, maxiter=maxiter)
        out[name] = {"s_per_iter": res.seconds / max(res.iters, 1), "iters": res.iters}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): every GPU box here has one "
              "MI355X and the solver's Schur/Cholesky path runs on a single device.", file=sys.stderr)
        sys.exit(2)

    import torch

    from hdsdp_amd import instances as I
    from hdsdp_amd.solver import Solver, load_library, run_solver

    lib = load_library()
    desc, make, fixture, cpu_sample = WORKLOADS[args.workload]
    t_gen = time.perf_counter()
    inst = make(I)
    print(f"[bench] instance {inst.name} generated in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr,
          flush=True)
    precs = {"fp64": [("fp64", 0)], "mixed": [("mixed", 1)],
             "both": [("fp64", 0), ("mixed", 1)]}[args.precision]
    base = dict(relopt=1e-6, absopt=1e30, maxiter=args.maxiter)
    for k in range(args.warmup):
        name, p = precs[k % len(precs)]
        wm = args.warmup_maxiter if args.maxiter is None else min(args.maxiter, args.warmup_maxiter)
        run_solver(lib, inst, **dict(base, precision=p, maxiter=wm))
        print(f"[bench] warmup solve {k} ({name}, {wm} iterations) done", file=sys.stderr, flush=True)

    # Each step: the pre-solve (host CSC -> HBM upload, rank-one detection, scaling) runs
    # untimed, then HDSDPOptimize on the resident data is timed between two device
    # synchronisations; the upload-inclusive time is reported beside it.
    runs = {name: {"results": [], "elapsed": 0.0, "upload": 0.0} for name, _ in precs}
    for k in range(args.steps):
        name, p = precs[k % len(precs)]
        with Solver(lib, inst, profile=True, precision=p, **base) as sv:
            torch.cuda.synchronize()
            tu = time.perf_counter()
            sv.setup()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = sv.optimize()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
        r = runs[name]
        r["results"].append(res)
        r["upload"] += t0 - tu
        r["elapsed"] += t1 - t0
        print(f"[bench] step {k} ({name}) done: {res.iters} iterations, {t1 - t0:.3f} s", file=sys.stderr,
              flush=True)

    m, n = inst.m, inst.n
    sflops = schur_flops(args.workload, n, m)
    summary = {}
    for name, p in precs:
        r = runs[name]
        rs = r["results"]
        if not rs:
            continue
        iters = sum(x.iters for x in rs)
        prof = {k: [sum(x.profile[k][0] for x in rs), sum(x.profile[k][1] for x in rs)] for k in rs[0].profile}
        s = {
            "s_per_iter": r["elapsed"] / iters,
            "solve_s": r["elapsed"] / len(rs),
            "solve_s_to_gap": float(np.mean([x.seconds_to_gap for x in rs])),
            "iters_per_solve": rs[0].iters,
            "status": rs[-1].status,
            "steps": len(rs),
            "presolve_upload_s": r["upload"] / len(rs),
            "s_per_iter_incl_upload": (r["elapsed"] + r["upload"]) / iters,
            "phases_ms_per_iter": {k: v[0] / iters for k, v in prof.items()},
            "_prof": prof,
            "_last": rs[-1],
        }
        sc_ms, sc_cnt = prof["schur"]
        gr_ms, gr_cnt = prof["gram"]
        g16_ms, g16_cnt = prof.get("gram16", [0.0, 0])
        if sflops is not None and sc_cnt:
            # the bf16 figures come from the builds made of bf16 split products only (a
            # mixed solve makes its late, ill-conditioned builds in FP64)
            bf16 = p != 0 and g16_cnt > 0
            prod = BF16_PRODUCTS if bf16 else 1
            peak = BF16_PEAK_TFLOPS if bf16 else FP64_PEAK_TFLOPS
            g_ms, g_cnt = (g16_ms, g16_cnt) if bf16 else (gr_ms, gr_cnt)
            s["schur_build"] = {
                "algorithmic_flops": sflops,
                "avg_ms": sc_ms / sc_cnt,
                "tflops_fp64_equivalent": sflops / (sc_ms / sc_cnt / 1e3) / 1e12,
                "gram_avg_ms": g_ms / max(g_cnt, 1),
                "mfma_products_per_flop": prod,
                "mfma_tflops": prod * sflops / (g_ms / max(g_cnt, 1) / 1e3) / 1e12 if g_cnt else None,
                "peak_tflops": peak,
                "builds": sc_cnt,
            }
            if p != 0:
                s["schur_build"]["bf16_builds"] = g16_cnt
                s["schur_build"]["fp64_builds"] = sc_cnt - g16_cnt
                if gr_cnt > g16_cnt:
                    s["schur_build"]["fp64_gram_avg_ms"] = (gr_ms - g16_ms) / (gr_cnt - g16_cnt)
            s["schur_build"]["frac"] = (s["schur_build"]["mfma_tflops"] or 0.0) / peak
        mf_ms, mf_cnt = prof["mfactor"]
        if mf_cnt:
            s["schur_factor"] = {"avg_ms": mf_ms / mf_cnt, "count": mf_cnt,
                                 "fp64_equiv_tflops": (2.0 * m ** 3 / 3.0) / (mf_ms / mf_cnt / 1e3) / 1e12}
        summary[name] = s

    head = "mixed" if "mixed" in summary else "fp64"
    hs = summary[head]
    hprof = hs["_prof"]

    # roofline of the dominant kernel of the headline path
    if sflops is not None and "schur_build" in hs and hprof["gram"][1]:
        sb = hs["schur_build"]
        kern = "gemm_bf16p_kernel (bf16 plane products of the dense Schur build: U^T = E^T L^-T, " \
            "T = U L^-T + (U L^-T)^T, <T_p, T_q>)" if sb["mfma_products_per_flop"] > 1 else \
            "gemm_kernel / gemm_big_kernel (FP64 Gram products)"
        roof = {
            "bound": "mfma", "kernel": kern, "achieved": sb["mfma_tflops"], "peak": sb["peak_tflops"],
            "unit": "TFLOP/s", "frac": sb["frac"],
            "traffic": profile_value(f"dense1000_{head}_gram_traffic.md", "traffic_bytes_per_build")
            if args.workload == "C3" else None,
            "algorithmic_flops_per_build": sflops * sb["mfma_products_per_flop"],
            "avg_build_ms": sb["gram_avg_ms"],
        }
    else:
        # the Schur pair step (sparse constraints): 8 bytes per lower entry of M among the
        # small constraints + S^-1 read once, against the HBM peak
        if inst.factor is None:
            col_nnz = np.diff(inst.beg)[1:]
            n_small = int(np.count_nonzero(col_nnz <= 4))
            single = n_small > 0 and n <= 480 and bool(np.all(col_nnz[col_nnz <= 4] == 1))
        else:
            n_small, single = 0, False
        pair_bytes = 8.0 * n_small * (n_small + 1) / 2 + 8.0 * n * n
        pair_ms, pair_cnt = hprof["pairs"]
        avg_s = pair_ms / 1e3 / max(pair_cnt, 1)
        ach = pair_bytes / avg_s / 1e9 if pair_cnt else 0.0
        roof = {
            "bound": "hbm", "kernel": "schur_pair_single_kernel" if single else "schur_pair_kernel",
            "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": profile_value("theta400_schur_pairs_traffic.md", "traffic_bytes_per_launch")
            if args.workload == "C2" else None,
            "bytes_per_launch": pair_bytes, "avg_launch_us": avg_s * 1e6, "launches": pair_cnt,
        }

    line = {
        "metric": "IPM s/iter + solve s to 1e-6 gap, mixed vs FP64 GPU; Schur TFLOP/s vs peak",
        "value": hs["s_per_iter"],
        "unit": "s/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sum(runs[k]["elapsed"] for k in runs) / args.steps * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "bf16x3+f32+f64",
        "data": "synthetic (seeded instance generator, numpy PCG64)",
        "config": {"workload": f"{desc}, {args.precision} GPU, gap 1e-6", "name": args.workload, "n": n, "m": m,
                   "precision": args.precision, "value_is": f"{head} s/iter", "maxiter": args.maxiter},
    }
    for name, s in summary.items():
        line[name] = {k: v for k, v in s.items() if not k.startswith("_")}
    if "fp64" in summary and "mixed" in summary:
        f, x = summary["fp64"], summary["mixed"]
        rf, rx = f["_last"], x["_last"]
        line["mixed_vs_fp64"] = {
            "speedup_per_iter": f["s_per_iter"] / x["s_per_iter"],
            "speedup_per_solve": f["solve_s"] / x["solve_s"],
            "speedup_to_gap": f["solve_s_to_gap"] / x["solve_s_to_gap"],
            "north_star_target_per_iter": 5.0,
            "y_rel_err": float(np.linalg.norm(rx.y - rf.y) / np.linalg.norm(rf.y)),
            "pobj_rel_err": abs(rx.pobj - rf.pobj) / max(1.0, abs(rf.pobj)),
            "dobj_rel_err": abs(rx.dobj - rf.dobj) / max(1.0, abs(rf.dobj)),
            "gap_fp64": rf.rel_gap,
            "gap_mixed": rx.rel_gap,
        }
    line["roofline"] = roof
    if "fp64" in summary:
        line["parity"] = parity(summary["fp64"]["_last"], fixture if args.maxiter is None else None)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline(cpu_sample)
        if cb and "error" not in cb and cpu_sample[1](I).n != n:
            cb["gpu_same_sample"] = gpu_on_sample(lib, cpu_sample, precs)
        line["cpu_baseline"] = cb
    if args.workload == "C5" and args.adjoint_dirs > 0 and "fp64" in summary:
        print("[bench] adjoint check", file=sys.stderr, flush=True)
        line["adjoint"] = adjoint_report(lib, inst, args.adjoint_dirs)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
