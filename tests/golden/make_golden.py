"""Write golden fixtures from the reference HDSDP build (oracle/_ref/hdsdp_ref_run).

Each fixture holds the reference's status, iteration count, objectives, DIMACS
errors, per-iteration log and y* for one seeded instance, plus a checksum of
the instance so a drift in the generator is caught.  Where the reference's
default build factorises the dual matrix with PARDISO, the result of its
dense-dual build (LAPACK, the factorisation the GPU path computes) is kept as
well.  "spread" records how far the reference's own results move when only the
order of its arithmetic changes (thread count, MKL's instruction-set path, the
dual-matrix backend): no result on the same input can be asked to agree with the
reference more closely than the reference agrees with itself.  Run from the
repository root after `make -f oracle/Makefile.ref`:

    python tests/golden/make_golden.py [name ...]
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from hdsdp_amd import instances as I  # noqa: E402

RUNNER = os.path.join(ROOT, "oracle", "_ref", "hdsdp_ref_run")
RUNNER_DENSE = os.path.join(ROOT, "oracle", "_ref", "dense", "hdsdp_ref_run")
# the same sources built for x86-64-v3 with floating-point contraction (oracle/Makefile.ref)
RUNNER_FMA = os.path.join(ROOT, "oracle", "_ref", "fma", "hdsdp_ref_run")
RUNNER_DENSE_FMA = os.path.join(ROOT, "oracle", "_ref", "dense_fma", "hdsdp_ref_run")

# name -> (instance constructor, solver options, variants)
#   "sparse": the reference's default build factorises this instance's dual matrix with
#             PARDISO (interface/hdsdp_conic_sdp.c:65-97); the dense-dual build
#             (oracle/_ref/dense) is run as well and its result kept beside the default one
#   "heavy":  the spread runs use the default thread count only (one thread takes minutes)
#   "medium": the spread runs use both thread counts over every MKL path (as the light
#             cases), listed for instances whose runs take about a minute
#   "dense_primary": the primary result comes from the dense-dual build (the default build
#             stops with a segmentation fault on this instance: its primal method symmetrises
#             the dual step as a dense n x n array while PARDISO holds it sparse,
#             interface/hdsdp_psdp.c:218), and so do the spread runs
# Options may carry "psdp": 1, the reference's default primal method (interface/hdsdp.c:384);
# every other case runs with it off.
CASES = {
    "c1_gap1e-6": (lambda: I.maxcut(200, 0.05, 1, "pm1"), {"relopt": 1e-6, "absopt": 1e30}, {"sparse"}),
    "c1_default": (lambda: I.maxcut(200, 0.05, 1, "pm1"), {}, {"sparse"}),
    "c2_gap1e-6": (lambda: I.lovasz_theta(400, 0.1, 2), {"relopt": 1e-6, "absopt": 1e30}, {"heavy"}),
    "c5_gap1e-6": (lambda: I.maxcut(2000, 0.01, 5, "unif"), {"relopt": 1e-6, "absopt": 1e30}, {"sparse", "heavy"}),
    "c5_gap1e-8": (lambda: I.maxcut(2000, 0.01, 5, "unif"), {"relopt": 1e-8, "absopt": 1e30}, {"sparse", "heavy"}),
    "theta_n50": (lambda: I.lovasz_theta(50, 0.3, 21), {"relopt": 1e-6, "absopt": 1e30}, set()),
    "maxcut_n60": (lambda: I.maxcut(60, 0.2, 22, "unif"), {"relopt": 1e-6, "absopt": 1e30}, {"sparse"}),
    "dense_n20_m30": (lambda: I.dense_random(20, 30, 23), {"relopt": 1e-6, "absopt": 1e30}, set()),
    "dense_n40_m60": (lambda: I.dense_random(40, 60, 24), {"relopt": 1e-6, "absopt": 1e30}, set()),
    # the dense (C3) and rank-one (C4) families at sizes the CPU reference solves in seconds
    "dense_n120_m240": (lambda: I.dense_random(120, 240, 33), {"relopt": 1e-6, "absopt": 1e30}, set()),
    "rankone_n60_m500": (lambda: I.rank_one(60, 500, 5, 45, form="csc"), {"relopt": 1e-6, "absopt": 1e30}, set()),
    "dense_n200_m400": (lambda: I.dense_random(200, 400, 34), {"relopt": 1e-6, "absopt": 1e30}, {"medium"}),
    "rankone_n200_m2000": (lambda: I.rank_one(200, 2000, 20, 4, form="csc"), {"relopt": 1e-6, "absopt": 1e30},
                           {"heavy"}),
    # C = -5 I leaves no interior dual point: phase A stops with suspected infeasibility and
    # the self-dual embedding runs until tau vanishes (interface/hdsdp_algo.c:1859-1875)
    "hsd_dense_n10_m15": (lambda: I.with_data(I.dense_random(10, 15, 7), C=-np.eye(10) * 5.0), {"relopt": 1e-6},
                          set()),
    "hsd_dense_n60_m120": (lambda: I.with_data(I.dense_random(60, 120, 35), C=-np.eye(60) * 5.0), {"relopt": 1e-6},
                           set()),
    # the primal method on (it starts once the dual gap is below 10% and ends the solve)
    "theta_n50_psdp": (lambda: I.lovasz_theta(50, 0.3, 21), {"psdp": 1}, set()),
    "maxcut_n60_psdp": (lambda: I.maxcut(60, 0.2, 22, "unif"), {"psdp": 1}, {"dense_primary"}),
    "c1_psdp": (lambda: I.maxcut(200, 0.05, 1, "pm1"), {"psdp": 1}, {"dense_primary"}),
    "c2_psdp": (lambda: I.lovasz_theta(400, 0.1, 2), {"psdp": 1}, {"heavy"}),
    "c5_psdp": (lambda: I.maxcut(2000, 0.01, 5, "unif"), {"psdp": 1}, {"dense_primary", "heavy"}),
}

# MKL's instruction-set code paths (MKL_CBWR): the same reference build on the same input,
# as it runs on CPUs of different generations -- each path sums in its own order.  Heavy
# cases run the first three.
CBWR_PATHS = ("", "COMPATIBLE", "AVX2", "AVX", "SSE4_2", "SSE4_1", "SSSE3", "SSE3", "SSE2", "AVX512")

# the reference prints its first primal-method log line after two timing lines without a
# line break (interface/hdsdp_psdp.c:175-178)
PSDP_PREFIX = re.compile(r"^KKT Build up: \S+ sKKT Factorize: \S+ s")
LOG_RE = re.compile(r"^\s+(\d+)\s+([+-]\d\.\d+e[+-]\d+)\s+([+-]\d\.\d+e[+-]\d+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)")


def instance_digest(inst) -> str:
    h = hashlib.sha256()
    for a in (inst.beg, inst.idx, inst.elem, inst.b):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def run_reference(inst, opts, timeout=3000, runner=RUNNER, cbwr=""):
    env = dict(os.environ)
    env.pop("MKL_CBWR", None)
    if cbwr:
        env["MKL_CBWR"] = cbwr
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "inst.bin")
        ypath = os.path.join(td, "y.bin")
        I.write_instance(inst, path)
        args = [runner, path, ypath, "psdp=0"] + [f"{k}={v}" for k, v in opts.items()]
        r = subprocess.run(args, capture_output=True, text=True, timeout=timeout, env=env)
        out = r.stdout
        res = json.loads(out.split("RESULT ", 1)[1].splitlines()[0])
        y = np.fromfile(ypath) if os.path.exists(ypath) else None
    log = []
    phase = -1
    for line in out.splitlines():
        if "infeasible dual method" in line:
            phase = 0
        elif "self-dual method" in line:
            phase = 1
        elif "feasible dual method" in line:
            phase = 2
        elif "Primal solver is initialized" in line:
            phase = 3
        line = PSDP_PREFIX.sub("", line)
        mt = LOG_RE.match(line)
        if mt and phase >= 0:
            log.append([phase, int(mt.group(1)), float(mt.group(2)), float(mt.group(3)), float(mt.group(5)),
                        float(mt.group(6))])
    return res, y, log, out


def spread_runs(inst, opts, flags, primary):
    """Every reference run of the instance that differs from the primary one only in how
    its arithmetic is ordered or fused: the build with FMA contraction, the dense-dual
    builds (sparse cases), one thread against the default count (light cases) and MKL's
    instruction-set paths.  Returns the largest pairwise distance of y*, pObj and dObj
    over all of them (primary included), relative to the primary result, and the runs'
    status and iteration counts."""
    builds = [("default", RUNNER), ("fma", RUNNER_FMA)]
    if "dense_primary" in flags:
        builds = [("dense", RUNNER_DENSE), ("dense_fma", RUNNER_DENSE_FMA)]
    elif "sparse" in flags:
        builds += [("dense", RUNNER_DENSE), ("dense_fma", RUNNER_DENSE_FMA)]
    threads = [None] if "heavy" in flags else [None, 1]
    paths = CBWR_PATHS[:3] if "heavy" in flags else CBWR_PATHS
    runs, results = [], []
    for bname, runner in builds:
        for th in threads:
            for cb in paths:
                o = dict(opts)
                if th:
                    o["threads"] = th
                res, y, _, _ = run_reference(inst, o, runner=runner, cbwr=cb)
                runs.append({"build": bname, "threads": res.get("threads"), "cbwr": cb or "default",
                             "status": res["status"], "iters": res["iters"]})
                results.append((y, res["pobj"], res["dobj"]))
                print(f"    {bname:9s} threads {res.get('threads')} cbwr {cb or 'default':10s} "
                      f"status {res['status']} iters {res['iters']}", flush=True)
    yn = max(float(np.linalg.norm(primary[0])), 1e-300)
    pn, dn = max(1.0, abs(primary[1])), max(1.0, abs(primary[2]))
    sy = sp = sd = 0.0
    for a in range(len(results)):
        for b in range(a + 1, len(results)):
            ya, pa, da = results[a]
            yb, pb, db = results[b]
            sy = max(sy, float(np.linalg.norm(ya - yb)) / yn)
            sp = max(sp, abs(pa - pb) / pn)
            sd = max(sd, abs(da - db) / dn)
    return {"y": sy, "pobj": sp, "dobj": sd, "runs": runs}


def main(names):
    for name in names or CASES:
        make, opts, flags = CASES[name]
        inst = make()
        res, y, log, _ = run_reference(inst, opts, runner=RUNNER_DENSE if "dense_primary" in flags else RUNNER)
        fix = {"name": name, "instance": inst.name, "digest": instance_digest(inst), "options": opts,
               "status": res["status"], "iters": res["iters"], "pobj": res["pobj"], "dobj": res["dobj"],
               "mu": res["mu"], "dimacs": res["dimacs"], "ref_seconds": res["time"],
               "ref_threads": res.get("threads"), "ref_build": "dense" if "dense_primary" in flags else "default",
               "log": log}
        print(name, res["status"], res["iters"], res["dobj"], f"{res['time']:.2f}s", flush=True)
        if "sparse" in flags:
            dres, dy, dlog, _ = run_reference(inst, opts, runner=RUNNER_DENSE)
            fix["dense"] = {"status": dres["status"], "iters": dres["iters"], "pobj": dres["pobj"],
                            "dobj": dres["dobj"], "mu": dres["mu"], "dimacs": dres["dimacs"],
                            "ref_seconds": dres["time"], "log": dlog}
            np.save(os.path.join(HERE, name + "_dense_y.npy"), dy)
        fix["spread"] = spread_runs(inst, opts, flags, (y, res["pobj"], res["dobj"]))
        print("   spread", {k: v for k, v in fix["spread"].items() if k != "runs"}, flush=True)
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(fix, f, indent=1)
        np.save(os.path.join(HERE, name + "_y.npy"), y)


if __name__ == "__main__":
    main(sys.argv[1:])
