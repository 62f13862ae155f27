"""Write golden fixtures from the reference HDSDP build (oracle/_ref/hdsdp_ref_run).

Each fixture holds the reference's status, iteration count, objectives, DIMACS
errors, per-iteration log and y* for one seeded instance, plus a checksum of
the instance so a drift in the generator is caught.  Run from the repository
root after `make -f oracle/Makefile.ref`:

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

# name -> (instance constructor, solver options)
CASES = {
    "c1_gap1e-6": (lambda: I.maxcut(200, 0.05, 1, "pm1"), {"relopt": 1e-6, "absopt": 1e30}),
    "c1_default": (lambda: I.maxcut(200, 0.05, 1, "pm1"), {}),
    "c2_gap1e-6": (lambda: I.lovasz_theta(400, 0.1, 2), {"relopt": 1e-6, "absopt": 1e30}),
    "c5_gap1e-6": (lambda: I.maxcut(2000, 0.01, 5, "unif"), {"relopt": 1e-6, "absopt": 1e30}),
    "c5_gap1e-8": (lambda: I.maxcut(2000, 0.01, 5, "unif"), {"relopt": 1e-8, "absopt": 1e30}),
    "theta_n50": (lambda: I.lovasz_theta(50, 0.3, 21), {"relopt": 1e-6, "absopt": 1e30}),
    "maxcut_n60": (lambda: I.maxcut(60, 0.2, 22, "unif"), {"relopt": 1e-6, "absopt": 1e30}),
    "dense_n20_m30": (lambda: I.dense_random(20, 30, 23), {"relopt": 1e-6, "absopt": 1e30}),
    "dense_n40_m60": (lambda: I.dense_random(40, 60, 24), {"relopt": 1e-6, "absopt": 1e30}),
    # the dense (C3) and rank-one (C4) families at sizes the CPU reference solves in seconds
    "dense_n120_m240": (lambda: I.dense_random(120, 240, 33), {"relopt": 1e-6, "absopt": 1e30}),
    "rankone_n60_m500": (lambda: I.rank_one(60, 500, 5, 45, form="csc"), {"relopt": 1e-6, "absopt": 1e30}),
    "dense_n200_m400": (lambda: I.dense_random(200, 400, 34), {"relopt": 1e-6, "absopt": 1e30}),
    "rankone_n200_m2000": (lambda: I.rank_one(200, 2000, 20, 4, form="csc"), {"relopt": 1e-6, "absopt": 1e30}),
}

LOG_RE = re.compile(r"^\s+(\d+)\s+([+-]\d\.\d+e[+-]\d+)\s+([+-]\d\.\d+e[+-]\d+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)")


def instance_digest(inst) -> str:
    h = hashlib.sha256()
    for a in (inst.beg, inst.idx, inst.elem, inst.b):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def run_reference(inst, opts, timeout=3000):
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "inst.bin")
        ypath = os.path.join(td, "y.bin")
        I.write_instance(inst, path)
        args = [RUNNER, path, ypath, "psdp=0"] + [f"{k}={v}" for k, v in opts.items()]
        r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
        out = r.stdout
        res = json.loads(out.split("RESULT ", 1)[1].splitlines()[0])
        y = np.fromfile(ypath) if os.path.exists(ypath) else None
    log = []
    phase = -1
    for line in out.splitlines():
        if "infeasible dual method" in line:
            phase = 0
        elif "feasible dual method" in line:
            phase = 2
        mt = LOG_RE.match(line)
        if mt and phase >= 0:
            log.append([phase, int(mt.group(1)), float(mt.group(2)), float(mt.group(3)), float(mt.group(5)),
                        float(mt.group(6))])
    return res, y, log, out


def main(names):
    for name in names or CASES:
        make, opts = CASES[name]
        inst = make()
        res, y, log, _ = run_reference(inst, opts)
        fix = {"name": name, "instance": inst.name, "digest": instance_digest(inst), "options": opts,
               "status": res["status"], "iters": res["iters"], "pobj": res["pobj"], "dobj": res["dobj"],
               "mu": res["mu"], "dimacs": res["dimacs"], "ref_seconds": res["time"],
               "log": log}
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(fix, f, indent=1)
        np.save(os.path.join(HERE, name + "_y.npy"), y)
        print(name, res["status"], res["iters"], res["dobj"], f"{res['time']:.2f}s")


if __name__ == "__main__":
    main(sys.argv[1:])
