"""The reference's own host solver over the device ABI (the drop-in boundary).

tests/glue/hdsdp_gpu_glue.c is linked with the reference's unmodified object files
(oracle/Makefile.ref, target `glue`): its interface/hdsdp_algo.c, hdsdp.c and
hdsdp_schur.c drive the iterations, while the DENSE_SDP cone slots
(interface/hdsdp_conic.c:117-146) and the Schur solver's hdsdp_linsys_fp table
(linalg/def_hdsdp_linsolver.h:34-57) forward to include/hdsdp_gpu.h.  Two builds:

  hdsdp_glue_run         over hdsdp_amd/lib/libhdsdp_amd.so (the product, GPU)
  hdsdp_glue_run_oracle  over oracle/_build/libhdsdp_oracle.so (the CPU oracle's
                         implementation of the same ABI: checks the shim without a GPU)

A run must match the reference's stored result for the instance to the parity bar
(status, iteration count, y*, objectives within the reference's own spread) and must
have sent its work through the slots (the shim's call counts)."""
import json
import os
import types

import pytest

import devtools as dt
import golden.make_golden as mg

GLUE_DIR = os.path.join(dt.ROOT, "oracle", "_ref", "glue")
GLUE_GPU = os.path.join(GLUE_DIR, "hdsdp_glue_run")
GLUE_ORACLE = os.path.join(GLUE_DIR, "hdsdp_glue_run_oracle")


def run_glue(binary, name, timeout=1200, shadow=False):
    if not os.path.exists(binary):
        pytest.skip("boundary shim not built (needs the reference sources at build time)")
    inst, fix = dt.load_golden(name)
    print(f"[glue] {os.path.basename(binary)} on {name}{' (shadow)' if shadow else ''}", flush=True)
    saved = os.environ.get("HDSDP_GLUE_SHADOW")
    os.environ["HDSDP_GLUE_SHADOW"] = "1" if shadow else "0"
    try:
        res, y, log, out = mg.run_reference(inst, dict(fix["options"]), timeout=timeout, runner=binary)
    finally:
        if saved is None:
            os.environ.pop("HDSDP_GLUE_SHADOW", None)
        else:
            os.environ["HDSDP_GLUE_SHADOW"] = saved
    assert "GLUE " in out, out[-2000:]
    calls = json.loads(out.split("GLUE ", 1)[1].splitlines()[0])
    assert res.get("retcode") == 0, out[-3000:]
    r = types.SimpleNamespace(status=res["status"], iters=res["iters"], pobj=res["pobj"], dobj=res["dobj"], y=y,
                              log=log)
    return r, calls, out, fix


def check_forwarded(r, calls, out):
    """The cone and the Schur solver were the device's, for every iteration."""
    assert "[glue] DENSE_SDP cone" in out and "Schur solver: GPU kind" in out, out[:3000]
    assert calls["schur_builds"] >= r.iters
    assert calls["kkt_factorizations"] >= r.iters
    assert calls["kkt_solves"] >= 2 * r.iters
    assert calls["factorizations"] > 0 and calls["ratio_tests"] > 0 and calls["updates"] > 0
    if r.status != 5:  # a solve that ends with suspected infeasibility recovers no X
        assert calls["primal_recoveries"] >= 1


@pytest.mark.parametrize("name", ["c1_gap1e-6", "maxcut_n60", "theta_n50", "dense_n40_m60", "hsd_dense_n10_m15"])
def test_reference_host_over_oracle_abi(name):
    r, calls, out, fix = run_glue(GLUE_ORACLE, name)
    check_forwarded(r, calls, out)
    dt.check_parity(fix, r, f"glue/oracle {name}")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1_gap1e-6", "c2_gap1e-6", "hsd_dense_n10_m15"])
def test_reference_host_over_device_abi(name):
    """C1, C2 and a self-dual embedding run through the reference's unmodified hdsdp_algo.c with the device slots."""
    r, calls, out, fix = run_glue(GLUE_GPU, name)
    check_forwarded(r, calls, out)
    dt.check_parity(fix, r, f"glue/gpu {name}")


# Shadow mode: the reference's own slot runs beside every device call and steers, so
# each device answer is compared with the reference's at the reference's own iterates.
# Bounds for the quantities the dual-scaling method uses as computed values: FP64
# rounding of S^-1-based products grows with cond(S) toward the final iterates (largest
# differences seen: M 5.1e-7 and A S^-1 7.8e-8 on C2 on the GPU, 6e-8 over the oracle;
# log det 1.9e-9; X 5.6e-8).  The step-length estimate and the Schur solves are reported
# but not bounded: the Lanczos bound depends on a Ritz residual and on the warm start
# the device's own previous call left, and the late Schur systems are ill-conditioned
# enough that two CG runs to the same residual tolerance differ in their solutions.
SHADOW_BOUNDS = {"schur_m": 1e-5, "asinv": 1e-6, "asinvrd": 1e-6, "trace_sinv": 1e-10, "logdet": 1e-8,
                 "primal_x": 1e-6, "hsd_terms": 1e-6}


def check_shadow(out, r, fix):
    assert "SHADOW " in out, out[-2000:]
    sh = json.loads(out.split("SHADOW ", 1)[1].splitlines()[0])
    print("shadow", json.dumps(sh), flush=True)
    # the iterates are the reference's own: its status, iterations and log
    assert r.status == fix["status"] and r.iters == fix["iters"]
    assert [(int(a[0]), int(a[1])) for a in r.log] == [(int(a[0]), int(a[1])) for a in fix["log"]]
    assert sh["interior_flips"] == 0
    for kind, bound in SHADOW_BOUNDS.items():
        worst, at, count = sh[kind]
        assert worst <= bound, (kind, worst, at, count)
    assert sh["schur_m"][2] >= r.iters and sh["ratio_test"][2] > 0 and sh["kkt_solve"][2] > 0


@pytest.mark.parametrize("name", ["c1_gap1e-6", "theta_n50", "dense_n40_m60", "rankone_n60_m500",
                                  "hsd_dense_n10_m15", "hsd_dense_n60_m120"])
def test_oracle_abi_at_reference_iterates(name):
    r, calls, out, fix = run_glue(GLUE_ORACLE, name, shadow=True)
    check_shadow(out, r, fix)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1_gap1e-6", "c2_gap1e-6", "dense_n120_m240", "rankone_n60_m500",
                                  "hsd_dense_n10_m15", "hsd_dense_n60_m120"])
def test_device_abi_at_reference_iterates(name):
    """Every device slot call of a reference run, against the reference's own answer."""
    r, calls, out, fix = run_glue(GLUE_GPU, name, shadow=True)
    check_shadow(out, r, fix)
