"""Build the MI355X library and the test-side libraries in-tree.

  python -m hdsdp_amd.build            # product: hdsdp_amd/lib/libhdsdp_amd.so
  python -m hdsdp_amd.build --all      # also oracle/_build and oracle/_ref (if the reference exists)

The product links the C host (hdsdp_amd/csrc/host) with the HIP kernels
(hdsdp_amd/csrc/gpu) compiled for gfx950 only.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hdsdp_amd")
HOST = os.path.join(PKG, "csrc", "host")
GPU = os.path.join(PKG, "csrc", "gpu")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib", "libhdsdp_amd.so")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 arithmetic is disabled for every HIP translation unit.  MFMA takes its
# accumulators in ordinary VGPRs: with the compiler's default AGPR form the GEMM loops
# copied every accumulator AGPR -> VGPR -> AGPR once per k slab (a drain of the matrix
# pipeline per slab) and needed 318 registers per lane, one wave per SIMD; the VGPR form
# runs the same MFMA instructions in 202 registers, two waves per SIMD.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-I" + INC, "-I" + HOST, "-I" + GPU,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result",
             "-mllvm", "-amdgpu-mfma-vgpr-form"]
C_FLAGS = ["-std=gnu11", "-O2", "-fPIC", "-Wall", "-Wno-unused-function", "-I" + INC, "-I" + HOST]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def build_product(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    cmds, objs = [], []
    for src in sorted(glob.glob(os.path.join(HOST, "*.c"))):
        o = os.path.join(OBJ, "host_" + os.path.basename(src)[:-2] + ".o")
        cmds.append(["gcc", *C_FLAGS, "-c", "-o", o, src])
        objs.append(o)
    for src in sorted(glob.glob(os.path.join(GPU, "*.hip"))):
        o = os.path.join(OBJ, "gpu_" + os.path.basename(src)[:-4] + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, "-c", "-o", o, src])
        objs.append(o)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs, "-Wl,-Bsymbolic"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> None:
    _run(["make", "-s", "-C", ROOT, "-f", "oracle/Makefile"])


def build_reference() -> bool:
    if not os.path.isdir("/root/reference"):
        return False
    # the reference builds, then the boundary test shim over the product and oracle libraries
    _run(["make", "-s", "-j8", "-C", ROOT, "-f", "oracle/Makefile.ref", "all", "glue"])
    return True


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--all", action="store_true")
    args = ap.parse_args(argv)
    print(build_product())
    if args.all:
        build_oracle()
        print("oracle built; reference built:", build_reference())


if __name__ == "__main__":
    main()
