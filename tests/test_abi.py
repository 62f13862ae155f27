"""The C-ABI boundary: the product library and the oracle build export every
entry point declared in include/*.h, the ctypes bindings cover exactly those,
and the Python driver refuses to run a library that is not the HIP build.
No GPU call is made here."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

import devtools as dt
from hdsdp_amd import _capi
from hdsdp_amd.solver import LIB_PATH, load_library

INCLUDE = os.path.join(dt.ROOT, "include")
DECL = re.compile(r"^(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", re.M)


def declared():
    names = set()
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h"):
            continue
        with open(os.path.join(INCLUDE, h)) as f:
            text = f.read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
        for mt in DECL.finditer(text):
            if mt.group(1) not in ("if", "while", "for", "return", "sizeof", "typedef"):
                names.add(mt.group(1))
    return names


def test_headers_declare_the_bound_api():
    names = declared()
    assert len(names) > 40
    assert names == set(_capi.EXPORTED), sorted(names ^ set(_capi.EXPORTED))


@pytest.mark.parametrize("which", ["product", "oracle"])
def test_library_exports_every_declared_symbol(which):
    path = LIB_PATH if which == "product" else dt.ORACLE_LIB
    if not os.path.exists(path):
        pytest.skip(f"{path} not built")
    lib = C.CDLL(path, mode=C.RTLD_LOCAL)
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_product_is_the_hip_build():
    if not os.path.exists(LIB_PATH):
        pytest.skip("product not built")
    lib = _capi.bind(C.CDLL(LIB_PATH, mode=C.RTLD_LOCAL))
    assert lib.hdsdp_dev_backend_name().decode() == "hip-gfx950"
    with open(LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"gfx950" in blob  # offload bundle for the MI355X


def test_driver_refuses_non_hip_library():
    if not os.path.exists(dt.ORACLE_LIB):
        pytest.skip("oracle not built")
    with pytest.raises(RuntimeError, match="not the HIP build"):
        load_library(dt.ORACLE_LIB)
    with pytest.raises(RuntimeError, match="not built"):
        load_library(os.path.join(dt.ROOT, "hdsdp_amd", "lib", "absent.so"))


def test_product_fails_loudly_without_a_gpu():
    """With no HIP device visible the product reports an error instead of
    computing anything on the host."""
    if not os.path.exists(LIB_PATH):
        pytest.skip("product not built")
    code = (
        "import ctypes as C, sys\n"
        f"sys.path.insert(0, {dt.ROOT!r})\n"
        "from hdsdp_amd import _capi\n"
        f"lib = _capi.bind(C.CDLL({LIB_PATH!r}))\n"
        "h = C.c_void_p()\n"
        "rc = lib.hdsdp_dev_create(C.byref(h), 4, 4, 0)\n"
        "print('RC', rc)\n"
    )
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, env=env)
    assert "RC " in r.stdout, r.stdout + r.stderr
    assert int(r.stdout.split("RC ")[1].split()[0]) != 0
    assert "no HIP device" in r.stderr
