"""SDPA sparse-format input (include/hdsdp_file_io.h, interface/hdsdp_file_io.c).

HReadSDPA of this build is checked
  * against the reference's own HReadSDPA (the reference build in oracle/_ref) on the
    same files: the same block sizes, right-hand side and CSC arrays, entry for entry
    (F0 negated into C, upper entries moved to the lower triangle, tiny entries dropped,
    the negative last block returned as the LP block);
  * by a round trip of seeded instances through write_sdpa: the instance read back hashes
    to the generator's, so the reference's stored solve result applies to it, and the
    host code over the oracle backend reproduces that result;
  * on malformed input (a short entry line, an index outside its block, a missing file).
No GPU is used."""
import ctypes as C
import os

import numpy as np
import pytest

import devtools as dt
import golden.make_golden as mg
from hdsdp_amd import _capi
from hdsdp_amd import instances as I

REF_LIB = os.path.join(dt.ROOT, "oracle", "_ref", "libhdsdp_ref.so")

P_INT, P_DBL = C.POINTER(C.c_int), C.POINTER(C.c_double)


def read_raw(lib, path):
    """All outputs of HReadSDPA as numpy arrays (None on failure)."""
    fn = lib.HReadSDPA
    fn.restype = C.c_int
    m, nb, ncols, nlp, nel = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
    dims, rhs = P_INT(), P_DBL()
    cbeg, cidx, celem = C.POINTER(P_INT)(), C.POINTER(P_INT)(), C.POINTER(P_DBL)()
    lbeg, lidx, lelem = P_INT(), P_INT(), P_DBL()
    rc = fn(path.encode(), C.byref(m), C.byref(nb), C.byref(dims), C.byref(rhs), C.byref(cbeg), C.byref(cidx),
            C.byref(celem), C.byref(ncols), C.byref(nlp), C.byref(lbeg), C.byref(lidx), C.byref(lelem), C.byref(nel))
    if rc != 0:
        return None
    mm = m.value
    out = {"m": mm, "nblk": nb.value, "ncols": ncols.value, "nlp": nlp.value, "nelem": nel.value,
           "dims": [dims[k] for k in range(nb.value)], "rhs": np.array([rhs[i] for i in range(mm)]), "blocks": []}
    for k in range(nb.value):
        beg = np.array([cbeg[k][j] for j in range(mm + 2)])
        nnz = int(beg[-1])
        out["blocks"].append((beg, np.array([cidx[k][e] for e in range(nnz)]),
                              np.array([celem[k][e] for e in range(nnz)])))
    if nlp.value > 0:
        beg = np.array([lbeg[j] for j in range(mm + 2)])
        nnz = int(beg[-1])
        out["lp"] = (beg, np.array([lidx[e] for e in range(nnz)]), np.array([lelem[e] for e in range(nnz)]))
    return out


def same(a, b):
    assert a is not None and b is not None
    for key in ("m", "nblk", "ncols", "nlp", "nelem", "dims"):
        assert a[key] == b[key], key
    assert np.array_equal(a["rhs"], b["rhs"])
    for (b1, i1, e1), (b2, i2, e2) in zip(a["blocks"], b["blocks"]):
        assert np.array_equal(b1, b2) and np.array_equal(i1, i2) and np.array_equal(e1, e2)
    if a["nlp"]:
        for x, y in zip(a["lp"], b["lp"]):
            assert np.array_equal(x, y)


HAND = """* two SDP blocks and a diagonal block
" a quoted comment line
3
3
2 3 -2
1.0 -2.5 3e-1
0 1 1 1 2.0
0 1 1 2 -1.0
0 2 3 1 4.0
1 1 2 2 1.0
1 2 2 3 0.5
1 3 1 1 -1.0
2 1 1 2 1e-13
2 2 1 1 3.0
2 2 1 1 0.25
2 3 2 2 7.0
3 1 2 1 -4.0
3 2 3 3 1.5
"""


def test_reader_matches_reference_on_hand_file(tmp_path, oracle):
    """Two SDP blocks, a diagonal block, comments, an upper and a lower entry of the same
    position, a repeated entry (kept twice), a dropped 1e-13 entry and F0 entries in every
    block: the same arrays as the reference's reader."""
    if not os.path.exists(REF_LIB):
        pytest.skip("reference build absent")
    path = str(tmp_path / "hand.dat-s")
    with open(path, "w") as f:
        f.write(HAND)
    ref = read_raw(C.CDLL(REF_LIB, mode=C.RTLD_LOCAL), path)
    got = read_raw(oracle, path)
    same(got, ref)
    assert got["dims"] == [2, 3] and got["nlp"] == 2 and got["nelem"] == 11
    # F0 (column 0) negated: block 0 entry (0, 0) = -2, (1, 0) = +1
    b0, i0, e0 = got["blocks"][0]
    assert list(i0[b0[0]:b0[1]]) == [0, 1] and list(e0[b0[0]:b0[1]]) == [-2.0, 1.0]


@pytest.mark.parametrize("name", ["maxcut_n60", "theta_n50", "dense_n20_m30"])
def test_round_trip_and_reference_reader(tmp_path, oracle, name):
    """write_sdpa -> HReadSDPA gives back the generator's instance bit for bit (so the
    reference's stored result applies), and the reference's reader reads the same arrays."""
    make = mg.CASES[name][0]
    inst = make()
    path = str(tmp_path / (name + ".dat-s"))
    I.write_sdpa(inst, path)
    back = I.read_sdpa(path, oracle, name=inst.name)
    assert back.n == inst.n and back.m == inst.m
    assert np.array_equal(back.beg, inst.beg) and np.array_equal(back.idx, inst.idx)
    assert np.array_equal(back.elem, inst.elem) and np.array_equal(back.b, inst.b)
    assert mg.instance_digest(back) == mg.instance_digest(inst)
    if os.path.exists(REF_LIB):
        same(read_raw(oracle, path), read_raw(C.CDLL(REF_LIB, mode=C.RTLD_LOCAL), path))


def test_solve_from_sdpa_matches_reference(tmp_path, oracle):
    """A solve of the instance read from its SDPA file reproduces the reference's result
    on that instance (host code over the oracle backend)."""
    inst, fix = dt.load_golden("maxcut_n60")
    path = str(tmp_path / "maxcut_n60.dat-s")
    I.write_sdpa(inst, path)
    back = I.read_sdpa(path, oracle)
    from hdsdp_amd.solver import run_solver

    res = run_solver(oracle, back, psdp=0, **fix["options"])
    dt.check_parity(fix, res, "oracle maxcut_n60 from SDPA")


@pytest.mark.parametrize("text", [
    "2\n1\n3\n1 1\n0 1 1 1\n",              # short entry line
    "2\n1\n3\n1 1\n1 1 4 1 1.0\n",          # row outside the block
    "2\n1\n3\n1 1\n3 1 1 1 1.0\n",          # constraint index beyond m
    "2\n2\n-3 3\n1 1\n1 1 1 1 1.0\n",       # diagonal block not last
    "2\n1\n3\n1\n",                          # right-hand side too short
])
def test_reader_rejects_malformed_input(tmp_path, oracle, text):
    path = str(tmp_path / "bad.dat-s")
    with open(path, "w") as f:
        f.write(text)
    assert read_raw(oracle, path) is None


def test_reader_missing_file(oracle):
    assert read_raw(oracle, "/nonexistent/file.dat-s") is None


def test_product_exports_the_reader():
    from hdsdp_amd.solver import LIB_PATH

    if not os.path.exists(LIB_PATH):
        pytest.skip("product not built")
    lib = C.CDLL(LIB_PATH, mode=C.RTLD_LOCAL)
    assert hasattr(lib, "HReadSDPA") and hasattr(lib, "HFreeSDPA")
    assert "HReadSDPA" in _capi.EXPORTED
