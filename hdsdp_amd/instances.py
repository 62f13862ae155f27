"""Seeded SDP instance generators in HDSDP's conic input format.

The solver takes one SDP block as a CSC matrix with ``m + 1`` columns over the
packed-lower row index of the ``n x n`` symmetric coefficient matrices: column 0
holds ``C`` and column ``i + 1`` holds ``A_i`` (``interface/def_hdsdp_user_data.h:11-21``,
``interface/hdsdp_conic_sdp.c:1367-1378``).  Packed-lower, column-major indexing is
``idx(i, j) = (2n - j - 1) j / 2 + i`` for ``i >= j`` (``interface/hdsdp_utils.h:48-51``).

Every generator draws from ``numpy.random.default_rng(seed)`` (PCG64), so the same
seed gives bit-identical instances on any host with the same numpy.
"""
from __future__ import annotations

import dataclasses
import struct

import numpy as np

__all__ = [
    "SDPInstance",
    "packed_index",
    "maxcut",
    "maxcut_graph",
    "lovasz_theta",
    "theta_graph",
    "dense_random",
    "rank_one",
    "write_instance",
    "read_instance",
    "write_sdpa",
    "read_sdpa",
    "with_data",
    "CONFIGS",
]


@dataclasses.dataclass
class SDPInstance:
    """One SDP block ``min <C, X> s.t. <A_i, X> = b_i, X psd``.

    ``beg``/``idx``/``elem`` is the CSC form described in the module docstring.
    ``factor`` is set instead for the factor-form rank-one input (``A_i = s_i a_i a_i^T``),
    where ``factor`` is ``n x m`` column-major and ``sign`` holds ``s_i``; then
    ``beg``/``idx``/``elem`` hold only column 0 (``C``).
    """

    name: str
    n: int
    m: int
    beg: np.ndarray  # int64, m + 2 entries
    idx: np.ndarray  # int32 packed-lower row indices
    elem: np.ndarray  # float64
    b: np.ndarray  # float64, m
    factor: np.ndarray | None = None  # float64, (n, m) Fortran order
    sign: np.ndarray | None = None  # float64, m

    @property
    def nnz(self) -> int:
        return int(self.beg[-1])

    def dense_coeff(self, col: int) -> np.ndarray:
        """Full symmetric ``n x n`` matrix of column ``col`` (0 = C, i + 1 = A_i)."""
        n = self.n
        out = np.zeros((n, n))
        if self.factor is not None and col > 0:
            a = self.factor[:, col - 1]
            return self.sign[col - 1] * np.outer(a, a)
        lo, hi = int(self.beg[col]), int(self.beg[col + 1])
        jj, ii = _unpack(n, self.idx[lo:hi].astype(np.int64))
        out[ii, jj] = self.elem[lo:hi]
        out[jj, ii] = self.elem[lo:hi]
        return out


def packed_index(n: int, i, j):
    """Packed-lower column-major index of entry ``(i, j)`` with ``i >= j``."""
    i = np.asarray(i, dtype=np.int64)
    j = np.asarray(j, dtype=np.int64)
    return (2 * n - j - 1) * j // 2 + i


def _unpack(n: int, p: np.ndarray):
    """Inverse of :func:`packed_index`: returns ``(j, i)`` (column, row)."""
    p = np.asarray(p, dtype=np.int64)
    starts = packed_index(n, np.arange(n), np.arange(n))  # diagonal positions
    j = np.searchsorted(starts, p, side="right") - 1
    i = p - starts[j] + j
    return j, i


def _assemble(name, n, cols, b):
    """Build CSC arrays from a list of (packed_idx, values) per column."""
    beg = np.zeros(len(cols) + 1, dtype=np.int64)
    for k, (pi, _) in enumerate(cols):
        beg[k + 1] = beg[k] + len(pi)
    idx = np.empty(int(beg[-1]), dtype=np.int32)
    elem = np.empty(int(beg[-1]), dtype=np.float64)
    for k, (pi, pv) in enumerate(cols):
        order = np.argsort(pi, kind="stable")
        idx[beg[k]:beg[k + 1]] = np.asarray(pi, dtype=np.int64)[order]
        elem[beg[k]:beg[k + 1]] = np.asarray(pv, dtype=np.float64)[order]
    return SDPInstance(name, n, len(cols) - 1, beg, idx, elem, np.asarray(b, dtype=np.float64))


def _random_graph(rng, n, p):
    iu, ju = np.triu_indices(n, 1)  # i < j
    keep = rng.random(iu.shape[0]) < p
    return iu[keep], ju[keep]


def with_data(inst: SDPInstance, b=None, C=None) -> SDPInstance:
    """``inst`` with the right-hand side ``b`` and/or the objective ``C`` (dense
    symmetric ``n x n``; its lower triangle is stored, zeros dropped) replaced."""
    out = dataclasses.replace(inst)
    if b is not None:
        out.b = np.array(b, dtype=np.float64).reshape(inst.m)
    if C is not None:
        n = inst.n
        Cm = np.asarray(C, dtype=np.float64).reshape(n, n)
        jj, ii = np.nonzero(np.tril(Cm).T)  # column-major order of the lower triangle
        vals = Cm[ii, jj]
        cpk = packed_index(n, ii, jj)
        order = np.argsort(cpk, kind="stable")
        c_idx, c_val = cpk[order].astype(np.int32), vals[order]
        lo, hi = int(inst.beg[1]), int(inst.beg[-1])
        shift = c_idx.shape[0] - lo
        out.beg = inst.beg.copy()
        out.beg[1:] += shift
        out.idx = np.concatenate([c_idx, inst.idx[lo:hi]])
        out.elem = np.concatenate([c_val, inst.elem[lo:hi]])
    return out


def maxcut_graph(n: int, i, j, w, name: str = "maxcut") -> SDPInstance:
    """MaxCut relaxation of the weighted graph with edges ``(i[k], j[k])``, ``i < j``:
    ``C = -L/4`` (``L`` the weighted Laplacian), ``diag(X) = 1``."""
    i = np.asarray(i, dtype=np.int64)
    j = np.asarray(j, dtype=np.int64)
    w = np.asarray(w, dtype=np.float64)
    deg = np.zeros(n)
    np.add.at(deg, i, w)
    np.add.at(deg, j, w)
    # C = -(D - W)/4: diagonal -deg/4, off-diagonal +w/4 at (max, min)
    rows = np.concatenate([np.arange(n), j])  # j > i so (j, i) is lower
    colsj = np.concatenate([np.arange(n), i])
    vals = np.concatenate([-deg / 4.0, w / 4.0])
    nz = vals != 0.0
    cpk = packed_index(n, rows[nz], colsj[nz])
    cols = [(cpk, vals[nz])]
    for k in range(n):
        cols.append((np.array([packed_index(n, k, k)]), np.array([1.0])))
    return _assemble(name, n, cols, np.ones(n))


def maxcut(n: int, p: float, seed: int, weights: str = "pm1") -> SDPInstance:
    """MaxCut relaxation on G(n, p): ``C = -L/4``, ``diag(X) = 1``.

    ``weights`` is ``"pm1"`` (uniform on {-1, +1}) or ``"unif"`` (U(0, 1)).
    """
    rng = np.random.default_rng(seed)
    i, j = _random_graph(rng, n, p)
    if weights == "pm1":
        w = rng.choice(np.array([-1.0, 1.0]), size=i.shape[0])
    elif weights == "unif":
        w = rng.random(i.shape[0])
    else:
        raise ValueError(weights)
    return maxcut_graph(n, i, j, w, f"maxcut_n{n}_p{p}_s{seed}")


def theta_graph(n: int, i, j, name: str = "theta") -> SDPInstance:
    """Lovasz theta SDP of the graph with edges ``(i[k], j[k])``, ``i < j``:
    ``C = -J``, ``tr X = 1``, ``X_ij = 0`` per edge; the optimum is ``-theta``."""
    npk = n * (n + 1) // 2
    cols = [(np.arange(npk), -np.ones(npk))]
    cols.append((packed_index(n, np.arange(n), np.arange(n)), np.ones(n)))
    for a, c in zip(i, j):
        cols.append((np.array([packed_index(n, c, a)]), np.array([1.0])))
    b = np.zeros(len(cols) - 1)
    b[0] = 1.0
    return _assemble(name, n, cols, b)


def lovasz_theta(n: int, p: float, seed: int) -> SDPInstance:
    """Lovasz theta SDP on G(n, p): ``C = -J``, ``tr X = 1``, ``X_ij = 0`` per edge."""
    rng = np.random.default_rng(seed)
    i, j = _random_graph(rng, n, p)
    return theta_graph(n, i, j, f"theta_n{n}_p{p}_s{seed}")


def dense_random(n: int, m: int, seed: int) -> SDPInstance:
    """Strictly feasible dense SDP: ``A_i`` packed N(0,1)/sqrt(n), ``C = sum y0_i A_i + I``,
    ``b_i = tr(A_i)`` (so ``X = I`` and ``y0`` are interior)."""
    rng = np.random.default_rng(seed)
    npk = n * (n + 1) // 2
    diag = packed_index(n, np.arange(n), np.arange(n))
    beg = np.arange(m + 2, dtype=np.int64) * npk
    idx = np.tile(np.arange(npk, dtype=np.int32), m + 1)
    elem = np.empty((m + 1) * npk)
    y0 = rng.standard_normal(m)
    cvec = np.zeros(npk)
    b = np.empty(m)
    scale = 1.0 / np.sqrt(n)
    for k in range(m):
        a = rng.standard_normal(npk) * scale
        elem[(k + 1) * npk:(k + 2) * npk] = a
        b[k] = a[diag].sum()
        cvec += y0[k] * a
    cvec[diag] += 1.0
    elem[:npk] = cvec
    return SDPInstance(f"dense_n{n}_m{m}_s{seed}", n, m, beg, idx, elem, b)


def rank_one(n: int, m: int, r: int, seed: int, form: str = "factor") -> SDPInstance:
    """Rank-one constraints ``A_k = a_k a_k^T``, ``a_k ~ N(0, I/n)``;
    ``b_k = a_k^T X0 a_k`` with ``X0 = Z Z^T / r + 1e-3 I``; ``C = I``.

    ``form="factor"`` returns the factor form (``factor``/``sign``, ``C`` the only
    CSC column); ``form="csc"`` expands every ``A_k`` into packed-lower CSC, the
    reference's input format (n(n+1)/2 entries per constraint).
    """
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, m)) / np.sqrt(n)
    z = rng.standard_normal((n, r))
    za = z.T @ a  # r x m
    b = (za * za).sum(axis=0) / r + 1e-3 * (a * a).sum(axis=0)
    diag = packed_index(n, np.arange(n), np.arange(n))
    name = f"rankone_n{n}_m{m}_s{seed}"
    if form == "csc":
        il, jl = np.tril_indices(n)  # row i >= column j
        pk = packed_index(n, il, jl)
        order = np.argsort(pk)
        il, jl, pk = il[order], jl[order], pk[order]
        cols = [(diag, np.ones(n))]
        for k in range(m):
            cols.append((pk, a[il, k] * a[jl, k]))
        inst = _assemble(name, n, cols, b)
        return inst
    if form != "factor":
        raise ValueError(form)
    beg = np.zeros(m + 2, dtype=np.int64)
    beg[1:] = n
    inst = SDPInstance(name, n, m, beg, diag.astype(np.int32),
                       np.ones(n), b, factor=np.asfortranarray(a), sign=np.ones(m))
    return inst


# The five benchmark configurations (seed = configuration number).
CONFIGS = {
    "C1": lambda: maxcut(200, 0.05, seed=1, weights="pm1"),
    "C2": lambda: lovasz_theta(400, 0.1, seed=2),
    "C3": lambda: dense_random(1000, 2000, seed=3),
    "C4": lambda: rank_one(2000, 20000, 20, seed=4),
    "C5": lambda: maxcut(2000, 0.01, seed=5, weights="unif"),
}

_MAGIC = b"HDSDPIN1"


def write_instance(inst: SDPInstance, path: str) -> None:
    """Binary instance file: magic, n, m, nnz, beg[m+2] (i64), idx (i32), elem (f64), b (f64)."""
    if inst.factor is not None:
        raise ValueError("factor-form instances have no CSC file form")
    with open(path, "wb") as f:
        f.write(_MAGIC)
        f.write(struct.pack("<iiq", inst.n, inst.m, inst.nnz))
        f.write(np.ascontiguousarray(inst.beg, dtype="<i8").tobytes())
        f.write(np.ascontiguousarray(inst.idx, dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(inst.elem, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(inst.b, dtype="<f8").tobytes())


def read_instance(path: str, name: str = "file") -> SDPInstance:
    with open(path, "rb") as f:
        if f.read(8) != _MAGIC:
            raise ValueError("not an instance file")
        n, m, nnz = struct.unpack("<iiq", f.read(16))
        beg = np.frombuffer(f.read(8 * (m + 2)), dtype="<i8").copy()
        idx = np.frombuffer(f.read(4 * nnz), dtype="<i4").copy()
        elem = np.frombuffer(f.read(8 * nnz), dtype="<f8").copy()
        b = np.frombuffer(f.read(8 * m), dtype="<f8").copy()
    return SDPInstance(name, n, m, beg, idx, elem, b)


def write_sdpa(inst: SDPInstance, path: str, comment: str | None = None) -> None:
    """SDPA sparse format of a one-block instance: ``F0 = -C``, ``F_i = A_i``, ``c = b``
    (the SDPA primal ``min <F0, X>`` form the reference reads, negating F0 into C,
    ``interface/hdsdp_file_io.c:234-236, 248-250``).  Entries are written upper-triangular
    (row <= column), constraint by constraint in CSC order."""
    if inst.factor is not None:
        raise ValueError("factor-form instances have no SDPA form here")
    n, m = inst.n, inst.m
    with open(path, "w") as f:
        f.write(f"* {comment or inst.name}\n")
        f.write(f"{m}\n1\n{n}\n")
        f.write(" ".join(f"{v:.17g}" for v in inst.b) + "\n")
        for k in range(m + 1):
            lo, hi = int(inst.beg[k]), int(inst.beg[k + 1])
            if hi == lo:
                continue
            jj, ii = _unpack(n, inst.idx[lo:hi].astype(np.int64))
            vals = inst.elem[lo:hi] * (-1.0 if k == 0 else 1.0)
            for j, i, v in zip(jj, ii, vals):
                f.write(f"{k} 1 {j + 1} {i + 1} {v:.17g}\n")


def read_sdpa(path: str, lib, name: str | None = None) -> SDPInstance:
    """One SDP block read by the library's ``HReadSDPA`` (include/hdsdp_file_io.h)."""
    import ctypes as C

    P_INT, P_DBL = C.POINTER(C.c_int), C.POINTER(C.c_double)
    m, nb, ncols, nlp, nel = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
    dims, rhs = P_INT(), P_DBL()
    cbeg, cidx, celem = C.POINTER(P_INT)(), C.POINTER(P_INT)(), C.POINTER(P_DBL)()
    lbeg, lidx, lelem = P_INT(), P_INT(), P_DBL()
    rc = lib.HReadSDPA(path.encode(), C.byref(m), C.byref(nb), C.byref(dims), C.byref(rhs), C.byref(cbeg),
                       C.byref(cidx), C.byref(celem), C.byref(ncols), C.byref(nlp), C.byref(lbeg), C.byref(lidx),
                       C.byref(lelem), C.byref(nel))
    if rc != 0:
        raise ValueError(f"HReadSDPA failed on {path} ({rc})")
    try:
        if nb.value != 1 or nlp.value != 0:
            raise ValueError(f"{path}: {nb.value} SDP blocks and an LP block of {nlp.value}; one SDP block expected")
        n, mm = dims[0], m.value
        beg = np.ctypeslib.as_array(cbeg[0], shape=(mm + 2,)).astype(np.int64)
        nnz = int(beg[-1])
        idx = np.ctypeslib.as_array(cidx[0], shape=(max(nnz, 1),))[:nnz].astype(np.int32)
        elem = np.ctypeslib.as_array(celem[0], shape=(max(nnz, 1),))[:nnz].astype(np.float64)
        b = np.ctypeslib.as_array(rhs, shape=(mm,)).astype(np.float64)
        return SDPInstance(name or path, n, mm, beg, idx.copy(), elem.copy(), b.copy())
    finally:
        lib.HFreeSDPA(nb.value, dims, rhs, cbeg, cidx, celem, lbeg, lidx, lelem)
