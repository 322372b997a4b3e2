"""ctypes/numpy front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
against.  Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of
bench.py import it; the product package (apspie_amd) never does.

Every function restates a piece of the reference apspie code; the C file
oracle/oracle.c carries the file:line citations.  Graphs are passed around as
``(m, rowptr int64[m+1], col int32[nnz], val float32[nnz] | None)``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_MT_PATH = HERE / "_ref" / "libmt19937ar.so"

SR_PLUS_TIMES, SR_BOOL, SR_MIN_PLUS, SR_MAX_FIRST = 0, 1, 2, 3

_lib = None


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref/ when the reference tree is present)."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-C", str(HERE), "all"], check=True,
                       stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _declare(L):
    vp, i32, i64, u32, f32, f64 = (C.c_void_p, C.c_int32, C.c_int64, C.c_uint32,
                                   C.c_float, C.c_double)
    L.or_mt_stream.argtypes = [u32, i64, vp]
    L.or_mt_res53_stream.argtypes = [u32, i64, vp]
    L.or_parse_mtx.argtypes = [C.c_char_p, i64, vp, vp, vp, vp, vp, vp, vp]
    L.or_parse_mtx.restype = i64
    L.or_symmetrize.argtypes = [i64, vp, vp, vp]
    L.or_symmetrize.restype = i64
    L.or_coo_to_csr.argtypes = [i32, i64, vp, vp, vp, vp, vp, vp]
    L.or_rmat_edges.argtypes = [C.c_int, C.c_int, f64, f64, f64, u32, vp, vp]
    L.or_lattice_edges.argtypes = [i32, u32, vp, vp]
    L.or_lattice_edges.restype = i64
    L.or_assign_weights.argtypes = [i32, vp, vp, u32, vp]
    L.or_bfs.argtypes = [i32, vp, vp, i32, vp]
    L.or_bfs.restype = i32
    L.or_sssp.argtypes = [i32, vp, vp, vp, i32, vp]
    L.or_sssp.restype = i32
    L.or_check_bfs.argtypes = [i32, vp, vp, vp, i32]
    L.or_check_bfs.restype = i32
    L.or_check_sssp.argtypes = [i32, vp, vp, vp, vp, i32, C.c_float]
    L.or_check_sssp.restype = i32
    L.or_mis_greedy.argtypes = [i32, vp, vp, vp, vp]
    L.or_mis_luby.argtypes = [i32, vp, vp, vp, f32, vp]
    L.or_tc.argtypes = [i32, vp, vp, vp]
    L.or_tc.restype = C.c_uint64
    L.or_components.argtypes = [i32, vp, vp, vp]
    L.or_spmspv.argtypes = [i32, vp, vp, vp, C.c_int, i64, vp, vp, vp, vp, vp]
    L.or_spmspv.restype = i64
    L.or_spgemm.argtypes = [i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.or_spgemm.restype = i64
    L.or_mxm.argtypes = [i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.or_mxm.restype = i64
    L.or_dcsc.argtypes = [i32, vp, vp, vp, i32, i32, vp, vp, vp, vp, vp]
    L.or_dcsc.restype = i64


# --------------------------------------------------------------------------- RNG
def mt_stream(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint32)
    lib().or_mt_stream(seed, n, _p(out))
    return out


def mt_res53_stream(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib().or_mt_res53_stream(seed, n, _p(out))
    return out


# ----------------------------------------------------------------- construction
def parse_mtx(text: str):
    """Returns (m, n, row, col, val, weighted) with the reference's column swap."""
    raw = text.encode()
    # header entry count bounds the arrays
    cap = 0
    for line in text.splitlines():
        if line.startswith("%") or not line.strip():
            continue
        cap = int(line.split()[2])
        break
    row = np.empty(2 * cap + 1, np.int32)
    col = np.empty(2 * cap + 1, np.int32)
    val = np.empty(2 * cap + 1, np.float32)
    m = C.c_int32()
    n = C.c_int32()
    z = C.c_int64()
    w = C.c_int32()
    k = lib().or_parse_mtx(raw, cap, C.byref(m), C.byref(n), C.byref(z),
                           _p(row), _p(col), _p(val), C.byref(w))
    if k < 0:
        raise ValueError(f"bad MatrixMarket text ({k})")
    return m.value, n.value, row, col, val, k, bool(w.value)


def symmetrize(row, col, val, nnz):
    """In place on arrays of capacity >= 2*nnz; returns trimmed copies."""
    k = lib().or_symmetrize(nnz, _p(row), _p(col), _p(val))
    if k < 0:
        raise MemoryError
    return row[:k].copy(), col[:k].copy(), val[:k].copy()


def coo_to_csr(m, row, col, val=None):
    nnz = len(row)
    rowptr = np.empty(m + 1, np.int64)
    col_out = np.empty(nnz, np.int32)
    val_out = np.empty(nnz, np.float32)
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    v = None if val is None else np.ascontiguousarray(val, np.float32)
    rc = lib().or_coo_to_csr(m, nnz, _p(row), _p(col), _p(v), _p(rowptr),
                             _p(col_out), _p(val_out))
    if rc:
        raise ValueError("row index out of range")
    return rowptr, col_out, val_out


def edges_to_csr(m, u, v):
    """Undirected edge list -> symmetric, dedup, loop-free CSR (rows sorted)."""
    nnz = len(u)
    row = np.empty(2 * nnz + 1, np.int32)
    col = np.empty(2 * nnz + 1, np.int32)
    val = np.ones(2 * nnz + 1, np.float32)
    row[:nnz] = u
    col[:nnz] = v
    r, c, w = symmetrize(row, col, val, nnz)
    rowptr, cc, _ = coo_to_csr(m, r, c, w)
    return rowptr, cc


def mtx_graph(text: str):
    """.mtx text -> symmetric CSR (the -undirected path of the drivers)."""
    m, n, row, col, val, k, _ = parse_mtx(text)
    r, c, w = symmetrize(row, col, val, k)
    rowptr, cc, vv = coo_to_csr(max(m, n), r, c, w)
    return max(m, n), rowptr, cc, vv


def rmat_edges(scale, ef=16, a=0.57, b=0.19, c=0.19, seed=1):
    ne = ef << scale
    src = np.empty(ne, np.int32)
    dst = np.empty(ne, np.int32)
    lib().or_rmat_edges(scale, ef, a, b, c, seed, _p(src), _p(dst))
    return src, dst


def rmat_graph(scale, ef=16, seed=1):
    s, d = rmat_edges(scale, ef, seed=seed)
    return (1 << scale,) + edges_to_csr(1 << scale, s, d)


def lattice_edges(W, seed=1):
    cap = 2 * W * (W - 1) + (2 * W * (W - 1)) // 1000 + 1
    eu = np.empty(cap, np.int32)
    ev = np.empty(cap, np.int32)
    k = lib().or_lattice_edges(W, seed, _p(eu), _p(ev))
    if k < 0:
        raise MemoryError
    return eu[:k].copy(), ev[:k].copy()


def lattice_graph(W, seed=1):
    eu, ev = lattice_edges(W, seed)
    return (W * W,) + edges_to_csr(W * W, eu, ev)


def assign_weights(m, rowptr, col, seed=2):
    val = np.zeros(len(col), np.float32)
    rc = lib().or_assign_weights(m, _p(rowptr), _p(col), seed, _p(val))
    if rc:
        raise ValueError("graph is not symmetric")
    return val


def mis_priorities(m, seed=5489):
    """prio[v] = v-th draw of mt19937(seed)."""
    return mt_stream(seed, m)


# ------------------------------------------------------------------ algorithms
def bfs(m, rowptr, col, src):
    depth = np.empty(m, np.int32)
    sd = lib().or_bfs(m, _p(rowptr), _p(col), src, _p(depth))
    return depth, sd


def sssp(m, rowptr, col, w, src):
    dist = np.empty(m, np.float32)
    rc = lib().or_sssp(m, _p(rowptr), _p(col), _p(np.ascontiguousarray(w, np.float32)),
                       src, _p(dist))
    if rc:
        raise ValueError("bad source")
    return dist


def check_bfs(m, rowptr, col, depth, src):
    """0 if the depths satisfy the invariants that fix them exactly (on a
    symmetric CSR), else the number of the first failed check."""
    return int(lib().or_check_bfs(m, _p(np.ascontiguousarray(rowptr, np.int64)), _p(col),
                                  _p(np.ascontiguousarray(depth, np.int32)), src))


def check_sssp(m, rowptr, col, w, dist, src, inf=np.float32(1.70141e38)):
    """0 if the distances satisfy the invariants that fix them exactly
    (positive weights; unreached = inf), else the first failed check."""
    return int(lib().or_check_sssp(m, _p(np.ascontiguousarray(rowptr, np.int64)), _p(col),
                                   _p(np.ascontiguousarray(w, np.float32)),
                                   _p(np.ascontiguousarray(dist, np.float32)), src,
                                   float(inf)))


def mis_greedy(m, rowptr, col, prio):
    out = np.empty(m, np.int32)
    lib().or_mis_greedy(m, _p(rowptr), _p(col), _p(np.ascontiguousarray(prio, np.uint32)),
                        _p(out))
    return out


def mis_luby(m, rowptr, col, prio, delta=0.1):
    out = np.empty(m, np.int32)
    rounds = lib().or_mis_luby(m, _p(rowptr), _p(col),
                               _p(np.ascontiguousarray(prio, np.uint32)), delta, _p(out))
    return out, rounds


def tc(m, rowptr, col):
    per_row = np.empty(m, np.int32)
    total = lib().or_tc(m, _p(rowptr), _p(col), _p(per_row))
    return int(total), per_row


def components(m, rowptr, col):
    lab = np.empty(m, np.int32)
    lib().or_components(m, _p(rowptr), _p(col), _p(lab))
    return lab


def spmspv(m, rowptr, col, aval, semiring, x_ind, x_val=None, mask=None):
    x_ind = np.ascontiguousarray(x_ind, np.int32)
    x_val = None if x_val is None else np.ascontiguousarray(x_val, np.float32)
    mask = None if mask is None else np.ascontiguousarray(mask, np.int32)
    aval = None if aval is None else np.ascontiguousarray(aval, np.float32)
    y_ind = np.empty(m, np.int32)
    y_val = np.empty(m, np.float32)
    ny = lib().or_spmspv(m, _p(rowptr), _p(col), _p(aval), semiring, len(x_ind),
                         _p(x_ind), _p(x_val), _p(mask), _p(y_ind), _p(y_val))
    return y_ind[:ny].copy(), y_val[:ny].copy()


def mxv_dense(m, n, rowptr, col, aval, x, y=None):
    """Dense mXv (mXv.cuh:110-247): returns (y, total).  The frontier is the
    entries of x above 0.5 (bitify, mXv.cuh:50-55) in index order (Scan +
    streamCompact, :142-147); y (n entries) is zeroed (preprocessFlag, :185)
    and receives the sparse mXv of the frontier with x's values (the same
    expand / sort / reduce-by-key stages, :153-205, restated by or_spmspv);
    total = the sum of the frontier rows' lengths (:162, returned at :229).
    An empty frontier returns y as given and 0 (:143-145)."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros(n, np.float32) if y is None else np.array(y, np.float32)
    x_ind = np.nonzero(x > np.float32(0.5))[0].astype(np.int32)
    if len(x_ind) == 0:
        return y, 0
    total = int((rowptr[x_ind + 1] - rowptr[x_ind]).sum())
    yi, yv = spmspv(max(m, n), rowptr, col, aval, SR_PLUS_TIMES, x_ind, x[x_ind])
    y[:] = 0
    y[yi] = yv
    return y, total


def spgemm(m, k, n, a_rowptr, a_col, a_val, b_rowptr, b_col, b_val):
    L = lib()
    a_val = None if a_val is None else np.ascontiguousarray(a_val, np.float32)
    b_val = None if b_val is None else np.ascontiguousarray(b_val, np.float32)
    c_rowptr = np.empty(m + 1, np.int64)
    nnz = L.or_spgemm(m, k, n, _p(a_rowptr), _p(a_col), _p(a_val), _p(b_rowptr),
                      _p(b_col), _p(b_val), _p(c_rowptr), None, None)
    c_col = np.empty(nnz, np.int32)
    c_val = np.empty(nnz, np.float32)
    L.or_spgemm(m, k, n, _p(a_rowptr), _p(a_col), _p(a_val), _p(b_rowptr), _p(b_col),
                _p(b_val), _p(c_rowptr), _p(c_col), _p(c_val))
    return c_rowptr, c_col, c_val


def mxm(p, q, r, a_colptr, a_row, a_val, b_colptr, b_row, b_val):
    """C = A*B column by column (mXm): A p x q and B q x r in CSC; returns C's
    (colptr int64, row indices, values) in CSC."""
    L = lib()
    a_colptr = np.ascontiguousarray(a_colptr, np.int64)
    b_colptr = np.ascontiguousarray(b_colptr, np.int64)
    a_row = np.ascontiguousarray(a_row, np.int32)
    b_row = np.ascontiguousarray(b_row, np.int32)
    a_val = None if a_val is None else np.ascontiguousarray(a_val, np.float32)
    b_val = None if b_val is None else np.ascontiguousarray(b_val, np.float32)
    c_colptr = np.empty(r + 1, np.int64)
    nnz = L.or_mxm(p, q, r, _p(a_colptr), _p(a_row), _p(a_val), _p(b_colptr), _p(b_row),
                   _p(b_val), _p(c_colptr), None, None)
    if nnz < 0:
        raise MemoryError("or_mxm")
    c_row = np.empty(nnz, np.int32)
    c_val = np.empty(nnz, np.float32)
    L.or_mxm(p, q, r, _p(a_colptr), _p(a_row), _p(a_val), _p(b_colptr), _p(b_row), _p(b_val),
             _p(c_colptr), _p(c_row), _p(c_val))
    return c_colptr, c_row, c_val


def dcsc(m, rowptr, col, val, part_size, part_num):
    """Row-block DCSC (csr_to_dcsc): (part_ptr, col_ind, col_off, row_ind, val)."""
    if part_size < 1 or part_num < 1 or (part_num - 1) * part_size > m:
        raise ValueError("parts do not fit the rows")
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    val = None if val is None else np.ascontiguousarray(val, np.float32)
    nnz = int(rowptr[-1])
    pp = np.empty(part_num + 1, np.int64)
    ci = np.empty(max(nnz, 1), np.int32)
    co = np.empty(nnz + 1, np.int64)
    ri = np.empty(max(nnz, 1), np.int32)
    vo = np.empty(max(nnz, 1), np.float32)
    L = lib().or_dcsc(m, _p(rowptr), _p(col), _p(val), part_size, part_num, _p(pp), _p(ci),
                      _p(co), _p(ri), _p(vo))
    if L < 0:
        raise MemoryError("or_dcsc")
    return pp, ci[:L].copy(), co[:L + 1].copy(), ri[:nnz].copy(), vo[:nnz].copy()


def lower(m, rowptr, col):
    """Strictly-lower part L (col < row) of a CSR, rows kept sorted."""
    counts = np.zeros(m, np.int64)
    rows = np.repeat(np.arange(m, dtype=np.int64), np.diff(rowptr))
    keep = col < rows
    np.add.at(counts, rows[keep], 1)
    lrp = np.zeros(m + 1, np.int64)
    np.cumsum(counts, out=lrp[1:])
    return lrp, col[keep].astype(np.int32)


# ------------------------------------------------------------ reference build
def ref_mt():
    """The reference's own mt19937ar.cpp compiled into _ref/ (None if absent)."""
    if not REF_MT_PATH.exists():
        return None
    L = C.CDLL(str(REF_MT_PATH))
    init = getattr(L, "_Z12init_genrandm")
    init.argtypes = [C.c_ulong]
    i32 = getattr(L, "_Z13genrand_int32v")
    i32.restype = C.c_ulong
    r53 = getattr(L, "_Z13genrand_res53v")
    r53.restype = C.c_double
    return init, i32, r53
