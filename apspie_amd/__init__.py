"""apspie_amd -- MI355X (gfx950) SpMSpV graph engine and hash SpGEMM.

Python front end over the C ABI of include/apspie.h (libapspie.so, built
in-tree by ``python -m apspie_amd.build``).  It mirrors the reference's entry
points -- bfs / sssp / mis / tc over a CSR matrix -- with numpy arrays on the
host side; device-resident use goes through the same functions with
``mem=DEVICE`` and raw device pointers.

There is no fallback: if libapspie.so is missing or no GPU is visible, calls
raise.
"""
from __future__ import annotations

import ctypes as C
import weakref
from pathlib import Path

import numpy as np

__all__ = ["Context", "Graph", "ApspieError", "lib", "HOST", "DEVICE", "SSSP_INF",
           "SR_PLUS_TIMES", "SR_BOOL", "SR_MIN_PLUS", "SR_MAX_FIRST",
           "PATH_AUTO", "PATH_SORT", "PATH_ATOMIC",
           "gen_rmat", "gen_lattice", "gen_weights", "read_mtx", "csr_write", "csr_read"]

LIB_PATH = Path(__file__).resolve().parent / "libapspie.so"

HOST, DEVICE = 0, 1
SR_PLUS_TIMES, SR_BOOL, SR_MIN_PLUS, SR_MAX_FIRST = 0, 1, 2, 3
PATH_AUTO, PATH_SORT, PATH_ATOMIC = 0, 1, 2
Q_BFS, Q_SSSP, Q_TC = 0, 1, 2
SSSP_INF = np.float32(1.70141e38)

_lib = None


class ApspieError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"apspie error {status}: {msg}")
        self.status = status


class _Csr(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int64),
                ("rowptr", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p)]


class _Coo(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int64),
                ("row", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("edges_traversed", C.c_int64), ("vertices_reached", C.c_int64),
                ("iterations", C.c_int32), ("host_syncs", C.c_int32),
                ("elapsed_ms", C.c_double), ("kernel_ms", C.c_double),
                ("kernel_launches", C.c_int64), ("edges_checked", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class _ServerCfg(C.Structure):
    _fields_ = [("max_batch", C.c_int32), ("slots", C.c_int32), ("hbm_budget", C.c_int64),
                ("queue_capacity", C.c_int32)]


class _ServerStats(C.Structure):
    _fields_ = [("batches", C.c_int64), ("queries", C.c_int64), ("busy_ms", C.c_double),
                ("traverse_ms", C.c_double), ("copy_ms", C.c_double), ("alloc_ms", C.c_double),
                ("retries", C.c_int64), ("retry_ms", C.c_double)]


def lib():
    """The loaded libapspie.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ApspieError(-6, f"{LIB_PATH} is missing: run `python -m apspie_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    vp, i32, i64, u32, f64, pp = (C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_double,
                                  C.POINTER(C.c_void_p))
    sig = {
        "apspie_abi_version": ([], C.c_int),
        "apspie_status_str": ([C.c_int], C.c_char_p),
        "apspie_last_error": ([vp], C.c_char_p),
        "apspie_ctx_create": ([C.c_int, pp], C.c_int),
        "apspie_ctx_destroy": ([vp], C.c_int),
        "apspie_ctx_set_stream": ([vp, vp], C.c_int),
        "apspie_ctx_stream": ([vp], vp),
        "apspie_ctx_set_hbm_budget": ([vp, i64], C.c_int),
        "apspie_ctx_hbm_used": ([vp], i64),
        "apspie_ctx_set_option": ([vp, C.c_char_p, i64], C.c_int),
        "apspie_ctx_get_option": ([vp, C.c_char_p, C.POINTER(i64)], C.c_int),
        "apspie_ctx_trim": ([vp], C.c_int),
        "apspie_ctx_reserve_arena": ([vp, i64], C.c_int),
        "apspie_ctx_arena_bytes": ([vp], i64),
        "apspie_ctx_sync": ([vp], C.c_int),
        "apspie_dev_alloc": ([vp, i64, pp], C.c_int),
        "apspie_dev_free": ([vp, vp, i64], C.c_int),
        "apspie_memcpy": ([vp, vp, vp, i64, C.c_int], C.c_int),
        "apspie_last_stats": ([vp, C.POINTER(Stats)], C.c_int),
        "apspie_graph_upload": ([vp, C.POINTER(_Csr), pp], C.c_int),
        "apspie_graph_wrap": ([vp, C.POINTER(_Csr), pp], C.c_int),
        "apspie_graph_from_coo": ([vp, C.POINTER(_Coo), C.c_int, C.c_int, pp], C.c_int),
        "apspie_graph_destroy": ([vp], C.c_int),
        "apspie_graph_info": ([vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64),
                               C.POINTER(i64)], C.c_int),
        "apspie_graph_download": ([vp, vp, vp, vp, vp], C.c_int),
        "apspie_graph_transpose": ([vp, vp, pp], C.c_int),
        "apspie_bfs": ([vp, vp, i32, vp, C.c_int], C.c_int),
        "apspie_bfs_noalloc": ([vp, vp, i32, vp, C.c_int], C.c_int),
        "apspie_mxv_dense": ([vp, vp, vp, vp, C.POINTER(i64), C.c_int], C.c_int),
        "apspie_components": ([vp, vp, vp, C.c_int], C.c_int),
        "apspie_sssp": ([vp, vp, i32, vp, C.c_int], C.c_int),
        "apspie_mis": ([vp, vp, vp, C.c_int, vp, C.c_int], C.c_int),
        "apspie_tc": ([vp, vp, C.POINTER(C.c_uint64), vp, C.c_int], C.c_int),
        "apspie_spmspv": ([vp, vp, C.c_int, i64, vp, vp, vp, C.POINTER(i64), vp, vp, C.c_int,
                           C.c_int], C.c_int),
        "apspie_spgemm": ([vp, vp, vp, pp], C.c_int),
        "apspie_mxm": ([vp, vp, vp, pp], C.c_int),
        "apspie_graph_to_dcsc": ([vp, vp, i32, i32, C.POINTER(i64), vp, vp, vp, vp, vp, C.c_int],
                                 C.c_int),
        "apspie_spgemm_masked": ([vp, vp, vp, vp, vp, C.POINTER(i64), C.c_int], C.c_int),
        "apspie_scan_exclusive": ([vp, vp, vp, i64, C.POINTER(i64), C.c_int], C.c_int),
        "apspie_sort_pairs": ([vp, vp, vp, i64, C.c_int, C.c_int], C.c_int),
        "apspie_bfs_batch": ([vp, vp, vp, i32, vp, C.c_int], C.c_int),
        "apspie_sssp_batch": ([vp, vp, vp, i32, vp, C.c_int], C.c_int),
        "apspie_autotune": ([vp, vp, i32], C.c_int),
        "apspie_autotune_hash": ([vp, vp], C.c_int),
        "apspie_tune_save": ([vp, vp, C.c_char_p], C.c_int),
        "apspie_tune_load": ([vp, vp, C.c_char_p, C.POINTER(i32)], C.c_int),
        "apspie_gen_rmat": ([C.c_int, C.c_int, f64, f64, f64, u32, vp, vp, C.POINTER(i64)],
                            C.c_int),
        "apspie_gen_lattice": ([i32, u32, vp, vp, C.POINTER(i64)], C.c_int),
        "apspie_gen_weights": ([i32, vp, vp, u32, vp], C.c_int),
        "apspie_read_mtx": ([C.c_char_p, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64), vp, vp,
                             vp], C.c_int),
        "apspie_csr_write": ([C.c_char_p, C.POINTER(_Csr)], C.c_int),
        "apspie_csr_read": ([C.c_char_p, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64),
                             C.POINTER(i32), vp, vp, vp], C.c_int),
        "apspie_graph_save": ([vp, vp, C.c_char_p], C.c_int),
        "apspie_graph_load": ([vp, C.c_char_p, pp], C.c_int),
        "apspie_server_create": ([C.c_int, C.POINTER(_ServerCfg), pp], C.c_int),
        "apspie_server_add_graph_coo": ([vp, C.POINTER(_Coo), C.c_int, C.c_int, C.POINTER(i32)],
                                        C.c_int),
        "apspie_server_add_graph_file": ([vp, C.c_char_p, C.POINTER(i32)], C.c_int),
        "apspie_server_last_error": ([vp], C.c_char_p),
        "apspie_server_busy": ([vp, i32, C.c_int, C.POINTER(i64), C.POINTER(i64),
                                C.POINTER(f64), C.c_int], C.c_int),
        "apspie_server_busy_stats": ([vp, i32, C.c_int, C.POINTER(_ServerStats), C.c_int],
                                     C.c_int),
        "apspie_server_destroy": ([vp], C.c_int),
        "apspie_server_add_graph": ([vp, C.POINTER(_Csr), C.POINTER(i32)], C.c_int),
        "apspie_server_submit": ([vp, i32, C.c_int, i32, C.POINTER(i64)], C.c_int),
        "apspie_server_submit_into": ([vp, i32, C.c_int, i32, vp, C.c_int, C.POINTER(i64)],
                                      C.c_int),
        "apspie_server_wait": ([vp, i64, vp, i64, C.POINTER(f64)], C.c_int),
        "apspie_server_wait_view": ([vp, i64, C.POINTER(vp), C.POINTER(i64), C.POINTER(f64)],
                                    C.c_int),
        "apspie_server_release": ([vp, i64], C.c_int),
        "apspie_server_drain": ([vp], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


def _p(a):
    if a is None:
        return None
    if isinstance(a, int):
        return a
    return a.ctypes.data_as(C.c_void_p)


def _check(st, ctx=None):
    if st != 0:
        L = lib()
        msg = L.apspie_status_str(st).decode()
        if ctx is not None:
            detail = L.apspie_last_error(ctx).decode()
            if detail:
                msg += f": {detail}"
        raise ApspieError(st, msg)


# ------------------------------------------------------------------ synthetic
def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    """Symmetric R-MAT CSR (int32 rowptr, int32 col)."""
    L = lib()
    cap = C.c_int64()
    _check(L.apspie_gen_rmat(scale, edge_factor, a, b, c, seed, None, None, C.byref(cap)))
    m = 1 << scale
    rowptr = np.empty(m + 1, np.int32)
    col = np.empty(cap.value, np.int32)
    nnz = C.c_int64()
    _check(L.apspie_gen_rmat(scale, edge_factor, a, b, c, seed, _p(rowptr), _p(col),
                             C.byref(nnz)))
    return rowptr, col[:nnz.value].copy()


def gen_lattice(width, seed=1):
    L = lib()
    cap = C.c_int64()
    _check(L.apspie_gen_lattice(width, seed, None, None, C.byref(cap)))
    m = width * width
    rowptr = np.empty(m + 1, np.int32)
    col = np.empty(cap.value, np.int32)
    nnz = C.c_int64()
    _check(L.apspie_gen_lattice(width, seed, _p(rowptr), _p(col), C.byref(nnz)))
    return rowptr, col[:nnz.value].copy()


def gen_weights(rowptr, col, seed=2):
    m = len(rowptr) - 1
    val = np.empty(len(col), np.float32)
    _check(lib().apspie_gen_weights(m, _p(rowptr), _p(col), seed, _p(val)))
    return val


def read_mtx(path):
    """COO of a MatrixMarket file (first index -> col, second -> row)."""
    L = lib()
    m, n, z = C.c_int32(), C.c_int32(), C.c_int64()
    bp = str(path).encode()
    _check(L.apspie_read_mtx(bp, C.byref(m), C.byref(n), C.byref(z), None, None, None))
    row = np.empty(z.value, np.int32)
    col = np.empty(z.value, np.int32)
    val = np.empty(z.value, np.float32)
    _check(L.apspie_read_mtx(bp, C.byref(m), C.byref(n), C.byref(z), _p(row), _p(col), _p(val)))
    k = z.value
    return m.value, n.value, row[:k], col[:k], val[:k]


def csr_write(path, rowptr, col, val=None, n=None):
    """Write host CSR arrays as a binary CSR cache file."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = None if val is None else np.ascontiguousarray(val, np.float32)
    m = len(rowptr) - 1
    csr = _Csr(m, m if n is None else n, len(col), _p(rowptr).value,
               _p(col).value if len(col) else None, None if val is None else _p(val).value)
    _check(lib().apspie_csr_write(str(path).encode(), C.byref(csr)))


def csr_read(path):
    """(rowptr, col, val or None, n) of a binary CSR cache file."""
    L = lib()
    bp = str(path).encode()
    m, n, z, hv = C.c_int32(), C.c_int32(), C.c_int64(), C.c_int32()
    _check(L.apspie_csr_read(bp, C.byref(m), C.byref(n), C.byref(z), C.byref(hv), None, None,
                             None))
    rowptr = np.empty(m.value + 1, np.int32)
    col = np.empty(max(z.value, 1), np.int32)
    val = np.empty(max(z.value, 1), np.float32) if hv.value else None
    _check(L.apspie_csr_read(bp, C.byref(m), C.byref(n), C.byref(z), C.byref(hv), _p(rowptr),
                             _p(col), _p(val)))
    k = z.value
    return rowptr, col[:k], (None if val is None else val[:k]), n.value


# ------------------------------------------------------------------- context
class Context:
    """One HIP stream, its scratch memory and tuning table on one device."""

    def __init__(self, device=0):
        self._L = lib()
        h = C.c_void_p()
        _check(self._L.apspie_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device
        self._graphs = weakref.WeakSet()

    def close(self):
        if self.h:
            for g in list(self._graphs):
                g.close()
            self._L.apspie_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- options / stats
    def set_option(self, name, value):
        _check(self._L.apspie_ctx_set_option(self.h, name.encode(), int(value)), self.h)

    def get_option(self, name):
        v = C.c_int64()
        _check(self._L.apspie_ctx_get_option(self.h, name.encode(), C.byref(v)), self.h)
        return v.value

    def set_hbm_budget(self, nbytes):
        _check(self._L.apspie_ctx_set_hbm_budget(self.h, int(nbytes)), self.h)

    def hbm_used(self):
        return self._L.apspie_ctx_hbm_used(self.h)

    def stream(self):
        return self._L.apspie_ctx_stream(self.h)

    def set_stream(self, s):
        _check(self._L.apspie_ctx_set_stream(self.h, s), self.h)

    def stats(self):
        s = Stats()
        _check(self._L.apspie_last_stats(self.h, C.byref(s)), self.h)
        return s.as_dict()

    def trim(self):
        _check(self._L.apspie_ctx_trim(self.h), self.h)

    def reserve_arena(self, nbytes):
        """Carve scratch buffers from one device block of nbytes (0 releases
        it): growing or dropping them then waits only for this context's
        stream."""
        _check(self._L.apspie_ctx_reserve_arena(self.h, int(nbytes)), self.h)

    def arena_bytes(self):
        return self._L.apspie_ctx_arena_bytes(self.h)

    # -- device memory (plain HIP allocations owned by the caller)
    def dev_alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def sync(self):
        _check(self._L.apspie_ctx_sync(self.h), self.h)

    # -- graphs
    def upload(self, rowptr, col, val=None, n=None):
        return Graph.upload(self, rowptr, col, val, n)

    def from_coo(self, m, row, col, val=None, n=None, symmetrize=True):
        return Graph.from_coo(self, m, row, col, val, n, symmetrize)

    def load(self, path):
        """A graph read from a binary CSR cache file into HBM."""
        h = C.c_void_p()
        _check(self._L.apspie_graph_load(self.h, str(path).encode(), C.byref(h)), self.h)
        return Graph(self, h)

    # -- algorithms (host numpy in/out unless mem=DEVICE with raw pointers)
    def bfs(self, g, source, out=None, mem=HOST):
        if mem == HOST:
            out = np.empty(g.m, np.int32) if out is None else out
        _check(self._L.apspie_bfs(self.h, g.h, int(source), _p(out), mem), self.h)
        return out

    def sssp(self, g, source, out=None, mem=HOST):
        if mem == HOST:
            out = np.empty(g.m, np.float32) if out is None else out
        _check(self._L.apspie_sssp(self.h, g.h, int(source), _p(out), mem), self.h)
        return out

    def components(self, g):
        """Connected-component labels (smallest vertex id of each component)."""
        out = np.empty(g.m, np.int32)
        _check(self._L.apspie_components(self.h, g.h, _p(out), HOST), self.h)
        return out

    def mxv_dense(self, g, x, y=None):
        """Dense mXv: (y, frontier edge count); x > 0.5 is the frontier.  y
        (host float32, n entries) is left as given when the frontier is
        empty (zeros when not given)."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(g.n, np.float32) if y is None else y
        t = C.c_int64()
        _check(self._L.apspie_mxv_dense(self.h, g.h, _p(x), _p(y), C.byref(t), HOST), self.h)
        return y, t.value

    def mis(self, g, prio=None, out=None, mem=HOST, prio_mem=HOST):
        if mem == HOST:
            out = np.empty(g.m, np.int32) if out is None else out
        if prio is not None and prio_mem == HOST:
            prio = np.ascontiguousarray(prio, np.uint32)
        _check(self._L.apspie_mis(self.h, g.h, _p(prio), prio_mem, _p(out), mem), self.h)
        return out

    def tc(self, g, per_row=True, out=None, mem=HOST):
        total = C.c_uint64()
        if per_row and mem == HOST and out is None:
            out = np.empty(g.m, np.int32)
        _check(self._L.apspie_tc(self.h, g.h, C.byref(total), _p(out) if per_row else None,
                                 mem), self.h)
        return int(total.value), out

    def spmspv(self, g, semiring, x_ind, x_val=None, mask=None, path=PATH_AUTO):
        x_ind = np.ascontiguousarray(x_ind, np.int32)
        x_val = None if x_val is None else np.ascontiguousarray(x_val, np.float32)
        mask = None if mask is None else np.ascontiguousarray(mask, np.int32)
        yi = np.empty(max(g.n, 1), np.int32)
        yv = np.empty(max(g.n, 1), np.float32)
        ny = C.c_int64()
        _check(self._L.apspie_spmspv(self.h, g.h, semiring, len(x_ind), _p(x_ind), _p(x_val),
                                     _p(mask), C.byref(ny), _p(yi), _p(yv), HOST, path), self.h)
        return yi[:ny.value].copy(), yv[:ny.value].copy()

    def spgemm(self, a, b):
        h = C.c_void_p()
        _check(self._L.apspie_spgemm(self.h, a.h, b.h, C.byref(h)), self.h)
        return Graph(self, h)

    def mxm(self, a_csc, b_csc):
        """C = A * B column by column (mXm): a_csc / b_csc are graphs whose
        rowptr/col/val hold A's and B's CSC arrays (m = columns, n = rows);
        the result is C in the same form."""
        h = C.c_void_p()
        _check(self._L.apspie_mxm(self.h, a_csc.h, b_csc.h, C.byref(h)), self.h)
        return Graph(self, h)

    def spgemm_masked(self, a, b, mask):
        """(A * B) o M on M's pattern: (values in M's CSR order, entries of M
        with at least one product)."""
        out = np.empty(max(mask.nnz, 1), np.float32)
        nnz = C.c_int64()
        _check(self._L.apspie_spgemm_masked(self.h, a.h, b.h, mask.h, _p(out), C.byref(nnz),
                                            HOST), self.h)
        return out[:mask.nnz], nnz.value

    def scan_exclusive(self, x):
        x = np.ascontiguousarray(x, np.int32)
        out = np.empty(max(len(x), 1), np.int32)
        tot = C.c_int64()
        _check(self._L.apspie_scan_exclusive(self.h, _p(x), _p(out), len(x), C.byref(tot), HOST),
               self.h)
        return out[:len(x)], tot.value

    def sort_pairs(self, keys, vals=None, bits=32):
        k = np.array(keys, np.uint32)
        v = None if vals is None else np.array(vals, np.uint32)
        _check(self._L.apspie_sort_pairs(self.h, _p(k), _p(v), len(k), bits, HOST), self.h)
        return k, v

    def bfs_batch(self, g, sources, out=None, mem=HOST):
        s = np.ascontiguousarray(sources, np.int32)
        if mem == HOST:
            out = np.empty((len(s), g.m), np.int32) if out is None else out
        _check(self._L.apspie_bfs_batch(self.h, g.h, _p(s), len(s), _p(out), mem), self.h)
        return out

    def sssp_batch(self, g, sources, out=None, mem=HOST):
        s = np.ascontiguousarray(sources, np.int32)
        if mem == HOST:
            out = np.empty((len(s), g.m), np.float32) if out is None else out
        _check(self._L.apspie_sssp_batch(self.h, g.h, _p(s), len(s), _p(out), mem), self.h)
        return out

    def autotune(self, g, batch=0):
        _check(self._L.apspie_autotune(self.h, g.h, int(batch)), self.h)

    def autotune_hash(self, g):
        """Tune the TC / SpGEMM tier bounds on this graph (A * A)."""
        _check(self._L.apspie_autotune_hash(self.h, g.h), self.h)

    def tune_save(self, g, path):
        """Save the tuned options as this GPU's entry for g's graph class."""
        _check(self._L.apspie_tune_save(self.h, g.h, str(path).encode()), self.h)

    def tune_load(self, g, path):
        """Apply the saved entry for this GPU and g's class; True if one matched."""
        f = C.c_int32()
        _check(self._L.apspie_tune_load(self.h, g.h, str(path).encode(), C.byref(f)), self.h)
        return bool(f.value)


class DeviceBuffer:
    """A device allocation made through the library (counts against the
    context's HBM budget); .ptr is the raw device address."""

    def __init__(self, ctx, nbytes):
        self.ctx, self.nbytes = ctx, int(nbytes)
        p = C.c_void_p()
        _check(lib().apspie_dev_alloc(ctx.h, self.nbytes, C.byref(p)), ctx.h)
        self.ptr = p.value

    def to_host(self, dtype, count):
        out = np.empty(count, dtype)
        _check(lib().apspie_memcpy(self.ctx.h, _p(out), self.ptr, out.nbytes, 1), self.ctx.h)
        return out

    def from_host(self, arr):
        arr = np.ascontiguousarray(arr)
        _check(lib().apspie_memcpy(self.ctx.h, self.ptr, _p(arr), arr.nbytes, 0), self.ctx.h)

    def free(self):
        if self.ptr and self.ctx.h:
            lib().apspie_dev_free(self.ctx.h, self.ptr, self.nbytes)
        self.ptr = None


class Graph:
    """An HBM-resident CSR matrix."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle
        ctx._graphs.add(self)
        L = lib()
        m, n, z, b = C.c_int32(), C.c_int32(), C.c_int64(), C.c_int64()
        _check(L.apspie_graph_info(handle, C.byref(m), C.byref(n), C.byref(z), C.byref(b)))
        self.m, self.n, self.nnz, self.hbm_bytes = m.value, n.value, z.value, b.value

    @classmethod
    def upload(cls, ctx, rowptr, col, val=None, n=None):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = None if val is None else np.ascontiguousarray(val, np.float32)
        if val is not None and len(val) == 0:
            val = np.zeros(1, np.float32)  # weighted but edgeless: keep a non-null pointer
        m = len(rowptr) - 1
        csr = _Csr(m, m if n is None else n, len(col), _p(rowptr).value,
                   _p(col).value if len(col) else None,
                   None if val is None else _p(val).value)
        h = C.c_void_p()
        _check(lib().apspie_graph_upload(ctx.h, C.byref(csr), C.byref(h)), ctx.h)
        return cls(ctx, h)

    @classmethod
    def from_coo(cls, ctx, m, row, col, val=None, n=None, symmetrize=True):
        row = np.ascontiguousarray(row, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = None if val is None else np.ascontiguousarray(val, np.float32)
        coo = _Coo(m, m if n is None else n, len(row), _p(row).value if len(row) else None,
                   _p(col).value if len(col) else None, None if val is None else _p(val).value)
        h = C.c_void_p()
        _check(lib().apspie_graph_from_coo(ctx.h, C.byref(coo), HOST, int(symmetrize),
                                           C.byref(h)), ctx.h)
        return cls(ctx, h)

    def download(self, values=False):
        rowptr = np.empty(self.m + 1, np.int32)
        col = np.empty(max(self.nnz, 1), np.int32)
        val = np.empty(max(self.nnz, 1), np.float32) if values else None
        _check(lib().apspie_graph_download(self.ctx.h, self.h, _p(rowptr), _p(col), _p(val)),
               self.ctx.h)
        return rowptr, col[:self.nnz], (None if val is None else val[:self.nnz])

    def to_dcsc(self, part_size, part_num, values=True):
        """Row-block DCSC (csr_to_dcsc): (part_ptr, col_ind, col_off, row_ind,
        val or None) as host arrays."""
        e = max(self.nnz, 1)
        pp = np.empty(part_num + 1, np.int32)
        ci = np.empty(e, np.int32)
        co = np.empty(e + 1, np.int32)
        ri = np.empty(e, np.int32)
        vv = np.empty(e, np.float32) if values else None
        L = C.c_int64()
        _check(lib().apspie_graph_to_dcsc(self.ctx.h, self.h, part_size, part_num, C.byref(L),
                                          _p(pp), _p(ci), _p(co), _p(ri), _p(vv), HOST),
               self.ctx.h)
        n = L.value
        return (pp, ci[:n], co[:n + 1], ri[:self.nnz],
                None if vv is None else vv[:self.nnz])

    def save(self, path):
        """Write the graph as a binary CSR cache file."""
        _check(lib().apspie_graph_save(self.ctx.h, self.h, str(path).encode()), self.ctx.h)

    def transpose(self):
        h = C.c_void_p()
        _check(lib().apspie_graph_transpose(self.ctx.h, self.h, C.byref(h)), self.ctx.h)
        return Graph(self.ctx, h)

    def close(self):
        if self.h:
            lib().apspie_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
