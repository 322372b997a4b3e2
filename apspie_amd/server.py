"""Python front end of the continuous-batching query server (apspie_server_*
in include/apspie.h)."""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import (HOST, Q_BFS, Q_SSSP, Q_TC, ApspieError, _Coo, _Csr, _ServerCfg, _ServerStats,
               _check, _p, lib)

__all__ = ["Server"]


class Server:
    """Resident graph pool + request queue + `slots` worker streams."""

    def __init__(self, device=0, slots=2, max_batch=64, hbm_budget=0, queue_capacity=0):
        self._L = lib()
        cfg = _ServerCfg(max_batch, slots, int(hbm_budget), queue_capacity)
        h = C.c_void_p()
        _check(self._L.apspie_server_create(device, C.byref(cfg), C.byref(h)))
        self.h = h
        self._keep = []  # host arrays must outlive the upload call only

    def close(self):
        if self.h:
            self._L.apspie_server_destroy(self.h)
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

    def add_graph(self, rowptr, col, val=None):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = None if val is None else np.ascontiguousarray(val, np.float32)
        m = len(rowptr) - 1
        csr = _Csr(m, m, len(col), _p(rowptr).value, _p(col).value if len(col) else None,
                   None if val is None else _p(val).value)
        gid = C.c_int32()
        _check(self._L.apspie_server_add_graph(self.h, C.byref(csr), C.byref(gid)))
        return gid.value

    def _admit_check(self, st):
        if st != 0:
            msg = self._L.apspie_status_str(st).decode()
            detail = self._L.apspie_server_last_error(self.h).decode()
            raise ApspieError(st, f"{msg}: {detail}" if detail else msg)

    def add_graph_coo(self, m, row, col, val=None, n=None, symmetrize=True):
        """Admit a graph built on the device from host COO triples."""
        row = np.ascontiguousarray(row, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = None if val is None else np.ascontiguousarray(val, np.float32)
        coo = _Coo(m, m if n is None else n, len(row), _p(row).value if len(row) else None,
                   _p(col).value if len(col) else None, None if val is None else _p(val).value)
        gid = C.c_int32()
        self._admit_check(self._L.apspie_server_add_graph_coo(self.h, C.byref(coo), HOST,
                                                              int(symmetrize), C.byref(gid)))
        return gid.value

    def add_graph_file(self, path):
        """Admit a graph from a binary CSR cache file."""
        gid = C.c_int32()
        self._admit_check(self._L.apspie_server_add_graph_file(self.h, str(path).encode(),
                                                               C.byref(gid)))
        return gid.value

    def submit(self, graph_id, kind, source=0):
        t = C.c_int64()
        _check(self._L.apspie_server_submit(self.h, graph_id, kind, int(source), C.byref(t)))
        return t.value

    def submit_into(self, graph_id, kind, source, out_ptr, mem):
        """A BFS / SSSP query whose result the server writes to out_ptr (m
        entries of caller memory, host or device per `mem`); finish it with
        wait_done.  A TC query ignores out_ptr (its count comes from wait)."""
        t = C.c_int64()
        _check(self._L.apspie_server_submit_into(self.h, graph_id, kind, int(source),
                                                 out_ptr, int(mem), C.byref(t)))
        return t.value

    def wait_done(self, ticket):
        """Wait for a submit_into query: latency_ms (the result is in the caller's memory)."""
        lat = C.c_double()
        _check(self._L.apspie_server_wait(self.h, ticket, None, 0, C.byref(lat)))
        return lat.value

    def wait(self, ticket, kind, m, copy=True):
        """Returns (result, latency_ms): int32[m] depths, float32[m] distances or an int.
        With copy=False a depth / distance result is a read-only view of the
        server's pinned host buffer (no copy); the buffer returns to the pool
        when the array is garbage-collected, and the view must not outlive
        the server."""
        lat = C.c_double()
        if kind == Q_TC:
            out = C.c_uint64()
            _check(self._L.apspie_server_wait(self.h, ticket, C.byref(out), 8, C.byref(lat)))
            return int(out.value), lat.value
        dtype = np.int32 if kind == Q_BFS else np.float32
        if copy:
            out = np.empty(m, dtype)
            _check(self._L.apspie_server_wait(self.h, ticket, _p(out), out.nbytes, C.byref(lat)))
            return out, lat.value
        data, nbytes = C.c_void_p(), C.c_int64()
        _check(self._L.apspie_server_wait_view(self.h, ticket, C.byref(data), C.byref(nbytes),
                                               C.byref(lat)))
        buf = (C.c_char * nbytes.value).from_address(data.value)
        view = np.frombuffer(buf, dtype=dtype, count=nbytes.value // 4)
        view.flags.writeable = False
        weakref.finalize(view, Server._release, weakref.ref(self), ticket)
        return view, lat.value

    @staticmethod
    def _release(server_ref, ticket):
        srv = server_ref()
        if srv is not None and srv.h:  # a closed server already freed its buffers
            srv._L.apspie_server_release(srv.h, ticket)

    def busy(self, graph_id, kind, reset=False):
        """(batches, queries, busy_ms) on one graph for one kind since the last reset."""
        b, q, ms = C.c_int64(), C.c_int64(), C.c_double()
        _check(self._L.apspie_server_busy(self.h, graph_id, kind, C.byref(b), C.byref(q),
                                          C.byref(ms), int(reset)))
        return b.value, q.value, ms.value

    def busy_stats(self, graph_id, kind, reset=False):
        """The busy counters split by phase: dict of batches, queries, busy_ms,
        traverse_ms, copy_ms, alloc_ms, and the attempts that ran out of
        budget or memory (retries, retry_ms: their wall time and the trims
        after them)."""
        st = _ServerStats()
        _check(self._L.apspie_server_busy_stats(self.h, graph_id, kind, C.byref(st), int(reset)))
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_}

    def drain(self):
        _check(self._L.apspie_server_drain(self.h))
