"""The reference-named C++ layer (include/apspie.hpp) driven by a
reference-style program (tests/cpp/mirror_driver.cpp), checked against the
oracle on the reference's MatrixMarket fixtures."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import apspie_amd as A
from oracle import oracle as O
from tests import graphs as T

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parent / "cpp" / "mirror_driver"


def run_driver(name, src):
    path = T.GOLDEN_DIR / "mtx" / f"{name}.mtx"
    r = subprocess.run([str(DRIVER), str(path), str(src)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        tag, *vals = line.split()
        out[tag] = vals
    return out


@pytest.mark.parametrize("name", ["chesapeake", "test_bc", "test_cc", "test_mesh", "test_mis"])
def test_mirror_driver_matches_oracle(name):
    g = T.fixture(name)
    src = int(np.argmax(g.deg))
    out = run_driver(name, src)
    want, _ = O.bfs(g.m, g.rowptr, g.col, src)
    assert [int(x) for x in out["bfs"]] == want.tolist()
    assert [int(x) for x in out["spmspvBfs"]] == want.tolist()
    d = O.sssp(g.m, g.rowptr, g.col, g.w, src)
    d = np.where(d == -1, A.SSSP_INF, d).astype(np.float32)
    assert np.array_equal(np.array([float(x) for x in out["sssp"]], np.float32), d)
    mis = O.mis_greedy(g.m, g.rowptr, g.col, O.mis_priorities(g.m))
    assert [int(x) for x in out["mis"]] == mis.tolist()
    total, rows = O.tc(g.m, g.rowptr, g.col)
    assert int(out["tc_total"][0]) == total
    assert [int(x) for x in out["tc_rows"]] == rows.tolist()
    c_rp, _, _ = O.spgemm(g.m, g.m, g.m, g.rowptr, g.col, None, g.rowptr, g.col, None)
    assert int(out["spgemm_nnz"][0]) == int(c_rp[-1])
    cp, ri, cv = O.mxm(g.m, g.m, g.m, g.rowptr, g.col, g.w, g.rowptr, g.col, g.w)
    assert [int(x) for x in out["mxm_colptr"]] == cp.tolist()
    assert [int(x) for x in out["mxm_rows"]] == ri.tolist()
    assert np.array_equal(np.array([float(x) for x in out["mxm_vals"]], np.float32), cv)
    ps, pn = (int(x) for x in out["dcsc_parts"])
    pp, ci, co, ri, _ = O.dcsc(g.m, g.rowptr, g.col, None, ps, pn)
    assert [int(x) for x in out["dcsc_part_ptr"]] == pp.tolist()
    assert [int(x) for x in out["dcsc_col_ind"]] == ci.tolist()
    assert [int(x) for x in out["dcsc_col_off"]] == co.tolist()
    assert [int(x) for x in out["dcsc_row_ind"]] == ri.tolist()
    # mXm with the reference's argument list: the same C as the d_matrix form
    assert out["mxm3_colptr"] == out["mxm_colptr"]
    assert out["mxm3_rows"] == out["mxm_rows"] and out["mxm3_vals"] == out["mxm_vals"]
    # dense mXv (mXv.cuh:110): the driver's x, frontier = entries above 0.5
    i = np.arange(g.m)
    x = np.where(i % 5 == 0, i % 3 + 1, np.where(i % 7 == 1, 0.5, np.where(i % 7 == 2, 0.25, 0)))
    y, total = O.mxv_dense(g.m, g.m, g.rowptr, g.col, g.w, x.astype(np.float32))
    assert int(out["mxv_dense_total"][0]) == total
    assert np.array_equal(np.array([float(v) for v in out["mxv_dense_y"]], np.float32), y)
    assert out["mxv_dense_empty"] == ["0", "1"]  # no frontier: 0, y unchanged
    # sparse mXv (mXv.cuh:249): x = the source's row with its weights
    r0, r1 = g.rowptr[src], g.rowptr[src + 1]
    yi, yv = O.spmspv(g.m, g.rowptr, g.col, g.w, O.SR_PLUS_TIMES, g.col[r0:r1], g.w[r0:r1])
    assert int(out["mxv_sparse_nnz"][0]) == len(yi)
    assert [int(v) for v in out["mxv_sparse_ind"]] == yi.tolist()
    assert np.array_equal(np.array([float(v) for v in out["mxv_sparse_val"]], np.float32), yv)
    assert out["mxv_sparse_zero"] == ["0"]
