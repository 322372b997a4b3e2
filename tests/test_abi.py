"""CPU-side checks of the product library: it loads, exports exactly the C ABI
that include/apspie.h declares, and its host-side graph construction matches
the oracle (no GPU needed)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import apspie_amd as A
from oracle import oracle as O
from tests import graphs as T

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "apspie.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(apspie_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = A.lib()
    names = declared()
    assert len(names) >= 40
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", str(A.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = sorted({ln.split()[-1] for ln in out.splitlines()
                       if ln.split()[-1].startswith("apspie_") and " T " in ln})
    assert exported == names  # nothing extra, nothing missing


def test_abi_version_and_status_strings():
    L = A.lib()
    assert L.apspie_abi_version() == 1
    assert L.apspie_status_str(0) == b"ok"
    assert L.apspie_status_str(-4) == b"HBM budget exceeded"
    assert L.apspie_status_str(-99) == b"unknown status"


def test_null_arguments_are_rejected():
    L = A.lib()
    assert L.apspie_ctx_create(0, None) == -1
    assert L.apspie_ctx_destroy(None) == -1
    assert L.apspie_graph_destroy(None) == -1
    assert L.apspie_bfs(None, None, 0, None, 0) == -1
    assert L.apspie_tc(None, None, None, None, 0) == -1
    assert L.apspie_last_stats(None, None) == -1


@pytest.mark.parametrize("scale", [6, 10, 14])
def test_rmat_generator_matches_oracle(scale):
    rp, col = A.gen_rmat(scale)
    m, orp, oc = O.rmat_graph(scale)
    assert np.array_equal(rp, orp.astype(np.int32)) and np.array_equal(col, oc)
    w = A.gen_weights(rp, col)
    assert np.array_equal(w, O.assign_weights(m, orp, oc))


@pytest.mark.parametrize("W", [8, 32, 64, 200])
def test_lattice_generator_matches_oracle(W):
    rp, col = A.gen_lattice(W)
    m, orp, oc = O.lattice_graph(W)
    assert np.array_equal(rp, orp.astype(np.int32)) and np.array_equal(col, oc)


def test_generator_known_answers():
    g = T.GOLD["rmat10"]
    rp, col = A.gen_rmat(10)
    import hashlib
    assert hashlib.sha256(rp.tobytes()).hexdigest() == g["rowptr_sha"]
    assert hashlib.sha256(col.tobytes()).hexdigest() == g["col_sha"]


@pytest.mark.parametrize("name", T.FIXTURES)
def test_read_mtx_matches_oracle(name):
    path = T.GOLDEN_DIR / "mtx" / f"{name}.mtx"
    m, n, row, col, val = A.read_mtx(path)
    om, on, orow, ocol, oval, k, _ = O.parse_mtx(path.read_text())
    assert (m, n) == (om, on)
    assert np.array_equal(row, orow[:k]) and np.array_equal(col, ocol[:k])
    assert np.array_equal(val, oval[:k])


def test_read_mtx_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.mtx"
    p.write_text("%%MatrixMarket\n3 3 1\n4 1\n")
    with pytest.raises(A.ApspieError):
        A.read_mtx(p)
    with pytest.raises(A.ApspieError):
        A.read_mtx(tmp_path / "missing.mtx")


def test_no_gpu_is_a_clean_error():
    """Without a visible GPU the context refuses instead of falling back."""
    try:
        import torch
        have_gpu = torch.cuda.is_available()
    except Exception:
        have_gpu = False
    if have_gpu:
        pytest.skip("a GPU is visible")
    with pytest.raises(A.ApspieError) as e:
        A.Context(0)
    assert e.value.status in (-1, -3)
