"""The binary CSR cache format on the host (apspie_csr_write / apspie_csr_read;
no GPU): exact round trips of the reference's fixtures and R-MAT-14, the
documented header layout, and refusal of truncated, corrupted and foreign
files."""
import numpy as np
import pytest

import apspie_amd as A
from oracle import oracle as O
from tests import graphs as T


def roundtrip(tmp_path, rp, col, val, n=None):
    p = tmp_path / "g.csr"
    A.csr_write(p, rp, col, val, n)
    return p, A.csr_read(p)


@pytest.mark.parametrize("name", T.FIXTURES)
def test_fixture_roundtrip(tmp_path, name):
    g = T.fixture(name)
    p, (rp, col, val, n) = roundtrip(tmp_path, g.rowptr32, g.col, g.w)
    assert np.array_equal(rp, g.rowptr32) and np.array_equal(col, g.col)
    assert np.array_equal(val, g.w) and n == g.m


def test_rmat14_roundtrip_and_layout(tmp_path):
    rp, col = A.gen_rmat(14)
    w = A.gen_weights(rp, col)
    p, (rp2, col2, val2, n) = roundtrip(tmp_path, rp, col, w)
    assert np.array_equal(rp2, rp) and np.array_equal(col2, col) and np.array_equal(val2, w)
    raw = p.read_bytes()
    assert raw[:8] == b"APSPCSR1"
    ver, flags = np.frombuffer(raw[8:16], np.uint32)
    m, nn, nnz = np.frombuffer(raw[16:40], np.int64)
    assert (ver, flags, m, nn, nnz) == (1, 1, len(rp) - 1, len(rp) - 1, len(col))
    assert raw[48:64] == bytes(16)
    body = np.frombuffer(raw[64:], np.uint8)
    assert len(body) == 4 * (m + 1) + 8 * nnz
    assert np.array_equal(body[:4 * (m + 1)].view(np.int32), rp)
    # the checksum: FNV-1a constants over 8-byte words, zero-padded tail, h ^ h >> 29
    words = np.frombuffer(raw[64:] + bytes(-len(body) % 8), np.uint64)
    h = 0xcbf29ce484222325
    for x in words[:2000].tolist():  # the prefix, then the rest vectorised below
        h = ((h ^ x) * 0x100000001b3) & (2 ** 64 - 1)
    h = fnv_words(words[2000:], h)
    assert (h ^ (h >> 29)) == int(np.frombuffer(raw[40:48], np.uint64)[0])


def fnv_words(words, h):
    for x in words.tolist():
        h = ((h ^ x) * 0x100000001b3) & (2 ** 64 - 1)
    return h


def test_pattern_only_and_empty(tmp_path):
    g = T.fixture("test_mesh")
    p, (rp, col, val, n) = roundtrip(tmp_path, g.rowptr32, g.col, None)
    assert val is None and np.array_equal(col, g.col)
    p, (rp, col, val, n) = roundtrip(tmp_path, np.zeros(6, np.int32), np.zeros(0, np.int32),
                                     None, n=9)
    assert len(rp) == 6 and len(col) == 0 and n == 9


def test_refuses_bad_files(tmp_path):
    rp, col = A.gen_rmat(10)
    p, _ = roundtrip(tmp_path, rp, col, A.gen_weights(rp, col))
    raw = bytearray(p.read_bytes())
    bad = tmp_path / "bad.csr"
    for mutate in (lambda b: b[:-4],                                    # truncated
                   lambda b: b[:5000] + bytes([b[5000] ^ 1]) + b[5001:],  # one bit flipped
                   lambda b: b"XPSPCSR1" + b[8:],                       # foreign magic
                   lambda b: b[:8] + bytes([2]) + b[9:],                # unknown version
                   lambda b: b[:32] + (10 ** 12).to_bytes(8, "little") + b[40:]):  # nnz
        bad.write_bytes(bytes(mutate(bytes(raw))))
        with pytest.raises(A.ApspieError):
            A.csr_read(bad)
    with pytest.raises(A.ApspieError):
        A.csr_read(tmp_path / "missing.csr")


def test_mtx_to_cache_matches_oracle_csr(tmp_path):
    """.mtx -> oracle CSR (symmetrized) -> cache -> read: the arrays survive."""
    path = T.GOLDEN_DIR / "mtx" / "chesapeake.mtx"
    m, rp, c, _ = O.mtx_graph(path.read_text())
    p, (rp2, col2, val2, n) = roundtrip(tmp_path, rp.astype(np.int32), c, None)
    assert np.array_equal(rp2, rp.astype(np.int32)) and np.array_equal(col2, c)
