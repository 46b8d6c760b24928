"""Eigenpair files on the host (no GPU): the raw-binary format that stands in
for the reference's HDF5 storage example (examples/hdf5_storage.py:5-13).
Round trips of values and vectors, files spanning several checksum blocks,
and the errors a damaged or foreign file must raise."""
import os

import numpy as np
import pytest


def _vecs(n, d, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, d)) + 1j * rng.standard_normal((n, d))


def test_eigenpairs_round_trip(qk, tmp_path):
    p = tmp_path / "pairs.bin"
    vals = np.array([-3.5, -1.25, 0.0, 2.0])
    vecs = _vecs(4, 1000, 1)
    qk.save_eigenpairs(p, vals, vecs)
    v2, x2 = qk.load_eigenpairs(p)
    assert np.array_equal(v2, vals) and np.array_equal(x2, vecs)
    # values only
    v3, x3 = qk.load_eigenpairs(p, vectors=False)
    assert np.array_equal(v3, vals) and x3 is None
    q = tmp_path / "vals.bin"
    qk.save_eigenpairs(q, vals)
    v4, x4 = qk.load_eigenpairs(q)
    assert np.array_equal(v4, vals) and x4 is None


def test_eigenpairs_many_blocks(qk, tmp_path):
    """3 vectors of 100,003 complex values: 4.8 MB, several 1 MiB checksum
    blocks and a ragged last one; bit patterns (NaN payloads, -0.0,
    denormals) come back unchanged."""
    p = tmp_path / "big.bin"
    vecs = _vecs(3, 100003, 2)
    vecs[0, 5] = complex(-0.0, 5e-324)
    vecs[1, 7] = complex(np.inf, -np.inf)
    vals = np.array([1.0, np.nextafter(1.0, 2.0), -0.0])
    qk.save_eigenpairs(p, vals, vecs)
    v2, x2 = qk.load_eigenpairs(p)
    assert v2.view(np.uint64).tolist() == vals.view(np.uint64).tolist()
    assert np.array_equal(x2.view(np.uint64), vecs.view(np.uint64))


def test_eigenpairs_damaged_files(qk, tmp_path):
    p = tmp_path / "pairs.bin"
    qk.save_eigenpairs(p, [1.0, 2.0], _vecs(2, 5000, 3))
    raw = bytearray(p.read_bytes())
    # one flipped bit inside the vectors: the checksum catches it
    bad = bytearray(raw)
    bad[64 + 16 + 12345] ^= 0x10
    (tmp_path / "flip.bin").write_bytes(bytes(bad))
    with pytest.raises(RuntimeError, match="checksum"):
        qk.load_eigenpairs(tmp_path / "flip.bin")
    # truncated
    (tmp_path / "short.bin").write_bytes(bytes(raw[:-100]))
    with pytest.raises(RuntimeError, match="truncated"):
        qk.load_eigenpairs(tmp_path / "short.bin")
    # not a qkrylov file
    (tmp_path / "junk.bin").write_bytes(b"\0" * 200)
    with pytest.raises(RuntimeError, match="not a qkrylov file"):
        qk.load_eigenpairs(tmp_path / "junk.bin")
    # missing
    with pytest.raises(RuntimeError, match="cannot open"):
        qk.load_eigenpairs(tmp_path / "none.bin")


def test_save_result_forms(qk, tmp_path):
    """save_result stores what the reference example stores: the Lanczos
    energy and eigenvector, the Davidson eigenvalues and eigenvectors."""
    lr = qk.LanczosResult(-2.0, _vecs(1, 16, 4)[0], 5, True)
    qk.save_result(tmp_path / "l.bin", lr)
    v, x = qk.load_eigenpairs(tmp_path / "l.bin")
    assert v.tolist() == [-2.0] and np.array_equal(x[0], lr.eigenvector)
    dr = qk.DavidsonResult([-3.0, -1.0], list(_vecs(2, 16, 5)), 9, True)
    qk.save_result(tmp_path / "d.bin", dr)
    v, x = qk.load_eigenpairs(tmp_path / "d.bin")
    assert v.tolist() == [-3.0, -1.0] and np.array_equal(x, np.asarray(dr.eigenvectors))
    with pytest.raises(TypeError):
        qk.save_result(tmp_path / "x.bin", 3.0)


def test_no_partial_file_left(qk, tmp_path):
    """Files are written under a temporary name and renamed."""
    p = tmp_path / "pairs.bin"
    qk.save_eigenpairs(p, [1.0], _vecs(1, 10, 6))
    assert sorted(os.listdir(tmp_path)) == ["pairs.bin"]
