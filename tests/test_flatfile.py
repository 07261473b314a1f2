"""Flat model files (.pnt/.col/.nrm, src/qubatron/model.c:53-111) and the
qmc quantise / sort / dedup step (src/qubatron/qmc.c:41-327)."""
import ctypes as C

import numpy as np
import pytest

from qubatron_amd import _ffi


def test_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    n = 1234
    p, c, nr = (rng.uniform(0, 1, (n, 3)).astype(np.float32) for _ in range(3))
    paths = [str(tmp_path / f"m.{e}").encode() for e in ("pnt", "col", "nrm")]
    h = _ffi.host()
    assert h.qbt_flat_write(*paths, _ffi.fptr(p), _ffi.fptr(c), _ffi.fptr(nr), n) == 0
    assert h.qbt_flat_count(paths[0]) == n
    a, b, d = (np.zeros((n, 3), np.float32) for _ in range(3))
    assert h.qbt_flat_read(*paths, _ffi.fptr(a), _ffi.fptr(b), _ffi.fptr(d), n) == n
    assert np.array_equal(a, p) and np.array_equal(b, c) and np.array_equal(d, nr)
    # files are raw little-endian float32 triples
    assert np.array_equal(np.fromfile(paths[0].decode(), dtype="<f4").reshape(-1, 3), p)
    # too small a destination, missing files, short files
    assert h.qbt_flat_read(*paths, _ffi.fptr(a), _ffi.fptr(b), _ffi.fptr(d), n - 1) == _ffi.QBT_EINVAL
    assert h.qbt_flat_count(b"/nonexistent.pnt") == _ffi.QBT_EIO
    open(paths[1].decode(), "wb").write(b"\0" * 12)
    assert h.qbt_flat_read(*paths, _ffi.fptr(a), _ffi.fptr(b), _ffi.fptr(d), n) == _ffi.QBT_EIO


def _qmc(p, size, levels):
    p = p.astype(np.float32).copy()
    c = np.arange(p.size, dtype=np.float32).reshape(-1, 3)
    nr = -c
    rng = np.zeros((p.shape[0] * 3 + 3,), np.int32)
    nr_ = C.c_size_t(0)
    n = _ffi.host().qbt_qmc_quantize(_ffi.fptr(p), _ffi.fptr(c), _ffi.fptr(nr), p.shape[0], size, levels,
                                     _ffi.iptr(rng), C.byref(nr_))
    return n, p[:n], c[:n], rng[: 3 * nr_.value].reshape(-1, 3)


def test_qmc_quantize_known_answer():
    """size 8, levels 1: division 4, precision 2. Cells sorted by x, then y,
    then z; the first point of a cell is kept (ties keep input order); points
    outside [0, 4) cells are dropped; a range record (first kept index, x, y)
    is written whenever the (x, y) column changes, starting from a zeroed
    previous record (so a first point in column (0, 0) writes none)."""
    p = np.array([[5.0, 1.0, 1.0],   # cell (2,0,0)
                  [1.0, 1.0, 1.0],   # cell (0,0,0)
                  [1.5, 0.5, 0.2],   # cell (0,0,0) duplicate -> dropped
                  [1.0, 3.0, 7.9],   # cell (0,1,3)
                  [9.0, 1.0, 1.0],   # x cell 4: outside -> dropped
                  [-0.1, 1.0, 1.0],  # x cell -1: outside -> dropped
                  [1.0, 3.5, 1.0],   # cell (0,1,0)
                  [5.5, 1.5, 1.5]])  # cell (2,0,0) duplicate -> dropped
    n, pp, cc, rng = _qmc(p, 8.0, 1)
    assert n == 4
    assert np.array_equal(pp, np.array([[1.0, 1.0, 1.0], [1.0, 3.5, 1.0], [1.0, 3.0, 7.9], [5.0, 1.0, 1.0]],
                                       np.float32))
    assert cc[:, 0].tolist() == [3.0, 18.0, 9.0, 0.0]  # colours travel with their points
    assert rng.tolist() == [[1, 0, 1], [3, 2, 0]]


def test_qmc_then_octree_is_spatially_ordered():
    """After quantisation every leaf cell holds one point, so the octree built
    from the output has one leaf per point."""
    import qubatron_amd as q
    rng = np.random.default_rng(1)
    p = rng.uniform(0, 1800, (50000, 3)).astype(np.float32)
    n, pp, _, _ = _qmc(p, 1800.0, 7)  # cells of size 1800/256 = leaf size at depth 8
    t = q.octree_from_points(pp, [0, 1800, 1800, 1800], 8)
    leaves = (t.nodes()[:, :8] == 0).all(1).sum()
    assert leaves == n
