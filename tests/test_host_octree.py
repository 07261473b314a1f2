"""Host octree builder (libqbt_host): node numbering and edits of
src/qubatron/octree.c, pinned by closed forms and by an independent
prefix-dictionary construction."""
import numpy as np
import pytest

import qubatron_amd as q

BC = np.array([0.0, 1800.0, 1800.0, 1800.0], np.float32)


def _grid_octets(p, levels, edge=1800.0):
    """Octets from integer cell coordinates (float64): x bit of level l is bit
    L-1-l of floor(x 2^L / w); octet bit1 / bit2 are set for the LOW y / z
    half (octree.c:104-109 adds 2 / 4 when the y / z sub-cell is even)."""
    c = np.floor(np.asarray(p, np.float64) * (2 ** levels) / edge).astype(np.int64)
    out = []
    for l in range(levels):
        sh = levels - 1 - l
        xb, yb, zb = (c[0] >> sh) & 1, (c[1] >> sh) & 1, (c[2] >> sh) & 1
        out.append(int(xb + (2 if yb == 0 else 0) + (4 if zb == 0 else 0)))
    return out


def _dict_build(points, levels, edge=1800.0):
    """Independent construction: nodes are numbered by first visit of their
    cell prefix in point order; slot 8 is the first visitor."""
    index = {(): 0}
    model = {(): 0}
    order = [()]
    for i, p in enumerate(points):
        path = tuple(_grid_octets(p, levels, edge))
        for d in range(1, levels + 1):
            pre = path[:d]
            if pre not in index:
                index[pre] = len(order)
                model[pre] = i
                order.append(pre)
    nodes = np.zeros((len(order), 12), np.int32)
    for pre in order:
        nodes[index[pre], 8] = model[pre]
        if pre:
            nodes[index[pre[:-1]], pre[-1]] = index[pre]
    return nodes


def _interior_points(rng, n, levels, edge=1800.0):
    """Points at least 1e-3 cells away from every cell boundary, so float32
    division and the float64 grid agree."""
    u = edge / 2 ** levels
    cells = rng.integers(0, 2 ** levels, (n, 3))
    frac = rng.uniform(0.01, 0.99, (n, 3))
    return ((cells + frac) * u).astype(np.float32)


@pytest.mark.parametrize("levels,n", [(3, 200), (6, 3000), (10, 20000)])
def test_matches_prefix_dictionary(levels, n):
    rng = np.random.default_rng(levels)
    pts = _interior_points(rng, n, levels)
    pts = np.concatenate([pts, pts[: n // 10]])  # duplicates keep the first inserter
    t = q.octree_from_points(pts, BC, levels)
    ref = _dict_build(pts, levels)
    assert t.len == ref.shape[0]
    assert np.array_equal(t.nodes(), ref)


def test_single_point_closed_form():
    p = np.array([10.0, 690.0, 10.0], np.float32)
    t = q.HostOctree(BC, 12)
    path = np.zeros(12, np.uint8)
    t.insert_point(7, p, path=path)
    assert t.len == 13
    nodes = t.nodes()
    assert list(path) == _grid_octets(p, 12)
    for l in range(12):  # node l+1 is the child of node l along the path
        assert nodes[l, path[l]] == l + 1
        assert np.count_nonzero(nodes[l, :8]) == 1
        assert nodes[l + 1, 8] == 7
    assert np.all(nodes[12, :8] == 0) and np.all(nodes[:, 9:] == 0)


def test_first_inserter_keeps_leaf():
    t = q.HostOctree(BC, 8)
    p = np.array([500.5, 500.5, 500.5], np.float32)
    t.insert_point(5, p)
    n1 = t.len
    t.insert_point(9, p + np.float32(0.01))  # same leaf
    assert t.len == n1
    m, parent = t.model_index(p)
    assert m == 5


def test_wrap_at_edge():
    """p == edge truncates to an even cell at every level: octet 6 (low y,
    low z, low x) all the way down (octree.c:104-109, quirk Q10)."""
    t = q.HostOctree(BC, 5)
    path = np.zeros(5, np.uint8)
    t.insert_point(0, np.array([1800.0, 1800.0, 1800.0], np.float32), path=path)
    assert list(path) == [6] * 5


def test_negative_coordinates():
    t = q.HostOctree(BC, 4)
    path = np.zeros(4, np.uint8)
    t.insert_point(0, np.array([-0.5, 10.0, 10.0], np.float32), path=path)  # truncates to cell 0
    assert path[0] == 6
    with pytest.raises(q.QbtError):  # x bit -1 with odd y, z cells: octet -1
        t.insert_point(1, np.array([-1000.0, 1000.0, 1000.0], np.float32))
    with pytest.raises(q.QbtError):
        t.insert_point(1, np.array([np.nan, 1.0, 1.0], np.float32))


def test_insert_path_equals_insert_point():
    rng = np.random.default_rng(1)
    pts = _interior_points(rng, 2000, 12)
    a = q.octree_from_points(pts, BC, 12)
    b = q.HostOctree(BC, 12)
    paths = q.octree_paths(pts, BC, 12).astype(np.int32)
    for i, pth in enumerate(paths):
        b.insert_path(i, pth[0:4], pth[4:8], pth[8:12])
    assert np.array_equal(a.nodes(), b.nodes())


def test_remove_point_and_octindarr():
    rng = np.random.default_rng(2)
    pts = _interior_points(rng, 500, 8)
    t = q.octree_from_points(pts, BC, 8)
    before = t.nodes()
    m, parent = t.remove_point(pts[17])
    after = t.nodes()
    assert m == 17 or before[before[:, 8] == m].size  # the leaf's first inserter
    diff = np.argwhere(before != after)
    assert diff.shape[0] == 1 and diff[0, 0] == parent  # exactly one slot zeroed
    assert after[parent, diff[0, 1]] == 0
    m2, _ = t.remove_point(pts[17])  # path gone: untouched
    assert m2 == -1
    # re-insert: octindarr names the parent whose slot is set and the new node
    oa = np.zeros(9, np.int32)
    n0 = t.len
    t.insert_point(17, pts[17], octindarr=oa)
    assert t.len == n0 + 1
    assert oa[8] == n0 and oa[7] == parent


def test_model_index_readonly():
    rng = np.random.default_rng(3)
    pts = _interior_points(rng, 300, 6)
    t = q.octree_from_points(pts, BC, 6)
    before = t.nodes()
    for i in (0, 5, 299):
        m, o = t.model_index(pts[i])
        assert before[before[o][:8][before[o][:8] > 0]][:, 8].tolist().count(m) >= 1
    assert np.array_equal(before, t.nodes())


def test_hash_and_reset():
    t = q.HostOctree(BC, 6)
    h0 = t.hash()
    t.insert_point(0, np.array([1.0, 2.0, 3.0], np.float32))
    assert t.hash() != h0
    t.reset()
    assert t.len == 1 and t.hash() == h0
