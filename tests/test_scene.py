"""Seeded scene generator: determinism, budgets, geometry."""
import numpy as np
import pytest

import qubatron_amd as q


def test_deterministic_and_thread_independent():
    s = q.Scene(2, seed=7, npoints=200000)
    a = s.points(nthreads=1)
    b = s.points(nthreads=8)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    c1 = s.points(first=0, count=70000)
    c2 = s.points(first=70000, count=130000)
    for x, y1, y2 in zip(a, c1, c2):
        assert np.array_equal(x, np.concatenate([y1, y2]))
    other = q.Scene(2, seed=8, npoints=200000).points()
    assert not np.array_equal(a[0], other[0])


@pytest.mark.parametrize("config", [1, 2, 3, 4, 5])
def test_presets(config):
    s = q.Scene(config, seed=0, npoints=90000)
    sp = s.spec
    assert sp.n_planes + sp.n_objects + sp.n_cluster == s.npoints
    assert s.depth == (10 if config == 1 else 12)
    assert len(s.lights) == {1: 1, 2: 1, 3: 2, 4: 4, 5: 1}[config]
    pnt, nrm, col = s.points()
    inset = sp.inset
    assert pnt.min() >= inset and pnt.max() <= s.size - inset
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-5)
    assert col.min() >= 0.0 and col.max() < 1.0
    # the first n_planes points lie on the six cube faces
    fp = pnt[: sp.n_planes]
    on = (fp == inset) | (fp == np.float32(s.size - inset))
    assert np.all(on.any(axis=1))
    if config == 3:
        assert sp.n_planes == int(0.6 * s.npoints) and sp.n_cluster == s.npoints // 9
        assert sp.nboxes == 200 and sp.nspheres == 50 and sp.nellipsoids == 15


def test_full_counts():
    assert q.Scene(1).npoints == 1_000_000
    assert q.Scene(2).npoints == 10_000_000
    assert q.Scene(3).npoints == 90_000_000


def test_lights_and_poses_are_free():
    s = q.Scene(3, seed=0, npoints=9000)
    for L in s.lights:
        assert not s.blocked(L, 20.0)
    pos, ang = s.poses(64, seed=1)
    for p in pos:
        assert not s.blocked(p, 40.0)
    assert np.all((ang[:, 0] >= 0) & (ang[:, 0] < 2 * np.pi))
    assert np.all(np.abs(ang[:, 1]) <= np.pi / 4 + 1e-6)
    pos2, _ = s.poses(64, seed=1)
    assert np.array_equal(pos, pos2)
