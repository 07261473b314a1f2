"""The GPU traversal code (qubatron_amd/csrc/device/trace.hpp) compiled for
the host with g++ (-ffp-contract=off) and compared with the oracle on the
CPU: every layout (direct, compact), cube arithmetic (closed form, running
sum) and tree mode (single, dual) on seeded scenes and rays, including
rays through the cube's mid-planes and the OCTTEST tie cases. Status, leaf
node, entry point bits and expansion counts must be identical. This catches
traversal bugs without a GPU; the GPU tests repeat the comparison on the
device build.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import qubatron_amd as q
from oracle import pyoracle as po
from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "hosttrace", "host_trace.cpp")


@pytest.fixture(scope="module")
def ht(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hosttrace") / "libhosttrace.so")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wno-unknown-pragmas",
           "-Wno-enum-compare", "-I" + os.path.join(ROOT, "tests", "hosttrace"),
           "-I" + os.path.join(ROOT, "qubatron_amd", "csrc", "device"), "-o", out, SRC]
    subprocess.run(cmd, check=True)
    lib = C.CDLL(out)
    f32p, i32p, i64p = C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    lib.ht_trace.argtypes = [i32p, C.c_size_t, i32p, C.c_size_t, f32p, C.c_int, C.c_int, C.c_int, f32p, f32p,
                             C.c_size_t, i32p, i32p, f32p, i64p]
    lib.ht_trace_patched.argtypes = [i32p, i32p, C.c_size_t, f32p, C.c_int, C.c_int, C.c_int, f32p, f32p,
                                     C.c_size_t, i32p, i32p, f32p, i64p]
    return lib


def host_trace(lib, nodes, basecube, maxlevel, org, dirs, layout, grid, dyn=None):
    nodes = np.ascontiguousarray(nodes, np.int32)
    org = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = org.shape[0]
    st, lf = np.empty(n, np.int32), np.empty(n, np.int32)
    isp, ex = np.empty((n, 4), np.float32), np.empty(n, np.int64)
    bc = np.asarray(basecube, np.float32)
    i32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    f32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    d = np.ascontiguousarray(dyn, np.int32) if dyn is not None else None
    rc = lib.ht_trace(i32(nodes), nodes.shape[0], i32(d) if d is not None else None, 0 if d is None else d.shape[0],
                      f32(bc), maxlevel, layout, grid, f32(org), f32(dirs), n, i32(st), i32(lf), f32(isp),
                      ex.ctypes.data_as(C.POINTER(C.c_int64)))
    assert rc == 0
    return st, lf, isp, ex


def oracle_trace(nodes, basecube, maxlevel, org, dirs, dyn=None):
    kw = {}
    if dyn is not None:
        kw = dict(dyn_nodes=dyn)
    osc = po.OracleScene(nodes, None, None, basecube, maxlevel, **kw)
    return osc.trace_many(org, dirs)


def check(lib, nodes, basecube, maxlevel, org, dirs, dyn=None, grids=(0, 1)):
    ref = oracle_trace(nodes, basecube, maxlevel, org, dirs, dyn)
    ref_exp = ref["exps"]
    for layout in (0, 1):
        for grid in grids:
            st, lf, isp, ex = host_trace(lib, nodes, basecube, maxlevel, org, dirs, layout, grid, dyn)
            what = f"layout {layout} grid {grid}"
            np.testing.assert_array_equal(st, ref["status"], err_msg=what)
            np.testing.assert_array_equal(lf, ref["leaf"], err_msg=what)
            np.testing.assert_array_equal(isp.view(np.uint32), ref["isp"].view(np.uint32), err_msg=what)
            if dyn is None:  # the oracle counts static expansions; the dual walk adds the dynamic ones
                np.testing.assert_array_equal(ex, ref_exp, err_msg=what)


def rays(rng, n, bc=(0.0, 1800.0, 1800.0, 1800.0), inside=0.7):
    lo = np.array([bc[0], bc[1] - bc[3], bc[2] - bc[3]])
    w = bc[3]
    org = np.where(rng.uniform(size=(n, 1)) < inside, lo + rng.uniform(0, w, (n, 3)),
                   lo + rng.uniform(-0.5 * w, 1.5 * w, (n, 3))).astype(np.float32)
    tgt = (lo + rng.uniform(0, w, (n, 3))).astype(np.float32)
    return org, (tgt - org).astype(np.float32)


@pytest.mark.parametrize("levels,npts,seed", [(3, 40, 1), (5, 300, 2), (7, 2000, 3), (9, 20000, 4)])
def test_random_trees(ht, levels, npts, seed):
    rng = np.random.default_rng(seed)
    bc = (0.0, 1800.0, 1800.0, 1800.0)
    pts = rng.uniform(1.0, 1799.0, (npts, 3)).astype(np.float32)
    nodes = q.octree_from_points(pts, bc, levels).nodes()
    org, dirs = rays(rng, 3000)
    check(ht, nodes, bc, levels, org, dirs)


def test_config1_scene_views(ht):
    sc, pnt, nrm, col, nodes = H.scene(1, npoints=200000)
    rng = np.random.default_rng(5)
    org, dirs = rays(rng, 4000)
    # axis rays and rays from the mid-planes (the start-on-plane quirk)
    org2 = np.array([[900, 900, 900]] * 6 + [[900, 1700, 900]] * 6, np.float32)
    d6 = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    check(ht, nodes, sc.basecube, sc.depth, np.vstack([org, org2]), np.vstack([dirs, d6, d6]))


def test_octtest_and_ties(ht):
    bc = (0.0, 1800.0, 1800.0, 1800.0)
    pts = np.array([[10, 10, 10], [1790, 10, 10], [10, 1790, 10], [10, 10, 1790], [900, 900, 900]], np.float32)
    nodes = q.octree_from_points(pts, bc, 12).nodes()
    org = np.array([[900, 900, 3000], [900, 900, 3000], [-100, 900, 900], [900, 900, 900], [10, 10, 3000]],
                   np.float32)
    tgt = np.array([[10, 10, 10], [900, 900, 900], [1790, 10, 10], [10, 10, 10], [10, 10, 10]], np.float32)
    check(ht, nodes, bc, 12, org, tgt - org)


def test_dual_tree(ht):
    rng = np.random.default_rng(8)
    bc = (0.0, 1800.0, 1800.0, 1800.0)
    a = q.octree_from_points(rng.uniform(1, 1799, (3000, 3)).astype(np.float32), bc, 7).nodes()
    b = q.octree_from_points(rng.uniform(1, 1799, (500, 3)).astype(np.float32), bc, 7).nodes()
    org, dirs = rays(rng, 3000)
    check(ht, a, bc, 7, org, dirs, dyn=b)


def test_non_grid_basecube_running_sum(ht):
    rng = np.random.default_rng(9)
    bc = (0.3, 1000.7, 1001.1, 1000.7)
    lo = np.array([bc[0], bc[1] - bc[3], bc[2] - bc[3]])
    pts = (lo + rng.uniform(0.5, bc[3] - 0.5, (5000, 3))).astype(np.float32)
    nodes = q.octree_from_points(pts, bc, 8).nodes()
    org, dirs = rays(rng, 3000, bc)
    check(ht, nodes, bc, 8, org, dirs, grids=(0,))


@pytest.mark.parametrize("seed,frac,leaves_only", [(11, 0.08, False), (12, 0.6, True)])
def test_zeroed_slots_on_the_compact_layout(ht, seed, frac, leaves_only):
    """In-place range updates that zero child slots (octree_remove_point,
    octree.c:182-218) keep the compact records and clear live bits; the
    child block keeps its slots (alloc), so the ranks of the other children
    stay. Traces on that layout (and on the direct layout of the edited array)
    equal the oracle on the edited array: status, leaf, entry point bits and
    expansion counts. Slots are zeroed at every depth (whole subtrees drop)."""
    rng = np.random.default_rng(seed)
    bc = (0.0, 1800.0, 1800.0, 1800.0)
    levels = 8
    pts = rng.uniform(1.0, 1799.0, (20000, 3)).astype(np.float32)
    nodes = q.octree_from_points(pts, bc, levels).nodes()
    live = nodes.copy()
    nz = np.argwhere(live[:, :8] > 0)
    if leaves_only:  # slots that point at leaves, as a punched hole removes them
        nz = nz[(nodes[live[nz[:, 0], nz[:, 1]], :8] == 0).all(1)]
    pick = nz[rng.uniform(size=len(nz)) < frac]
    live[pick[:, 0], pick[:, 1]] = 0
    live[rng.integers(0, live.shape[0], 500), 8] += 7  # slot 8 changes travel too
    org, dirs = rays(rng, 3000)
    ref = oracle_trace(live, bc, levels, org, dirs)
    assert (ref["status"] == po.HIT).sum() > 300
    i32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    f32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    n = org.shape[0]
    bca = np.asarray(bc, np.float32)
    for layout in (0, 1):
        for grid in (0, 1):
            st, lf = np.empty(n, np.int32), np.empty(n, np.int32)
            isp, ex = np.empty((n, 4), np.float32), np.empty(n, np.int64)
            rc = ht.ht_trace_patched(i32(nodes), i32(live), nodes.shape[0], f32(bca), levels, layout, grid,
                                     f32(org), f32(dirs), n, i32(st), i32(lf), f32(isp),
                                     ex.ctypes.data_as(C.POINTER(C.c_int64)))
            assert rc == 0
            what = f"layout {layout} grid {grid}"
            np.testing.assert_array_equal(st, ref["status"], err_msg=what)
            np.testing.assert_array_equal(lf, ref["leaf"], err_msg=what)
            np.testing.assert_array_equal(isp.view(np.uint32), ref["isp"].view(np.uint32), err_msg=what)
            np.testing.assert_array_equal(ex, ref["exps"], err_msg=what)
