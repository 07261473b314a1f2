"""The particle pass on the GPU (qbt_particle_update, the reference's
particle_glc_update) and the particle shader's copy of the traversal
(qbt_trace_rays semantics 2) against the oracle, bit for bit: positions and
velocities out, and for the traces status, leaf, model index, entry point and
leaf cube. Cases: random particles (moving, dead, falling straight down with
zero x/z speed, falling out below y = -10), particles just above surfaces
that stop at a leaf corner, and rays aimed at the children that the
particle shader's slot address loses (nodes with 3 i mod 8192 == 8191).
"""
import numpy as np
import pytest

import qubatron_amd as q
from oracle import pyoracle as po
from tests import helpers as H
from tests.test_particles import BC, node_cells, quirk_nodes

pytestmark = pytest.mark.gpu


@pytest.fixture
def scene_ctx(gpu_ctx):
    sc, pnt, nrm, col, nodes = H.scene(1, npoints=300000)
    gpu_ctx.upload_octree(nodes)
    gpu_ctx.upload_model(col, nrm)
    gpu_ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)
    yield gpu_ctx, sc, pnt, nodes, po.OracleScene(nodes, col, nrm, sc.basecube, sc.depth)


def _particles(rng, pnt, n):
    pos = rng.uniform(5.0, 1795.0, (n, 3)).astype(np.float32)
    spd = rng.uniform(-15.0, 15.0, (n, 3)).astype(np.float32)
    k = n // 5
    spd[:k, 0] = spd[:k, 2] = 0.0                     # straight down or up
    spd[k:2 * k, 0] = -100000.0                       # dead
    near = pnt[rng.integers(0, pnt.shape[0], k)]      # just off a surface point
    pos[2 * k:3 * k] = near + rng.uniform(-3.0, 3.0, (k, 3)).astype(np.float32)
    spd[2 * k:3 * k] = rng.uniform(-2.0, 2.0, (k, 3)).astype(np.float32)
    pos[3 * k:3 * k + 50, 1] = rng.uniform(-9.0, 2.0, 50).astype(np.float32)  # about to fall out
    return pos, spd


def _same(a, b, what):
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), \
        f"{what}: {np.count_nonzero(np.any(a.view(np.uint32) != b.view(np.uint32), axis=-1))} rows differ"


@pytest.mark.parametrize("seed", [1, 2])
def test_particle_update_equals_oracle(scene_ctx, seed):
    ctx, sc, pnt, nodes, osc = scene_ctx
    rng = np.random.default_rng(seed)
    pos, spd = _particles(rng, pnt, 20000)
    print(f"[case] particle update seed {seed}, {pos.shape[0]} particles", flush=True)
    gp, gv = ctx.particle_update(pos, spd, sc.depth, float(sc.basecube[3]))
    op, ov = osc.particle_step(pos, spd)
    _same(gp, op, "pos_out")
    _same(gv, ov, "spd_out")
    stopped = np.count_nonzero((ov[:, 0] == -100000.0) & (spd[:, 0] != -100000.0))
    assert stopped > 100


def test_particle_glc_mirror_steps(scene_ctx):
    """Three passes through the ParticleGLC mirror (the reference's
    alloc_in / alloc_out / update), each fed the previous outputs."""
    ctx, sc, pnt, nodes, osc = scene_ctx
    rng = np.random.default_rng(9)
    pos, spd = _particles(rng, pnt, 4000)
    pg = q.ParticleGLC(ctx)
    p_ref, v_ref = pos, spd
    for step in range(3):
        print(f"[case] ParticleGLC step {step}", flush=True)
        pg.alloc_in(pos, spd, pos.nbytes)
        pg.alloc_out(pos.nbytes)
        pg.update(pos.shape[0], sc.depth, float(sc.basecube[3]))
        p_ref, v_ref = osc.particle_step(p_ref, v_ref)
        _same(pg.pos_out, p_ref, f"step {step} pos")
        _same(pg.spd_out, v_ref, f"step {step} spd")
        pos, spd = pg.pos_out.copy(), pg.spd_out.copy()


def test_particle_trace_lost_slots(gpu_ctx):
    rng = np.random.default_rng(11)
    pts = rng.uniform(1.0, 1799.0, (60000, 3)).astype(np.float32)
    nodes = q.octree_from_points(pts, BC, 7).nodes()
    hi_lost, _ = quirk_nodes(nodes.shape[0])
    cells = node_cells(nodes)
    targets = []
    for i in hi_lost:
        for c in nodes[i, 4:8]:
            if c > 0:
                lv, ix, iy, iz = cells[int(c)]
                w = 1800.0 / 2 ** lv
                targets.append((ix * w + 0.5 * w, 1800.0 - iy * w - 0.5 * w, 1800.0 - iz * w - 0.5 * w))
    targets = np.array(targets, np.float32)
    n = 6000
    org = rng.uniform(1.0, 1799.0, (n, 3)).astype(np.float32)
    dirs = (targets[rng.integers(0, len(targets), n)] - org).astype(np.float32)
    dirs[:200, 0] = 0.0  # zero components: the particle defaults
    dirs[200:400, 2] = 0.0
    gpu_ctx.upload_octree(nodes)
    gpu_ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)
    print("[case] particle-semantics traces at lost slots", flush=True)
    g = gpu_ctx.trace_rays(org, dirs, BC, 7, semantics=2)
    o = po.OracleScene(nodes, None, None, BC, 7).trace_many(org, dirs, semantics=po.SEM_PARTICLE)
    sh = po.OracleScene(nodes, None, None, BC, 7).trace_many(org, dirs, semantics=po.SEM_SHADER)
    for k in ("status", "model", "leaf"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    _same(g["isp"], o["isp"], "isp")
    _same(g["tlf"], o["tlf"], "tlf")
    assert np.count_nonzero(sh["leaf"] != o["leaf"]) > 0  # the lost slots changed some results
