"""qbt_render_noalloc (no reference counterpart; for callers that cannot
allocate device memory) against qbt_render, called through the C ABI.

The caller owns the output frames: this test allocates them with hipMalloc of
the HIP runtime libqbt.so links, as a C caller would. Two calls in a row, the
second with fewer and smaller views; each checked against qbt_render of the
same views (and the oracle) to the parity bar: hit, leaf, dynamic index,
shadow bits, flags, entry point and expansion counts bit-exact, colour within
1/255 and 1 RGBA8 unit. The context's device bytes must not change.
"""
import ctypes as C

import numpy as np
import pytest

import qubatron_amd as q
from qubatron_amd import _ffi
from tests import helpers as H

pytestmark = pytest.mark.gpu

ALL = q.OUT_ALL | q.RENDER_COUNT
_FIELDS = [("hit", np.int32, 1), ("leaf", np.int32, 1), ("shadow", np.uint8, 1), ("rgba8", np.uint32, 1),
           ("rgbaf", np.float32, 4), ("isp", np.float32, 4), ("flags", np.uint8, 1), ("dyn", np.int32, 1),
           ("exp", np.int32, 1)]


def _hip():
    _ffi.dev()  # libqbt.so is loaded, and with it its HIP runtime
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DeviceFrame:
    """Caller-owned device buffers of one view, filled with a sentinel byte."""

    def __init__(self, hip, w, h):
        self.hip, self.w, self.h, self.ptrs = hip, w, h, {}
        for name, dt, comp in _FIELDS:
            n = w * h * comp * np.dtype(dt).itemsize
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), n) == 0
            assert hip.hipMemset(p, 0xA5, n) == 0
            self.ptrs[name] = p.value

    def read(self):
        out = {}
        for name, dt, comp in _FIELDS:
            a = np.empty((self.h, self.w, comp) if comp > 1 else (self.h, self.w), dtype=dt)
            assert self.hip.hipMemcpy(a.ctypes.data, C.c_void_p(self.ptrs[name]), a.nbytes, 2) == 0  # D2H
            out[name] = a
        return out

    def free(self):
        for p in self.ptrs.values():
            self.hip.hipFree(C.c_void_p(p))
        self.ptrs = {}


def _check_call(ctx, hip, osc, views, poses, lights, what):
    """qbt_render, then qbt_render_noalloc of the same views; compare."""
    print(f"[case] {what}: {[(int(v.width), int(v.height)) for v in views]}, {len(lights)} lights", flush=True)
    ctx.render(views, lights, ALL)
    ref = [ctx.read_frame(i) for i in range(len(views))]
    st_ref = ctx.stats()
    frames = [DeviceFrame(hip, int(v.width), int(v.height)) for v in views]
    try:
        before = ctx.device_bytes()
        ctx.render_noalloc(views, lights, [f.ptrs for f in frames], ALL)
        st = ctx.stats()
        assert ctx.device_bytes() == before, "qbt_render_noalloc allocated device memory"
        got = [f.read() for f in frames]
    finally:
        for f in frames:
            f.free()
    for k in ("primary_rays", "shadow_rays", "hits", "expansions", "expansions_primary"):
        assert st[k] == st_ref[k], f"{what}: stats {k} {st[k]} != {st_ref[k]}"
    for i, (v, g, r) in enumerate(zip(views, got, ref)):
        mm = H.compare_frames(g, r)
        assert all(x == 0 for x in mm.values()), f"{what} view {i}: vs qbt_render {mm}"
        pos, ang = poses[i]
        o = osc.render(pos, ang, int(v.width), int(v.height), lights, shoot=int(v.shoot))
        mm = H.compare_frames(g, o)
        assert all(x == 0 for x in mm.values()), f"{what} view {i}: vs oracle {mm}"
        assert (g["hit"] >= 0).any()


def test_noalloc_matches_render_twice():
    sc, pnt, nrm, col, nodes = H.scene(1, npoints=200000)
    osc = H.oracle_scene(1, npoints=200000)
    hip = _hip()
    with q.Context(0) as ctx:
        ctx.upload_octree(nodes)
        ctx.upload_model(col, nrm)
        ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)
        lights2 = np.array([[900, 1700, 900], [300, 1500, 1400]], np.float32)
        pos, ang = sc.poses(3, seed=5)
        views = [q.view(pos[0], ang[0], 160, 120, sc.basecube, sc.depth),
                 q.view(pos[1], ang[1], 97, 61, sc.basecube, sc.depth, 1)]
        _check_call(ctx, hip, osc, views, [(pos[0], ang[0]), (pos[1], ang[1])], lights2, "first call")
        # second call: fewer and smaller views, one light; the scratch keeps
        # the first call's counters and sprite flags, which must not leak
        small = [q.view(pos[2], ang[2], 64, 48, sc.basecube, sc.depth)]
        _check_call(ctx, hip, osc, small, [(pos[2], ang[2])], sc.lights, "second call")


def test_noalloc_argument_checks():
    sc, pnt, nrm, col, nodes = H.scene(1, npoints=200000)
    hip = _hip()
    with q.Context(0) as ctx:
        ctx.upload_octree(nodes)
        ctx.upload_model(col, nrm)
        v = q.view((900.0, 900.0, 900.0), (np.pi, 0.0), 32, 16, sc.basecube, sc.depth)
        f = DeviceFrame(hip, 32, 16)
        try:
            d, h = _ffi.dev(), ctx.handle
            lights = np.asarray(sc.lights, np.float32)
            many = (_ffi.View * 257)(*([v] * 257))
            frs = (_ffi.Frame * 257)()
            assert d.qbt_render_noalloc(h, many, 257, _ffi.fptr(lights), 1, q.OUT_DEFAULT, frs) == _ffi.QBT_ERANGE
            one = (_ffi.View * 1)(v)
            fr = (_ffi.Frame * 1)()
            fr[0].w, fr[0].h = 32, 16
            fr[0].hit = f.ptrs["hit"]  # leaf, shadow, rgba8 missing
            assert d.qbt_render_noalloc(h, one, 1, _ffi.fptr(lights), 1, q.OUT_DEFAULT, fr) == _ffi.QBT_EINVAL
            for name, p in f.ptrs.items():
                setattr(fr[0], name, p)
            assert d.qbt_render_noalloc(h, one, 1, _ffi.fptr(lights), 1,
                                        q.OUT_DEFAULT | q.RENDER_SURROGATE, fr) == _ffi.QBT_ENOTSUP
            fr[0].w = 31  # frame size must match the view
            assert d.qbt_render_noalloc(h, one, 1, _ffi.fptr(lights), 1, q.OUT_DEFAULT, fr) == _ffi.QBT_EINVAL
            fr[0].w = 32
            assert d.qbt_render_noalloc(h, one, 1, _ffi.fptr(lights), 1, q.OUT_DEFAULT, fr) == _ffi.QBT_OK
            ctx.stats()
        finally:
            f.free()
