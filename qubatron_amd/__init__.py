"""qubatron_amd: MI355X-native sparse-octree ray traversal, shadow test and
shading (the per-pixel hot path of milgra/qubatron), behind the reference's
own connector interface.

Host mirror of the reference API:
  * ``HostOctree``  - ``src/qubatron/octree.c`` (create / insert_point /
    insert_path / remove_point / model_index), node numbering identical.
  * ``OctreeGLC``   - ``src/qubatron/octree_glc.c`` (init / update /
    upload_texbuffer_data), backed by the gfx950 library ``libqbt.so``.
  * ``Context``     - the explicit form: uploads, multi-view renders, batched
    ray queries, GPU octree builds, statistics.

There is no CPU fallback: without ``libqbt.so`` or a GPU, ``Context`` raises.
"""
import ctypes as C
import math

import numpy as np

from . import _ffi
from ._ffi import (QbtError, check, fptr, iptr, u8ptr, OUT_HIT, OUT_LEAF, OUT_SHADOW, OUT_RGBA8, OUT_RGBAF,
                   OUT_ISP, OUT_FLAGS, OUT_DYN, OUT_EXP, OUT_DEFAULT, OUT_ALL, RENDER_COUNT, RENDER_SURROGATE,
                   RENDER_COLLECT, MLP_INPUTS, MLP_PARAMS,
                   PX_DISCARD, PX_LEAF, PX_SHADED, PX_SPRITE, PX_PREDICTED)

__all__ = ["HostOctree", "Scene", "Context", "OctreeGLC", "QbtError", "view", "octree_from_points"]

NODE_INTS = 12


class HostOctree:
    """Flat octree on the host with the reference's node numbering
    (``src/qubatron/octree.c:55-252``)."""

    def __init__(self, basecube, levels):
        self._t = _ffi.Octree()
        bc = np.asarray(basecube, dtype=np.float32)
        check(_ffi.host().qbt_octree_init(C.byref(self._t), fptr(bc), int(levels)), "octree_init")

    def __del__(self):
        try:
            _ffi.host().qbt_octree_free(C.byref(self._t))
        except Exception:
            pass

    @classmethod
    def from_nodes(cls, nodes, basecube, levels):
        """A host octree holding a copy of an existing node array (n x 12)."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int32).reshape(-1, NODE_INTS)
        t = cls(basecube, levels)
        t.reserve(nodes.shape[0] + 1024)
        C.memmove(t._t.octs, nodes.ctypes.data, nodes.nbytes)
        t._t.len = nodes.shape[0]
        return t

    @property
    def len(self):
        return int(self._t.len)

    @property
    def levels(self):
        return int(self._t.levels)

    @property
    def basecube(self):
        return np.array(self._t.basecube[:], dtype=np.float32)

    def nodes(self):
        """Copy of the node array, shape (len, 12) int32."""
        n = self.len
        buf = C.cast(self._t.octs, C.POINTER(C.c_int32 * (n * NODE_INTS))).contents
        return np.frombuffer(buf, dtype=np.int32).reshape(n, NODE_INTS).copy()

    def nodes_view(self):
        """Zero-copy view (invalidated by the next insertion)."""
        n = self.len
        buf = C.cast(self._t.octs, C.POINTER(C.c_int32 * (n * NODE_INTS))).contents
        return np.frombuffer(buf, dtype=np.int32).reshape(n, NODE_INTS)

    def reset(self):
        _ffi.host().qbt_octree_reset(C.byref(self._t))

    def reserve(self, n):
        check(_ffi.host().qbt_octree_reserve(C.byref(self._t), int(n)), "octree_reserve")

    def insert_point(self, modind, p, index=0, octindarr=None, path=None):
        p = np.ascontiguousarray(p, dtype=np.float32)
        return check(_ffi.host().qbt_octree_insert_point(C.byref(self._t), int(index), int(modind), fptr(p),
                                                         iptr(octindarr), u8ptr(path)), "insert_point")

    def insert_path(self, modind, octs14, octs54, octs94, index=0):
        a = np.ascontiguousarray(octs14, dtype=np.int32)
        b = np.ascontiguousarray(octs54, dtype=np.int32)
        c = np.ascontiguousarray(octs94, dtype=np.int32)
        return check(_ffi.host().qbt_octree_insert_path(C.byref(self._t), int(index), int(modind), iptr(a),
                                                        iptr(b), iptr(c)), "insert_path")

    def insert_points(self, xyz, first_modind=0):
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        self.reserve(max(1024, int(xyz.shape[0] * 1.6) + 16))
        return check(_ffi.host().qbt_octree_insert_points(C.byref(self._t), fptr(xyz), xyz.shape[0],
                                                          int(first_modind)), "insert_points")

    def remove_point(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        m, o = C.c_int32(-1), C.c_int32(-1)
        check(_ffi.host().qbt_octree_remove_point(C.byref(self._t), fptr(p), C.byref(m), C.byref(o)),
              "remove_point")
        return m.value, o.value

    def model_index(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        m, o = C.c_int32(-1), C.c_int32(-1)
        check(_ffi.host().qbt_octree_model_index(C.byref(self._t), fptr(p), C.byref(m), C.byref(o)),
              "model_index")
        return m.value, o.value

    def hash(self):
        return int(_ffi.host().qbt_octree_hash(self._t.octs, self._t.len))


def octree_from_points(xyz, basecube, levels):
    t = HostOctree(basecube, levels)
    t.insert_points(xyz)
    return t


def octree_paths(xyz, basecube, levels, nthreads=8):
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    out = np.zeros((xyz.shape[0], levels), dtype=np.uint8)
    bc = np.asarray(basecube, dtype=np.float32)
    check(_ffi.host().qbt_octree_paths(fptr(bc), int(levels), fptr(xyz), xyz.shape[0], u8ptr(out), nthreads),
          "octree_paths")
    return out


class Scene:
    """Seeded synthetic scene preset (configs 1-5)."""

    def __init__(self, config, seed=0, npoints=0):
        self.spec = _ffi.SceneSpec()
        check(_ffi.host().qbt_scene_preset(C.byref(self.spec), int(config), int(seed), int(npoints)),
              "scene_preset")

    @property
    def npoints(self):
        return int(self.spec.npoints)

    @property
    def depth(self):
        return int(self.spec.depth)

    @property
    def size(self):
        return float(self.spec.size)

    @property
    def basecube(self):
        s = self.spec.size
        return np.array([0.0, s, s, s], dtype=np.float32)

    @property
    def lights(self):
        return np.array([self.spec.lights[i][:] for i in range(self.spec.nlights)], dtype=np.float32)

    def points(self, first=0, count=None, nthreads=8):
        if count is None:
            count = self.npoints - first
        pnt = np.empty((count, 3), dtype=np.float32)
        nrm = np.empty((count, 3), dtype=np.float32)
        col = np.empty((count, 3), dtype=np.float32)
        check(_ffi.host().qbt_scene_points(C.byref(self.spec), int(first), int(count), fptr(pnt), fptr(nrm),
                                           fptr(col), int(nthreads)), "scene_points")
        return pnt, nrm, col

    def poses(self, n, seed=1, margin=40.0):
        pos = np.zeros((n, 3), dtype=np.float32)
        ang = np.zeros((n, 2), dtype=np.float32)
        check(_ffi.host().qbt_scene_poses(C.byref(self.spec), int(seed), int(n), float(margin), fptr(pos),
                                          fptr(ang)), "scene_poses")
        return pos, ang

    def blocked(self, p, margin):
        p = np.ascontiguousarray(p, dtype=np.float32)
        return bool(_ffi.host().qbt_scene_blocked(C.byref(self.spec), fptr(p), float(margin)))


def view(camfp, angle, width, height, basecube, maxlevel, shoot=0):
    v = _ffi.View()
    v.camfp[:] = [float(x) for x in camfp]
    v.angle[:] = [float(angle[0]), float(angle[1])]
    v.width, v.height = float(width), float(height)
    v.shoot = int(shoot)
    v.basecube[:] = [float(x) for x in basecube]
    v.maxlevel = int(maxlevel)
    return v


_FIELDS = [  # name, out bit, dtype, components
    ("hit", OUT_HIT, np.int32, 1), ("leaf", OUT_LEAF, np.int32, 1), ("shadow", OUT_SHADOW, np.uint8, 1),
    ("rgba8", OUT_RGBA8, np.uint32, 1), ("rgbaf", OUT_RGBAF, np.float32, 4), ("isp", OUT_ISP, np.float32, 4),
    ("flags", OUT_FLAGS, np.uint8, 1), ("dyn", OUT_DYN, np.int32, 1), ("exp", OUT_EXP, np.int32, 1)]


class Context:
    """A libqbt context on one GPU (one HIP stream, its device buffers)."""

    def __init__(self, device=0, max_device_bytes=0):
        d = _ffi.dev()
        opts = _ffi.Opts(int(device), int(max_device_bytes))
        h = C.c_void_p()
        check(d.qbt_create(C.byref(opts), C.byref(h)), "qbt_create")
        self._h = h
        self._d = d
        self._views = {}

    def close(self):
        if self._h:
            self._d.qbt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def device_bytes(self):
        return int(self._d.qbt_device_bytes(self._h))

    # -- uploads (octree_glc_upload_texbuffer_data contract) --------------
    def upload(self, data, type_, size, itemsize, start, end, buftype):
        ptr = data.ctypes.data if isinstance(data, np.ndarray) else data
        return check(self._d.qbt_upload(self._h, C.c_void_p(ptr), int(type_), int(size), int(itemsize),
                                        int(start), int(end), int(buftype)), "qbt_upload")

    def upload_octree(self, nodes, dynamic=False, start_node=None, end_node=None):
        nodes = np.ascontiguousarray(nodes, dtype=np.int32).reshape(-1, NODE_INTS)
        size = nodes.shape[0] * 48
        s = 0 if start_node is None else start_node * 48
        e = size if end_node is None else end_node * 48
        bt = _ffi.BUFFER_DYNAMIC_OCTREE if dynamic else _ffi.BUFFER_STATIC_OCTREE
        return self.upload(nodes, _ffi.TYPE_INT, size, 16, s, e, bt)

    def upload_model(self, col, nrm, dynamic=False):
        col = np.ascontiguousarray(col, dtype=np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, dtype=np.float32).reshape(-1, 3)
        off = 3 if dynamic else 0
        self.upload(col, _ffi.TYPE_FLOAT, col.nbytes, 12, 0, col.nbytes, _ffi.BUFFER_STATIC_COLOR + off)
        self.upload(nrm, _ffi.TYPE_FLOAT, nrm.nbytes, 12, 0, nrm.nbytes, _ffi.BUFFER_STATIC_NORMAL + off)

    def upload_points(self, xyz, dynamic=False):
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        bt = _ffi.BUFFER_DYNAMIC_POINTS if dynamic else _ffi.BUFFER_STATIC_POINTS
        return self.upload(xyz, _ffi.TYPE_FLOAT, xyz.nbytes, 12, 0, xyz.nbytes, bt)

    def download_octree(self, nnodes, dynamic=False):
        out = np.empty((nnodes, NODE_INTS), dtype=np.int32)
        bt = _ffi.BUFFER_DYNAMIC_OCTREE if dynamic else _ffi.BUFFER_STATIC_OCTREE
        check(self._d.qbt_download(self._h, bt, C.c_void_p(out.ctypes.data), 0, nnodes * 48), "qbt_download")
        return out

    def build_octree(self, basecube, levels, npoints, dynamic=False):
        bc = np.asarray(basecube, dtype=np.float32)
        n = C.c_size_t(0)
        bt = _ffi.BUFFER_DYNAMIC_OCTREE if dynamic else _ffi.BUFFER_STATIC_OCTREE
        check(self._d.qbt_build_octree(self._h, bt, fptr(bc), int(levels), int(npoints), C.byref(n)),
              "qbt_build_octree")
        return int(n.value)

    # -- rendering ---------------------------------------------------------
    def render(self, views, lights, mask=OUT_DEFAULT):
        if isinstance(views, _ffi.View):
            views = [views]
        arr = (_ffi.View * len(views))(*views)
        lights = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 3)
        check(self._d.qbt_render(self._h, arr, len(views), fptr(lights), lights.shape[0], int(mask)), "qbt_render")
        outs = int(mask) & OUT_ALL
        if not (int(mask) & RENDER_COUNT):
            outs &= ~OUT_EXP  # expansion counts exist only in counting renders
        for i, v in enumerate(views):
            self._views[i] = (int(v.width), int(v.height), outs)

    def render_noalloc(self, views, lights, frames, mask=OUT_DEFAULT):
        """qbt_render_noalloc: render into caller-owned device frames, no
        device allocation. frames: one dict per view mapping output names of
        _FIELDS to device addresses (ints)."""
        if isinstance(views, _ffi.View):
            views = [views]
        arr = (_ffi.View * len(views))(*views)
        fr = (_ffi.Frame * len(views))()
        for i, (v, f) in enumerate(zip(views, frames)):
            fr[i].w, fr[i].h = int(v.width), int(v.height)
            for name, ptr in f.items():
                setattr(fr[i], name, int(ptr))
        lights = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 3)
        check(self._d.qbt_render_noalloc(self._h, arr, len(views), fptr(lights), lights.shape[0], int(mask), fr),
              "qbt_render_noalloc")

    def update(self, width, height, position, angle, lighta, quality, maxlevel, basesize, shoot):
        pos = np.asarray(position, dtype=np.float32)
        ang = np.asarray(angle, dtype=np.float32)
        check(self._d.qbt_update(self._h, float(width), float(height), fptr(pos), fptr(ang), float(lighta),
                                 int(quality), int(maxlevel), float(basesize), int(shoot)), "qbt_update")
        q = float(np.float32(quality) / np.float32(2.0))
        ow = np.float32(float(np.float32(width)) / (6.0 - q))
        oh = np.float32(float(np.float32(height)) / (6.0 - q))
        self._views[0] = (int(ow), int(oh), OUT_DEFAULT)

    def read_frame(self, view_index=0, fields=None):
        w, h, mask = self._views[view_index]
        fr = _ffi.Frame()
        out = {}
        for name, bit, dt, comp in _FIELDS:
            if not (mask & bit) or (fields is not None and name not in fields):
                continue
            a = np.empty((h, w, comp) if comp > 1 else (h, w), dtype=dt)
            out[name] = a
            setattr(fr, name, a.ctypes.data)
        check(self._d.qbt_read_frame(self._h, int(view_index), C.byref(fr)), "qbt_read_frame")
        return out

    def stats(self):
        st = _ffi.Stats()
        rc = self._d.qbt_get_stats(self._h, C.byref(st))
        check(rc, "qbt_get_stats")
        return {k: getattr(st, k) for k, _ in _ffi.Stats._fields_}

    def synchronize(self):
        check(self._d.qbt_synchronize(self._h), "qbt_synchronize")

    def layout_stats(self, dynamic=False):
        """(relayouts, patches, records) of the compact layout of a tree."""
        r, p, n = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        bt = _ffi.BUFFER_DYNAMIC_OCTREE if dynamic else _ffi.BUFFER_STATIC_OCTREE
        check(self._d.qbt_layout_stats(self._h, bt, C.byref(r), C.byref(p), C.byref(n)), "qbt_layout_stats")
        return int(r.value), int(p.value), int(n.value)

    def profile_begin(self):
        check(self._d.qbt_profile_begin(self._h), "qbt_profile_begin")

    def profile_end(self):
        p = _ffi.Profile()
        check(self._d.qbt_profile_end(self._h, C.byref(p)), "qbt_profile_end")
        out = {k: getattr(p, k) for k, _ in _ffi.Profile._fields_ if k != "reserved"}
        out["kind_launches"] = list(p.kind_launches)
        out["kind_ms"] = list(p.kind_ms)
        return out

    TUNING_FIELDS = ("path", "block_threads", "refill", "layout", "stack", "reorder", "graph", "strip", "merge")

    def tuning(self):
        t = _ffi.Tuning()
        check(self._d.qbt_get_tuning(self._h, C.byref(t)), "qbt_get_tuning")
        return {k: getattr(t, k) for k in self.TUNING_FIELDS}

    def set_tuning(self, **kw):
        """Set any of TUNING_FIELDS (None leaves a field as it is)."""
        t = _ffi.Tuning()
        check(self._d.qbt_get_tuning(self._h, C.byref(t)), "qbt_get_tuning")
        for k, v in kw.items():
            if k not in self.TUNING_FIELDS:
                raise TypeError(f"unknown tuning field {k!r}")
            if v is not None:
                setattr(t, k, int(v))
        check(self._d.qbt_set_tuning(self._h, C.byref(t)), "qbt_set_tuning")

    def autotune(self, views, lights, mask=OUT_DEFAULT, reps=3):
        """Run the library's autotuner on these views; returns (best tuning,
        list of (tuning, ms) for every candidate timed)."""
        if isinstance(views, _ffi.View):
            views = [views]
        arr = (_ffi.View * len(views))(*views)
        lights = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 3)
        res = (_ffi.TuneResult * 64)()
        n = C.c_int(0)
        check(self._d.qbt_autotune(self._h, arr, len(views), fptr(lights), lights.shape[0], int(mask), int(reps),
                                   res, 64, C.byref(n)), "qbt_autotune")
        for i, v in enumerate(views):
            self._views[i] = (int(v.width), int(v.height), int(mask) & OUT_ALL & ~OUT_EXP)
        table = [({k: getattr(r.tuning, k) for k in self.TUNING_FIELDS}, r.ms) for r in res[:min(n.value, 64)]]
        return self.tuning(), table

    def copy_bandwidth(self, nbytes=1 << 30, reps=10):
        ms, gbps = C.c_double(0), C.c_double(0)
        check(self._d.qbt_copy_bandwidth(self._h, int(nbytes), int(reps), C.byref(ms), C.byref(gbps)),
              "qbt_copy_bandwidth")
        return ms.value, gbps.value

    # -- learned shadow surrogate ------------------------------------------
    def surrogate_init(self, seed=0):
        check(self._d.qbt_surrogate_init(self._h, int(seed)), "qbt_surrogate_init")

    def surrogate_params(self):
        out = np.empty(_ffi.MLP_PARAMS, dtype=np.float32)
        check(self._d.qbt_surrogate_get_params(self._h, fptr(out), out.size), "qbt_surrogate_get_params")
        return out

    def surrogate_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        check(self._d.qbt_surrogate_set_params(self._h, fptr(p), p.size), "qbt_surrogate_set_params")

    def surrogate_set_band(self, lo=0.1, hi=0.9):
        check(self._d.qbt_surrogate_set_band(self._h, float(lo), float(hi)), "qbt_surrogate_set_band")

    def surrogate_reserve(self, nsamples, stride=1):
        check(self._d.qbt_surrogate_reserve(self._h, int(nsamples), int(stride)), "qbt_surrogate_reserve")

    def surrogate_samples(self):
        """(samples held, of which lit)."""
        n, lit = C.c_size_t(0), C.c_size_t(0)
        check(self._d.qbt_surrogate_samples(self._h, C.byref(n), C.byref(lit)), "qbt_surrogate_samples")
        return int(n.value), int(lit.value)

    def surrogate_read_samples(self, start=0, count=None):
        """Samples as an (n, 10) array: hit point, normal, light, label."""
        if count is None:
            count = self.surrogate_samples()[0] - start
        out = np.empty((count, 10), dtype=np.float32)
        check(self._d.qbt_surrogate_read_samples(self._h, int(start), int(count), fptr(out)),
              "qbt_surrogate_read_samples")
        return out

    def surrogate_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, _ffi.MLP_INPUTS)
        p = np.empty(x.shape[0], dtype=np.float32)
        check(self._d.qbt_surrogate_forward(self._h, fptr(x), x.shape[0], fptr(p)), "qbt_surrogate_forward")
        return p

    def surrogate_predict(self, isp, nrm, light, basecube):
        isp = np.ascontiguousarray(isp, dtype=np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, dtype=np.float32).reshape(-1, 3)
        light = np.ascontiguousarray(light, dtype=np.float32).reshape(-1, 3)
        bc = np.asarray(basecube, dtype=np.float32)
        p = np.empty(isp.shape[0], dtype=np.float32)
        check(self._d.qbt_surrogate_predict(self._h, fptr(isp), fptr(nrm), fptr(light), isp.shape[0], fptr(bc),
                                            fptr(p)), "qbt_surrogate_predict")
        return p

    def surrogate_grad(self, x, y):
        """(mean BCE loss, gradient in flat parameter order)."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, _ffi.MLP_INPUTS)
        y = np.ascontiguousarray(y, dtype=np.float32).ravel()
        g = np.empty(_ffi.MLP_PARAMS, dtype=np.float32)
        loss = C.c_float(0.0)
        check(self._d.qbt_surrogate_grad(self._h, fptr(x), fptr(y), x.shape[0], fptr(g), C.byref(loss)),
              "qbt_surrogate_grad")
        return float(loss.value), g

    def surrogate_step(self, x, y, lr):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, _ffi.MLP_INPUTS)
        y = np.ascontiguousarray(y, dtype=np.float32).ravel()
        loss = C.c_float(0.0)
        check(self._d.qbt_surrogate_step(self._h, fptr(x), fptr(y), x.shape[0], float(lr), C.byref(loss)),
              "qbt_surrogate_step")
        return float(loss.value)

    def surrogate_train(self, basecube, steps, batch=4096, lr=1e-3, seed=0):
        bc = np.asarray(basecube, dtype=np.float32)
        loss = C.c_float(0.0)
        check(self._d.qbt_surrogate_train(self._h, fptr(bc), int(steps), int(batch), float(lr), int(seed),
                                          C.byref(loss)), "qbt_surrogate_train")
        return float(loss.value)

    def trace_rays(self, origins, dirs, basecube, maxlevel, semantics=0):
        o = np.ascontiguousarray(origins, dtype=np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        n = o.shape[0]
        bc = np.asarray(basecube, dtype=np.float32)
        status = np.empty(n, np.int32)
        model = np.empty(n, np.int32)
        leaf = np.empty(n, np.int32)
        isp = np.empty((n, 4), np.float32)
        tlf = np.empty((n, 4), np.float32)
        check(self._d.qbt_trace_rays(self._h, fptr(o), fptr(d), n, int(semantics), fptr(bc), int(maxlevel),
                                     iptr(status), iptr(model), iptr(leaf), fptr(isp), fptr(tlf)),
              "qbt_trace_rays")
        return dict(status=status, model=model, leaf=leaf, isp=isp, tlf=tlf)

    def particle_update(self, pos, spd, maxlevel, basesize):
        """The reference's particle pass (``particle_glc_update``,
        ``src/qubatron/particle_glc.c:118-156``): returns (pos_out, spd_out),
        n x 3 float32 each."""
        p = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        v = np.ascontiguousarray(spd, dtype=np.float32).reshape(-1, 3)
        po, vo = np.empty_like(p), np.empty_like(v)
        check(self._d.qbt_particle_update(self._h, fptr(p), fptr(v), p.shape[0], int(maxlevel), float(basesize),
                                          fptr(po), fptr(vo)), "qbt_particle_update")
        return po, vo


class OctreeGLC:
    """Python mirror of ``octree_glc_t`` (``src/qubatron/octree_glc.c:65-85``):
    the same three calls, executed by libqbt on the GPU."""

    STATIC_COLOR, STATIC_NORMAL, STATIC_OCTREE = 0, 1, 2
    DYNAMIC_COLOR, DYNAMIC_NORMAL, DYNAMIC_OCTREE = 3, 4, 5

    def __init__(self, base_path="", device=0):
        self.ctx = Context(device)

    def upload_texbuffer_data(self, data, type_, size, itemsize, start, end, buftype):
        return self.ctx.upload(data, type_, size, itemsize, start, end, buftype)

    def update(self, width, height, position, angle, lighta, quality, maxlevel, basesize, shoot):
        return self.ctx.update(width, height, position, angle, lighta, quality, maxlevel, basesize, shoot)

    def frame(self):
        return self.ctx.read_frame(0)


class ParticleGLC:
    """Python mirror of ``particle_glc_t`` (``src/qubatron/particle_glc.c``):
    ``alloc_in`` takes the particle positions and velocities, ``alloc_out``
    sizes the outputs, ``update`` runs the particle pass on the GPU and leaves
    the results in ``pos_out`` / ``spd_out`` (float32, n x 3), as the
    reference's read-back does (``:147-150``)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.pos_in = self.spd_in = None
        self.pos_out = self.spd_out = None

    def alloc_in(self, posdata, spddata, size):
        n = size // 12
        self.pos_in = np.ascontiguousarray(posdata, dtype=np.float32).reshape(-1, 3)[:n]
        self.spd_in = np.ascontiguousarray(spddata, dtype=np.float32).reshape(-1, 3)[:n]

    def alloc_out(self, size):
        n = size // 12
        self.pos_out = np.zeros((n, 3), np.float32)
        self.spd_out = np.zeros((n, 3), np.float32)

    def update(self, model_count, maxlevel, basesize):
        p, v = self.ctx.particle_update(self.pos_in[:model_count], self.spd_in[:model_count], maxlevel, basesize)
        self.pos_out[:model_count] = p
        self.spd_out[:model_count] = v
