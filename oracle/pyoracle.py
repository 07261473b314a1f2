"""ctypes wrapper of liboracle.so, the CPU restatement of the reference
traversal (``src/qubatron/shaders/octree_fsh.c`` / ``src/qubatron/octree.c``).

TEST INFRASTRUCTURE: only tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg use it, to compare against. The product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

SEM_SHADER, SEM_CPU, SEM_PARTICLE = 0, 1, 2
HIT, MISS, DISCARD, ERR = 1, 0, -1, -2
PX_DISCARD, PX_LEAF, PX_SHADED, PX_SPRITE = 1, 2, 4, 8
MLP_IN, MLP_H = 42, 64
MLP_NPARAM = 64 * 42 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)


class Tree(C.Structure):
    _fields_ = [("octs", i32p), ("len", C.c_size_t)]


class SceneC(C.Structure):
    _fields_ = [("stat", Tree), ("dyn", Tree), ("col_s", f32p), ("nrm_s", f32p), ("npts_s", C.c_size_t),
                ("col_d", f32p), ("nrm_d", f32p), ("npts_d", C.c_size_t), ("basecube", C.c_float * 4),
                ("maxlevel", C.c_int)]


class TraceOut(C.Structure):
    _fields_ = [("status", C.c_int), ("isp", C.c_float * 4), ("tlf", C.c_float * 4), ("socti", C.c_int32),
                ("docti", C.c_int32), ("ms", C.c_int32), ("md", C.c_int32), ("col", C.c_float * 4),
                ("nrm", C.c_float * 4), ("expansions", C.c_int64), ("expansions_dyn", C.c_int64),
                ("iterations", C.c_int64)]


class ViewC(C.Structure):
    _fields_ = [("camfp", C.c_float * 3), ("angle", C.c_float * 2), ("W", C.c_float), ("H", C.c_float),
                ("shoot", C.c_int)]


class FrameC(C.Structure):
    _fields_ = [("w", C.c_int), ("h", C.c_int), ("hit", C.c_void_p), ("leaf", C.c_void_p), ("dyn", C.c_void_p),
                ("shadow", C.c_void_p), ("flags", C.c_void_p), ("rgba", C.c_void_p), ("rgba8", C.c_void_p),
                ("isp", C.c_void_p), ("expansions", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise FileNotFoundError(f"{LIB} not built (make -C oracle)")
        L = C.CDLL(LIB)
        L.orc_trace.argtypes = [C.POINTER(SceneC), f32p, f32p, C.c_int, C.POINTER(TraceOut)]
        L.orc_particle_step.argtypes = [C.POINTER(SceneC), f32p, f32p, f32p, f32p]
        L.orc_render.argtypes = [C.POINTER(SceneC), C.POINTER(ViewC), C.c_int, f32p, C.POINTER(FrameC), C.c_int,
                                 C.c_int, C.c_int]
        L.orc_camera.argtypes = [f32p, f32p, f32p, f32p]
        L.orc_camera.restype = None
        L.orc_pixel_dir.argtypes = [C.POINTER(ViewC), f32p, f32p, C.c_int, C.c_int, f32p]
        L.orc_pixel_dir.restype = None
        L.orc_pack_rgba8.argtypes = [f32p]
        L.orc_pack_rgba8.restype = C.c_uint32
        if hasattr(L, "orc_mlp_forward"):
            L.orc_mlp_features.argtypes = [f32p, f32p, f32p, f32p, f32p]
            L.orc_mlp_features.restype = None
            L.orc_bf16_round.argtypes = [C.c_float]
            L.orc_bf16_round.restype = C.c_float
            L.orc_mlp_forward.argtypes = [f32p, f32p, C.c_int, f32p, f32p, C.c_int]
            L.orc_mlp_forward.restype = None
            L.orc_mlp_backward.argtypes = [f32p, f32p, f32p, C.c_int, f32p, C.c_int]
            L.orc_mlp_backward.restype = C.c_float
            L.orc_mlp_adam.argtypes = [f32p, f32p, f32p, f32p, C.c_int, C.c_float, C.c_float, C.c_float,
                                       C.c_float]
            L.orc_mlp_adam.restype = None
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(f32p)


class OracleScene:
    """Keeps numpy arrays alive for the C scene struct."""

    def __init__(self, nodes, col=None, nrm=None, basecube=(0.0, 1800.0, 1800.0, 1800.0), maxlevel=10,
                 dyn_nodes=None, dyn_col=None, dyn_nrm=None):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.int32).reshape(-1, 12)
        self.col = None if col is None else np.ascontiguousarray(col, dtype=np.float32).reshape(-1, 3)
        self.nrm = None if nrm is None else np.ascontiguousarray(nrm, dtype=np.float32).reshape(-1, 3)
        self.dyn_nodes = None if dyn_nodes is None else np.ascontiguousarray(dyn_nodes, dtype=np.int32).reshape(-1, 12)
        self.dyn_col = None if dyn_col is None else np.ascontiguousarray(dyn_col, dtype=np.float32).reshape(-1, 3)
        self.dyn_nrm = None if dyn_nrm is None else np.ascontiguousarray(dyn_nrm, dtype=np.float32).reshape(-1, 3)
        s = SceneC()
        s.stat = Tree(self.nodes.ctypes.data_as(i32p), self.nodes.shape[0])
        if self.dyn_nodes is not None:
            s.dyn = Tree(self.dyn_nodes.ctypes.data_as(i32p), self.dyn_nodes.shape[0])
        if self.col is not None:
            s.col_s, s.nrm_s, s.npts_s = _f(self.col), _f(self.nrm), self.col.shape[0]
        if self.dyn_col is not None:
            s.col_d, s.nrm_d, s.npts_d = _f(self.dyn_col), _f(self.dyn_nrm), self.dyn_col.shape[0]
        s.basecube[:] = [float(x) for x in basecube]
        s.maxlevel = int(maxlevel)
        self.c = s

    def trace(self, pos, dirv, semantics=SEM_SHADER):
        p = np.ascontiguousarray(pos, dtype=np.float32)
        d = np.ascontiguousarray(dirv, dtype=np.float32)
        out = TraceOut()
        lib().orc_trace(C.byref(self.c), _f(p), _f(d), int(semantics), C.byref(out))
        return out

    def trace_many(self, origins, dirs, semantics=SEM_SHADER):
        o = np.ascontiguousarray(origins, dtype=np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        n = o.shape[0]
        res = dict(status=np.empty(n, np.int32), model=np.empty(n, np.int32), leaf=np.empty(n, np.int32),
                   isp=np.empty((n, 4), np.float32), tlf=np.empty((n, 4), np.float32),
                   exps=np.empty(n, np.int64))
        out = TraceOut()
        L = lib()
        for i in range(n):
            L.orc_trace(C.byref(self.c), _f(o[i]), _f(d[i]), int(semantics), C.byref(out))
            st = out.status
            res["status"][i] = st
            hit = st == HIT
            res["model"][i] = out.ms if hit else -1
            res["leaf"][i] = out.socti if hit else -1
            res["isp"][i] = out.isp[:] if hit else (0, 0, 0, 0)
            res["tlf"][i] = out.tlf[:] if hit else (0, 0, 0, 0)
            res["exps"][i] = out.expansions
        return res

    def particle_step(self, pos, spd):
        """One transform-feedback pass of the particle shader over n particles
        (particle_vsh.c:345-373): returns (pos_out, spd_out), n x 3 each."""
        p = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        v = np.ascontiguousarray(spd, dtype=np.float32).reshape(-1, 3)
        po, vo = np.empty_like(p), np.empty_like(v)
        L = lib()
        for i in range(p.shape[0]):
            if L.orc_particle_step(C.byref(self.c), _f(p[i]), _f(v[i]), _f(po[i]), _f(vo[i])) != 0:
                raise RuntimeError("orc_particle_step: node index out of range")
        return po, vo

    def render(self, camfp, angle, W, H, lights, shoot=0, rows=None, nthreads=8):
        v = ViewC()
        v.camfp[:] = [float(x) for x in camfp]
        v.angle[:] = [float(angle[0]), float(angle[1])]
        v.W, v.H, v.shoot = float(W), float(H), int(shoot)
        w, h = int(W), int(H)
        fr = FrameC()
        fr.w, fr.h = w, h
        out = dict(hit=np.full((h, w), -1, np.int32), leaf=np.full((h, w), -1, np.int32),
                   dyn=np.full((h, w), -1, np.int32), shadow=np.zeros((h, w), np.uint8),
                   flags=np.zeros((h, w), np.uint8), rgbaf=np.zeros((h, w, 4), np.float32),
                   rgba8=np.zeros((h, w), np.uint32), isp=np.zeros((h, w, 4), np.float32),
                   exp=np.zeros((h, w), np.int32))
        fr.hit, fr.leaf, fr.dyn = out["hit"].ctypes.data, out["leaf"].ctypes.data, out["dyn"].ctypes.data
        fr.shadow, fr.flags = out["shadow"].ctypes.data, out["flags"].ctypes.data
        fr.rgba, fr.rgba8 = out["rgbaf"].ctypes.data, out["rgba8"].ctypes.data
        fr.isp, fr.expansions = out["isp"].ctypes.data, out["exp"].ctypes.data
        lights = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 3)
        r0, r1 = (0, h) if rows is None else rows
        rc = lib().orc_render(C.byref(self.c), C.byref(v), lights.shape[0], _f(lights), C.byref(fr), r0, r1,
                              int(nthreads))
        if rc != 0:
            raise RuntimeError(f"orc_render failed ({rc})")
        return out


def camera(angle):
    a = np.asarray(angle, dtype=np.float32)
    qz = np.zeros(4, np.float32)
    vx = np.zeros(3, np.float32)
    qx = np.zeros(4, np.float32)
    lib().orc_camera(_f(a), _f(qz), _f(vx), _f(qx))
    return qz, vx, qx


def pixel_dirs(angle, W, H, pixels):
    """csv of the given (px, py) pixels."""
    qz, _, qx = camera(angle)
    v = ViewC()
    v.angle[:] = [float(angle[0]), float(angle[1])]
    v.W, v.H = float(W), float(H)
    out = np.zeros((len(pixels), 3), np.float32)
    for i, (px, py) in enumerate(pixels):
        lib().orc_pixel_dir(C.byref(v), _f(qz), _f(qx), int(px), int(py), _f(out[i]))
    return out


# ---------------- shadow-visibility MLP (no reference counterpart) ---------

def mlp_features(isp, nrm, light, basecube):
    isp = np.ascontiguousarray(isp, np.float32).reshape(-1, 3)
    nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
    light = np.ascontiguousarray(light, np.float32).reshape(-1, 3)
    if light.shape[0] == 1 and isp.shape[0] > 1:
        light = np.repeat(light, isp.shape[0], 0)
    bc = np.asarray(basecube, np.float32)
    out = np.zeros((isp.shape[0], MLP_IN), np.float32)
    for i in range(isp.shape[0]):
        lib().orc_mlp_features(_f(isp[i]), _f(nrm[i]), _f(light[i]), _f(bc), _f(out[i]))
    return out


def mlp_forward(params, x, bf16=True):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_IN)
    p = np.zeros(x.shape[0], np.float32)
    lg = np.zeros(x.shape[0], np.float32)
    lib().orc_mlp_forward(_f(params), _f(x), x.shape[0], _f(p), _f(lg), int(bf16))
    return p, lg


def mlp_backward(params, x, y, bf16=True):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_IN)
    y = np.ascontiguousarray(y, np.float32)
    g = np.zeros(MLP_NPARAM, np.float32)
    loss = lib().orc_mlp_backward(_f(params), _f(x), _f(y), x.shape[0], _f(g), int(bf16))
    return loss, g


def mlp_adam(params, grads, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    for a in (params, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    g = np.ascontiguousarray(grads, np.float32)
    lib().orc_mlp_adam(_f(params), _f(g), _f(m), _f(v), int(t), lr, b1, b2, eps)


def bf16_round(x):
    return float(lib().orc_bf16_round(float(x)))
