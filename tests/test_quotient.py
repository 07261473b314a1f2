"""The plane-test quotient of the GPU traversal (trace.hpp plane_k) against
IEEE division, on the CPU.

The reference divides: t = (c - p) / d (octree_fsh.c:67, 80, 93). The
traversal computes t from a per-ray reciprocal y = RN(1/d) as q = RN(a y),
e = RN(a - d q), t = RN(q + e y) (FMAs), and divides only outside
|d| in [2^-20, 2^100], |q| in [2^-80, 2^120] (which bound |a| below by
2^-101). Inside that range every step is exact or lands on a normal number,
so the result scales
with the exponents of a and d and depends only on their significands; the
GPU test tests/test_gpu_quotient.py checks all 2^46 significand pairs on the
device build. Here the host build of the same code (g++, -ffp-contract=off,
correctly rounded fmaf) is checked on hard operands, a full-range random
sample and a sweep of whole significand rows. Bit-exact agreement required.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "hosttrace", "quotient.cpp")


def _has_fma():
    try:
        return " fma " in open("/proc/cpuinfo").read()
    except OSError:
        return False


@pytest.fixture(scope="module")
def qc(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("quotient") / "libquotient.so")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
           "-Wno-unknown-pragmas", "-I" + os.path.join(ROOT, "tests", "hosttrace"),
           "-I" + os.path.join(ROOT, "qubatron_amd", "csrc", "device"), "-o", out, SRC]
    if _has_fma():
        cmd.insert(3, "-mfma")  # fmaf in one instruction; the other operations stay separate
    subprocess.run(cmd, check=True)
    lib = C.CDLL(out)
    f32p, i32p, u32p = C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_uint32)
    lib.qc_quotients.argtypes = [f32p, f32p, C.c_size_t, f32p, i32p]
    lib.qc_quotients.restype = None
    lib.qc_sweep.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_int]
    lib.qc_sweep.restype = C.c_uint64
    return lib


def _run(qc, a, d):
    a = np.ascontiguousarray(a, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    t = np.empty_like(a)
    fast = np.empty(a.shape, np.int32)
    f32 = lambda x: x.ctypes.data_as(C.POINTER(C.c_float))
    qc.qc_quotients(f32(a), f32(d), a.size, f32(t), fast.ctypes.data_as(C.POINTER(C.c_int32)))
    return t, fast


def _assert_ieee(qc, a, d, what):
    a = np.asarray(a, np.float32)
    d = np.asarray(d, np.float32)
    t, fast = _run(qc, a, d)
    with np.errstate(all="ignore"):
        ref = a / d
    nz = d != 0
    bad = nz & (t.view(np.uint32) != ref.view(np.uint32))
    if bad.any():
        i = np.flatnonzero(bad)[:5]
        raise AssertionError(f"{what}: {int(bad.sum())} quotients differ, e.g. a={a[i]} d={d[i]} t={t[i]} "
                             f"ieee={ref[i]}")
    # d == 0: the shader's FLT_MAX defaults are NaN here; callers reject both alike
    assert np.all(np.isnan(t[~nz]))
    return fast


def _f(bits):
    return np.asarray(bits, np.uint32).view(np.float32)


def test_hard_operands(qc):
    """Quotient significands near 2 and 1, divisors at and next to powers of
    two, |q| at both switch points (2^-80, 2^120), |d| at its switch points
    (2^-20, 2^100), |a| at the smallest the fast path admits (2^-101 with
    |d| = 2^-20), subnormal and tiny a, and exact coordinate differences."""
    one_m = np.float32(2.0) - np.float32(2.0 ** -23)  # 2 - ulp
    sig = np.concatenate([1 + np.arange(0, 64) * 2.0 ** -23, 2 - np.arange(1, 65) * 2.0 ** -23,
                          1 + np.arange(0, 64) * 2.0 ** -12, [1.5, 1.25, 1.75, 4.0 / 3, 1.1]]).astype(np.float32)
    A, D = np.meshgrid(sig, sig)
    A, D = A.ravel(), D.ravel()
    cases = [(A, D), (-A, D), (A, -D), (A * np.float32(2.0 ** 40), D * np.float32(2.0 ** -30)),
             (A * np.float32(2.0 ** -90), D * np.float32(2.0 ** 10))]
    # |q| around 2^-80 and 2^120; |d| around 2^-20 and 2^100; |a| around 2^-101
    e = np.float32(1.0 + 2.0 ** -23)
    for s in (np.float32(1.0), one_m, e):
        for q in (2.0 ** -80, 2.0 ** 120):
            for lo in (0.5, 1.0, 2.0):
                a = np.float32(q * lo) * s
                cases.append((np.array([a, -a]), np.array([np.float32(1.0), np.float32(1.0)])))
                cases.append((np.array([a * np.float32(2.0 ** 5)]), np.array([np.float32(2.0 ** 5) * e])))
        for dd in (2.0 ** -21, 2.0 ** -20, 2.0 ** -19, 2.0 ** 99, 2.0 ** 100, 2.0 ** 101):
            cases.append((np.float32([1.0, 1.5, 1e-20, 1e20]) * s, np.full(4, np.float32(dd) * s)))
        for aa in (2.0 ** -102, 2.0 ** -101, 2.0 ** -100, 2.0 ** -99, 2.0 ** -126, 2.0 ** -140):
            cases.append((np.full(4, np.float32(aa) * s), np.float32([2.0 ** -60, 2.0 ** -30, 2.0 ** -20, 1.0])))
        # q just above 2^-80 from the smallest a: d at 2^-20 and just above
        for dd in (2.0 ** -20, 2.0 ** -20 * (1 + 2.0 ** -23), 2.0 ** -19 * (1 - 2.0 ** -24)):
            a = (np.float32(2.0 ** -100) * sig[:32] * s).astype(np.float32)
            a = np.concatenate([a, -a] + [a * np.float32(2.0 ** -k) for k in range(1, 8)])
            cases.append((a, np.full(a.size, np.float32(dd))))
    sub = _f(np.concatenate([np.arange(1, 64), np.arange(0x7FFFC0, 0x800040)]))  # subnormals, the normal edge
    cases.append((sub, np.full(sub.size, np.float32(2.0 ** -60))))
    cases.append((sub, np.full(sub.size, np.float32(3.0 * 2.0 ** -90))))
    # exact coordinate differences of a 1800-cube at depth 12 against ray directions
    rng = np.random.default_rng(0)
    c = (rng.integers(0, 4097, 20000) * np.float32(1800.0 / 4096)).astype(np.float32)
    p = rng.uniform(-100, 1900, 20000).astype(np.float32)
    cases.append((c - p, rng.normal(size=20000).astype(np.float32) * np.float32(500.0)))
    cases.append((np.float32([0.0, -0.0, 1.0, -1.0]), np.float32([0.0, 0.0, 0.0, 0.0])))
    nfast = 0
    for i, (a, d) in enumerate(cases):
        nfast += int(_assert_ieee(qc, a, d, f"hard case {i}").sum())
    assert nfast > 50000  # most of these take the reciprocal path


def test_random_full_range(qc):
    """4 M operand pairs with uniformly random bit patterns (every exponent,
    both signs, finite a and d)."""
    rng = np.random.default_rng(1)
    n = 1 << 22
    a = _f(rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32))
    d = _f(rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32))
    keep = np.isfinite(a) & np.isfinite(d)
    fast = _assert_ieee(qc, a[keep], d[keep], "random bits")
    assert fast.sum() > n // 8


def test_random_scene_range(qc):
    """4 M pairs shaped like plane tests: coordinate differences up to a few
    thousand, directions of every magnitude a camera or shadow ray has."""
    rng = np.random.default_rng(2)
    n = 1 << 22
    a = (rng.uniform(-3000, 3000, n) * 2.0 ** rng.integers(-30, 1, n)).astype(np.float32)
    d = (rng.normal(size=n) * 2.0 ** rng.integers(-8, 12, n)).astype(np.float32)
    fast = _assert_ieee(qc, a, d, "scene range")
    assert fast.mean() > 0.99


def test_significand_rows(qc):
    """Whole rows of the 2^23 x 2^23 significand table: every a in [1, 2)
    against divisor significands at both ends and spread over the range."""
    ex = (C.c_uint32 * 16)()
    rows = [(0, 1, 3), ((1 << 23) - 3, 1, 3), (12345, 349525, 24)]
    for first, step, count in rows:
        bad = qc.qc_sweep(first, step, count, ex, 8)
        assert bad == 0, f"rows {first}+{step}k: {bad} differ, e.g. a={ex[0]:#x} d={ex[1]:#x}"
