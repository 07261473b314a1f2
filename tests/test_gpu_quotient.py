"""Exhaustive device check of the plane-test quotient (trace.hpp plane_k)
against the device's IEEE division.

plane_k takes t = (c - p) / d from a per-ray reciprocal y = RN(1/d) as
q = RN(a y), e = RN(a - d q), t = RN(q + e y) where |d| is in
[2^-20, 2^100] and |q| is in [2^-80, 2^120], and divides elsewhere. Inside
that range |a| > 2^-101, y, q and t are normal numbers and e is exact (the
exact a - d q is a multiple of 2^(e_a - 47) >= 2^-148), so every
step commutes with scaling a and d by powers of two, and signs are symmetric:
t = RN(a / d) for all such operands exactly when it holds for every pair of
significands a, d in [1, 2). This test checks all 2^46 pairs on the device
build (built with the library's flags), in 32 chunks of 2^18 divisor rows.
The reference divides (octree_fsh.c:67, 80, 93); the host build of the same
code is checked in tests/test_quotient.py.
"""
import ctypes as C
import os
import time

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "hosttrace", "libquotient_gpu.so")


def test_all_significand_pairs():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built: run __graft_entry__.build()")
    from qubatron_amd import _ffi
    _ffi.dev()  # the HIP runtime the library links
    lib = C.CDLL(LIB)
    lib.qg_rows.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint32), C.c_int]
    rows, chunk = 1 << 23, 1 << 18
    ex = (C.c_uint32 * 16)()
    total = 0
    t0 = time.time()
    for mb0 in range(0, rows, chunk):
        bad = C.c_ulonglong(0)
        assert lib.qg_rows(mb0, chunk, C.byref(bad), ex, 8) == 0, "HIP error in the quotient check"
        total += bad.value
        print(f"[quotient] divisor rows {mb0}..{mb0 + chunk - 1}: {bad.value} mismatches "
              f"({time.time() - t0:.1f}s)", flush=True)
        assert bad.value == 0, f"a={ex[0]:#010x} d={ex[1]:#010x} and {bad.value - 1} more"
    assert total == 0
