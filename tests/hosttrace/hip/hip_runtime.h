/*
 * Host definitions of the few HIP vector types and intrinsics that
 * trace.hpp uses, so that the GPU traversal code can be compiled by g++ and
 * checked against the oracle on the CPU (tests/test_host_trace.py). Test
 * build only; the library is always built by hipcc for gfx950.
 */
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>

#define __device__
#define __host__
#define __forceinline__ inline
#define QBT_OPAQUE(x) ((void) 0)

struct float4 {
    float x, y, z, w;
};
struct int4 {
    int x, y, z, w;
};
struct uint2 {
    unsigned x, y;
};
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline int4   make_int4(int x, int y, int z, int w) { return int4{x, y, z, w}; }
inline uint2  make_uint2(unsigned x, unsigned y) { return uint2{x, y}; }
inline int    __popc(unsigned x) { return __builtin_popcount(x); }
inline int    __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned) x); }
inline unsigned __builtin_amdgcn_ubfe(unsigned x, unsigned off, unsigned w)
{
    return (x >> (off & 31u)) & ((w & 31u) ? (1u << (w & 31u)) - 1u : 0u);
}
