/*
 * pixel.hpp - per-pixel program helpers of octree_fsh.c:381-464 shared by the
 * render kernels: camera rotation, GLSL-style normalize/dot/max, RGBA8
 * packing, attribute fetch, wave reductions and the LDS stack layout.
 * Every expression keeps the operand order of the reference (no FMA).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qbt_internal.hpp"
#include "trace.hpp"

#pragma clang fp contract(off)

namespace qbt {

__device__ __forceinline__ void cross3(const float a[3], const float b[3], float r[3])
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

/* v + 2 * cross(q.xyz, cross(q.xyz, v) + q.w * v) (octree_fsh.c:392-395) */
__device__ __forceinline__ void quat_rotate(const float q[4], const float v[3], float r[3])
{
    float c1[3], s[3], c2[3];
    cross3(q, v, c1);
    s[0] = c1[0] + q[3] * v[0];
    s[1] = c1[1] + q[3] * v[1];
    s[2] = c1[2] + q[3] * v[2];
    cross3(q, s, c2);
    r[0] = v[0] + 2.0f * c2[0];
    r[1] = v[1] + 2.0f * c2[1];
    r[2] = v[2] + 2.0f * c2[2];
}

__device__ __forceinline__ float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ __forceinline__ void normalize3(const float v[3], float r[3])
{
    float s = 1.0f / sqrtf(dot3(v, v));
    r[0]    = v[0] * s;
    r[1]    = v[1] * s;
    r[2]    = v[2] * s;
}

__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; } /* GLSL max */

__device__ __forceinline__ uint32_t pack_rgba8(float c0, float c1, float c2, float c3)
{
    float    c[4] = {c0, c1, c2, c3};
    uint32_t r    = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float x = c[k];
        if (!(x > 0.0f)) x = 0.0f;
        if (x > 1.0f) x = 1.0f;
        r |= ((uint32_t) (x * 255.0f + 0.5f)) << (8 * k);
    }
    return r;
}

__device__ __forceinline__ void fetch_attr(const AttrDev& a, int m, float col[4], float nrm[4])
{
    if (m >= 0 && m < a.n) {
        float4 c = a.ca[2 * (size_t) m], v = a.ca[2 * (size_t) m + 1];
        col[0] = c.x, col[1] = c.y, col[2] = c.z, col[3] = 1.0f;
        nrm[0] = v.x, nrm[1] = v.y, nrm[2] = v.z, nrm[3] = 1.0f;
    } else {
        col[0] = col[1] = col[2] = col[3] = 0.0f;
        nrm[0] = nrm[1] = nrm[2] = nrm[3] = 0.0f;
    }
}

__device__ __forceinline__ void pixel_dir(const ViewDev& v, int px, int py, float csv[3])
{
    float c0[3] = {((float) px + 0.5f) - v.W / 2.0f, ((float) py + 0.5f) - v.H / 2.0f, 0.0f - (v.W / 2.0f) / 1.0f};
    float c1[3];
    quat_rotate(v.qz, c0, c1);
    quat_rotate(v.qx, c1, csv);
}

/* Work item -> pixel: item = 64 * tile + 8 * (py % 8) + px % 8 over 8x8
 * tiles. strip = 0: tiles in row-major order; strip = S > 0: the tile grid
 * cut into vertical strips S tiles wide (the last one narrower), strips left
 * to right, each in row-major order, so that the consecutive tiles that are
 * in flight together cover a compact screen region. False for the padding
 * items of edge tiles. */
__device__ __forceinline__ bool item_pixel(uint32_t item, int w, int h, int strip, int& px, int& py)
{
    const uint32_t tx = (uint32_t) ((w + 7) >> 3), ty = (uint32_t) ((h + 7) >> 3);
    const uint32_t tile = item >> 6, sub = item & 63u;
    uint32_t       cx, cy;
    if (strip <= 0) {
        cx = tile % tx;
        cy = tile / tx;
    } else {
        const uint32_t S = (uint32_t) strip, full = tx / S, inf = full * S * ty;
        if (tile < inf) {
            const uint32_t k = tile / (S * ty), u = tile - k * S * ty;
            cx               = k * S + u % S;
            cy               = u / S;
        } else {
            const uint32_t W = tx - full * S, u = tile - inf;
            cx               = full * S + u % W;
            cy               = u / W;
        }
    }
    px = (int) cx * 8 + (int) (sub & 7u);
    py = (int) cy * 8 + (int) (sub >> 3);
    return px < w && py < h;
}

__device__ __forceinline__ unsigned wave_sum(unsigned x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

/* Per-lane stacks. Level l of a thread is one slot of W words (W = 2: the
 * entry word and the node word; W = 3 with the dynamic tree's node word),
 * slots interleaved across threads: word j of level l of thread i is at
 * base[(l * nt + i) * W + j]. The two words a push writes and a pop reads sit
 * side by side (one 64-bit LDS access, one address), and consecutive lanes
 * touch consecutive slots. gstack: L * W words per thread of the grid. */
__device__ __forceinline__ Stack stack_at(uint32_t* base, uint32_t tid, uint32_t nt, int dual)
{
    Stack     st;
    const int W = dual ? 3 : 2;
    st.stride   = W * (int) nt;
    st.pw       = base + (size_t) W * tid;
    st.ns       = reinterpret_cast<int*>(st.pw) + 1;
    st.nd       = dual ? reinterpret_cast<int*>(st.pw) + 2 : st.ns;
    return st;
}

__device__ __forceinline__ Stack make_stack_global(uint32_t* gstack, int L, int dual)
{
    return stack_at(gstack, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, dual);
}

__device__ __forceinline__ Stack make_stack(uint32_t* lds, int L, int dual)
{
    return stack_at(lds, threadIdx.x, blockDim.x, dual);
}

} // namespace qbt
