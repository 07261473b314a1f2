/*
 * trace.hpp - per-ray sparse-octree traversal for gfx950.
 *
 * Semantics (template parameter SEM): cube_trace_line of
 * src/qubatron/shaders/octree_fsh.c:138-379 (static + dynamic tree in
 * lockstep), octree_trace_line of src/qubatron/octree.c:341-537, or the
 * particle shader's copy (src/qubatron/shaders/particle_vsh.c:121-343: one
 * tree, (0, 0, 0, 0) plane-test defaults, its slot address). Results (leaf node, entry point bits,
 * miss/discard) are identical to the reference algorithm; only the storage
 * differs:
 *
 *  - The reference keeps per level the cube, four pending entry points
 *    (4 x float4) and octets, the `ispsi` word and two node indices
 *    (~108 B/level, octree_fsh.c:50-58). Here a level holds one 32-bit
 *    pending word + the node index (8 B/level, 12 with the dynamic tree) in
 *    LDS: the cube is a function of the integer cell coordinates (ix, iy, iz)
 *    of the current node, and every pending point is re-derived from the
 *    plane that produced it (same inputs, same fp32 operations, so the same
 *    bits): a mid-plane hit of that level's node, or the node's entry point,
 *    which is the point that was popped one level up (followed upwards to
 *    the root entry).
 *  - The node array is split into a 32-byte child block per node (two int4,
 *    one aligned 32 B gather per expansion) and an int32 model index array.
 *
 * Pending word: bits 0-2 count n, bits 3-5 next index k, entry j at bit
 * 8+5j: octet (3 bits) | source << 3 (0 = node entry point, 1 = z, 2 = x,
 * 3 = y mid-plane).
 *
 * Compile with -ffp-contract=off; fp32 division and sqrt correctly rounded.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace qbt {

constexpr int kMaxLevel = 24;
constexpr float kFltMax = 3.402823466e+38f;

/* traversal semantics */
enum : int {
    kSemShader   = 0, /* octree_fsh.c */
    kSemCpu      = 1, /* octree.c octree_trace_line */
    kSemParticle = 2  /* particle_vsh.c */
};

/* An empty volatile asm on a value: the compiler must keep what depends on
 * it inside its branch (a rarely taken slow path would otherwise be
 * if-converted and computed on every lane). The host test build of this
 * header defines it as nothing. */
#ifndef QBT_OPAQUE
#define QBT_OPAQUE(x) asm volatile("" : "+v"(x))
#endif

enum TraceStatus : int { ST_MISS = 0, ST_HIT = 1, ST_DISCARD = -1, ST_ABORT = -2 };

struct TreeDev {
    const int4* __restrict__ child; /* direct layout: 2 per node */
    const int* __restrict__ model;  /* slot 8, 1 per node (of the layout traversed) */
    int n;                          /* node count of the layout traversed */
    const uint2* __restrict__ rec;  /* compact layout: {first child, live | alloc << 8} */
    const int* __restrict__ ref;    /* compact layout: reference node index */
};

struct AttrDev {
    const float4* __restrict__ ca; /* [2m] = colour (r,g,b,1), [2m+1] = normal (x,y,z,1) */
    int n;
};

struct SceneDev {
    TreeDev stat, dyn;
    AttrDev astat, adyn;
    float bx, by, bz, bw; /* basecube */
    int maxlevel;
    int iter_cap;
};

struct Ray {
    float px, py, pz, dx, dy, dz;
    float ix, iy, iz; /* RN(1/d) per axis (set by trace_begin), NaN where not usable */
};

/* RN(1 / d) when |d| is in [2^-20, 2^100] (the reciprocal is then a
 * normal number of magnitude at most 2^20), else NaN so that plane_k takes
 * the division */
__device__ __forceinline__ float recip_or_nan(float d)
{
    const float a = fabsf(d);
    return (a >= 0x1p-20f && a <= 0x1p+100f) ? 1.0f / d : __builtin_nanf("");
}

/* Plane test on axis k (0 x, 1 y, 2 z) at coordinate c: is_cube_xplane,
 * is_cube_yplane and is_cube_zplane (octree_fsh.c:62-99 / octree.c:302-339)
 * in one form. t = (c - p_k) / d_k; the other two coordinates are
 * p_j + d_j t, coordinate k is c; d_k == 0 gives the octree.c twin's
 * defaults (0, 0, 0, t = FLT_MAX), the particle shader's (0, 0, 0, 0)
 * (particle_vsh.c:64-65), and NaN in shader semantics (see below).
 *
 * The quotient is the correctly rounded one, as the reference's division,
 * from y = RN(1 / d_k): q = RN(a y), e = RN(a - d_k q) (one FMA),
 * t = RN(q + e y) (one FMA), a Markstein-style correction. It is taken
 * where |d_k| is in [2^-20, 2^100] (y not NaN) and |q| is in [2^-80, 2^120].
 * There |y| <= 2^20, so |a| >= |q| / (|y| (1 + 2^-24)) > 2^-101: y, q and t
 * are normal numbers and e is exact (a - d_k q is a multiple of
 * 2^(e_a - 47) >= 2^-148), so each step scales with the exponents of a and
 * d_k, and t is correct for all such operands exactly when it is for every
 * pair of significands in [1, 2). All 2^46
 * pairs are checked on the device (tests/test_gpu_quotient.py) and a sample
 * on the host (tests/test_quotient.py). Elsewhere, and where y is NaN, the
 * quotient is the division itself, on a branch that rays rarely take.
 * Everything else is selects, so that lanes of a wave do not split on the
 * axis or on d_k == 0. */
/* component k by value (a conditional over the Ray's members would select an
 * address and keep the Ray in scratch memory) */
__device__ __forceinline__ float axis3(int k, float x, float y, float z)
{
    return k == 0 ? x : (k == 1 ? y : z);
}

template <int SEM>
__device__ __forceinline__ float4 plane_k(int k, float c, const Ray& r)
{
    const float p  = axis3(k, r.px, r.py, r.pz);
    const float d  = axis3(k, r.dx, r.dy, r.dz);
    const float y  = axis3(k, r.ix, r.iy, r.iz);
    const float a  = c - p;
    const float q  = a * y;
    const float e  = __builtin_fmaf(-q, d, a);
    float       t  = __builtin_fmaf(e, y, q);
    const float aq = fabsf(q);
    const bool  ok = d != 0.0f;
    if (!(aq >= 0x1p-80f && aq <= 0x1p+120f) && ok) {
        float aa = a, dd = d;
        QBT_OPAQUE(aa);
        QBT_OPAQUE(dd);
        t = aa / dd;
    }
    const float X = r.px + r.dx * t, Y = r.py + r.dy * t, Z = r.pz + r.dz * t;
    float4      o;
    o.x = k == 0 ? c : X;
    o.y = k == 1 ? c : Y;
    o.z = k == 2 ? c : Z;
    o.w = t;
    if (SEM != kSemShader) { /* these defaults can pass the containment tests */
        o.x = ok ? o.x : 0.0f;
        o.y = ok ? o.y : 0.0f;
        o.z = ok ? o.z : 0.0f;
        o.w = ok ? o.w : (SEM == kSemCpu ? kFltMax : 0.0f);
    }
    /* Shader semantics: where d_k == 0, y is NaN and so are t and the other
     * two coordinates; every caller tests t > 0 and the containment of the
     * point before using it, and those comparisons are false for NaN just
     * as they are for the shader's FLT_MAX defaults (octree_fsh.c:64).
     * No select is spent on the defaults. */
    return o;
}

template <int SEM>
__device__ __forceinline__ float4 plane_x(float c, const Ray& r) { return plane_k<SEM>(0, c, r); }
template <int SEM>
__device__ __forceinline__ float4 plane_y(float c, const Ray& r) { return plane_k<SEM>(1, c, r); }
template <int SEM>
__device__ __forceinline__ float4 plane_z(float c, const Ray& r) { return plane_k<SEM>(2, c, r); }

struct Hit {
    float4 isp;
    int ns, nd;     /* leaf node indices (static, dynamic) as the shader holds them */
    int ix, iy, iz; /* leaf cell coordinates */
    unsigned exps;  /* node expansions */
};

/* Node cube (tlf x, y, z) at `level` with cell coordinates (ix, iy, iz).
 * GRID: every coordinate of the subdivision is an exact fp32 multiple of the
 * leaf edge (checked on the host), so x = bx + ix*w is the value the
 * reference's running sum reaches. Otherwise the running sum of the
 * reference (tlf.x += xsft*halfs, octree_fsh.c:342-347) is replayed from the
 * root along the path bits. */
template <bool GRID>
__device__ __forceinline__ void cube_at(const SceneDev& s, int level, int ix, int iy, int iz, float& x, float& y,
                                        float& z, float& w)
{
    w = ldexpf(s.bw, -level);
    if (GRID) { /* i * w is exact, so the FMA rounds as the separate add */
        x = __builtin_fmaf((float) ix, w, s.bx);
        y = __builtin_fmaf(-(float) iy, w, s.by);
        z = __builtin_fmaf(-(float) iz, w, s.bz);
    } else {
        x = s.bx, y = s.by, z = s.bz;
        for (int l = 0; l < level; l++) {
            float h  = ldexpf(s.bw, -(l + 1));
            int   sh = level - 1 - l;
            x        = x + (float) ((ix >> sh) & 1) * h;
            y        = y - (float) ((iy >> sh) & 1) * h;
            z        = z - (float) ((iz >> sh) & 1) * h;
        }
    }
}

/* Far corner (bxx, byy, bzz) and mid-planes (hx, hy, hz) of the node cube
 * (x, y, z, w), as the shader computes them (octree_fsh.c:255-256:
 * h = b + (t - b) * 0.5). On an exact grid every one of these values is a
 * multiple of the leaf edge that fp32 holds exactly, so h = t +- w / 2 is
 * the same number in one FMA. */
template <bool GRID>
__device__ __forceinline__ void cube_mid(float x, float y, float z, float w, float& bxx, float& byy, float& bzz,
                                         float& hx, float& hy, float& hz)
{
    bxx = x + w, byy = y - w, bzz = z - w;
    if (GRID) {
        hx = __builtin_fmaf(0.5f, w, x);
        hy = __builtin_fmaf(-0.5f, w, y);
        hz = __builtin_fmaf(-0.5f, w, z);
    } else {
        hx = bxx + (x - bxx) * 0.5f;
        hy = byy + (y - byy) * 0.5f;
        hz = bzz + (z - bzz) * 0.5f;
    }
}

/* Mid-plane hit `src` (1 z, 2 x, 3 y) of the node cube (x, y, z, w). */
template <int SEM, bool GRID>
__device__ __forceinline__ float4 midplane(int src, const Ray& r, float x, float y, float z, float w)
{
    float bxx, byy, bzz, hx, hy, hz;
    cube_mid<GRID>(x, y, z, w, bxx, byy, bzz, hx, hy, hz);
    const int   k = src == 1 ? 2 : src - 2; /* src 1 z, 2 x, 3 y */
    const float c = src == 1 ? hz : (src == 2 ? hx : hy);
    return plane_k<SEM>(k, c, r);
}

/* midplane() of the cube at `level` whose cell coordinates are (ix, iy,
 * iz) >> sh. On the closed-form grid only the plane's own coordinate of the
 * cube is formed, by the two FMAs that cube_at and cube_mid apply to it (the
 * sign folded into the exact product), so the bits are the same. */
template <int SEM, bool GRID>
__device__ __forceinline__ float4 midplane_cell(const SceneDev& s, const Ray& r, int src, int level, int ix, int iy,
                                                int iz, int sh)
{
    if constexpr (GRID) {
        const int   k  = src == 1 ? 2 : src - 2; /* src 1 z, 2 x, 3 y */
        const int   ci = (k == 0 ? ix : (k == 1 ? iy : iz)) >> sh;
        const float w  = ldexpf(s.bw, -level);
        const float sg = k == 0 ? 1.0f : -1.0f; /* x grows along +x, y and z along -y, -z */
        const float c0 = __builtin_fmaf(sg * (float) ci, w, axis3(k, s.bx, s.by, s.bz));
        return plane_k<SEM>(k, __builtin_fmaf(sg * 0.5f, w, c0), r);
    } else {
        float x, y, z, ww;
        cube_at<GRID>(s, level, ix >> sh, iy >> sh, iz >> sh, x, y, z, ww);
        return midplane<SEM, GRID>(src, r, x, y, z, ww);
    }
}

/* component selects by value (a conditional over lvalues selects an address
 * and forces the operands into scratch memory) */
__device__ __forceinline__ int sel(bool c, int a, int b) { return c ? a : b; }
__device__ __forceinline__ int pick4(int x, int y, int z, int w, int k)
{
    return sel(k == 0, x, sel(k == 1, y, sel(k == 2, z, w)));
}
__device__ __forceinline__ int pick8(int4 lo, int4 hi, int o)
{
    const bool h = (o & 4) != 0;
    return pick4(sel(h, hi.x, lo.x), sel(h, hi.y, lo.y), sel(h, hi.z, lo.z), sel(h, hi.w, lo.w), o & 3);
}
__device__ __forceinline__ unsigned exist_mask(int4 lo, int4 hi)
{
    return (lo.x > 0 ? 1u : 0u) | (lo.y > 0 ? 2u : 0u) | (lo.z > 0 ? 4u : 0u) | (lo.w > 0 ? 8u : 0u) |
           (hi.x > 0 ? 16u : 0u) | (hi.y > 0 ? 32u : 0u) | (hi.z > 0 ? 64u : 0u) | (hi.w > 0 ? 128u : 0u);
}

/* The child block of one node in either layout: the 8 child indices of the
 * direct layout (two int4), or the compact {first child, mask} record. */
template <bool CMP>
struct Block;

template <>
struct Block<false> {
    int4 lo, hi;
    __device__ __forceinline__ void clear() { lo = hi = make_int4(0, 0, 0, 0); }
    __device__ __forceinline__ void load(const TreeDev& t, int i)
    {
        const int4* p = t.child + 2 * (size_t) i;
        lo            = p[0];
        hi            = p[1];
    }
    __device__ __forceinline__ unsigned mask() const { return exist_mask(lo, hi); }
    __device__ __forceinline__ int      child(int o) const { return pick8(lo, hi, o); }
};

/* Compact record: x = index of the first child; y bits 0-7 = the children
 * that exist (live), bits 8-15 = the slots of the child block (alloc): child
 * o is record x + (the number of alloc bits below o). Live is a subset of
 * alloc; they differ only where an in-place range update zeroed a slot and
 * left its record in the block. */
template <>
struct Block<true> {
    uint2 r;
    __device__ __forceinline__ void     clear() { r = make_uint2(0u, 0u); }
    __device__ __forceinline__ void     load(const TreeDev& t, int i) { r = t.rec[i]; }
    __device__ __forceinline__ unsigned mask() const { return r.y & 0xFFu; }
    __device__ __forceinline__ int      child(int o) const
    {
        return ((r.y >> o) & 1u) ? (int) (r.x + (uint32_t) __popc((r.y >> 8) & ((1u << o) - 1u))) : 0;
    }
};

/* The particle shader addresses slot k of node i as texel
 * (3 i mod 8192 + k / 4, 3 i / 8192) (particle_vsh.c:110-119): the texel
 * column does not carry into the next row, so for a node whose three texels
 * straddle a row end the column passes 8191. Such a fetch is outside the
 * texture, for which GLSL ES leaves the result undefined; here it reads 0,
 * the value of a robust-access fetch. Slots 0-7 are affected only where
 * 3 i mod 8192 == 8191 (slots 4-7), the model index (slot 8) also where
 * 3 i mod 8192 == 8190. */
__device__ __forceinline__ bool particle_hi_lost(int i) { return ((uint32_t) i * 3u) % 8192u == 8191u; }
__device__ __forceinline__ bool particle_model_lost(int i) { return ((uint32_t) i * 3u) % 8192u >= 8190u; }

/* the child block of node i as the semantics SEM reads it (direct layout) */
template <int SEM>
__device__ __forceinline__ void load_block(Block<false>& b, const TreeDev& t, int i)
{
    b.load(t, i);
    if (SEM == kSemParticle && particle_hi_lost(i)) b.hi = make_int4(0, 0, 0, 0);
}
template <int SEM>
__device__ __forceinline__ void load_block(Block<true>& b, const TreeDev& t, int i)
{
    static_assert(SEM != kSemParticle, "the particle semantics addresses reference node indices: direct layout");
    b.load(t, i);
}

/* slot k (< 8) of node i at `level` (oct_from_octets_for_index,
 * octree_fsh.c:127-136): 0 for index 0 below the root. */
template <bool CMP, int SEM = kSemShader>
__device__ __forceinline__ int child_slot(const TreeDev& t, int i, int k, int level)
{
    if (i == 0 && level > 0) return 0;
    Block<CMP> b;
    load_block<SEM>(b, t, i);
    return b.child(k);
}

/* The calling lane's stack: the entry word, the node word and the dynamic
 * node word of level l are pw[l * stride], ns[l * stride], nd[l * stride]. */
struct Stack {
    uint32_t* pw;
    int* ns;
    int* nd;
    int stride;
};

/* Entry point of the node at `level` on the current path (level >= 1 uses
 * the entry popped one level up). */
template <int SEM, bool GRID>
__device__ __forceinline__ float4 entry_of(const SceneDev& s, const Ray& r, const Stack& st, int level, int ix, int iy,
                                        int iz, float4 start)
{
    for (int m = level; m > 0; m--) {
        uint32_t w   = st.pw[(m - 1) * st.stride];
        int      k   = (int) ((w >> 3) & 7u) - 1;
        int      src = (int) ((w >> (8 + 5 * k + 3)) & 3u);
        if (src != 0) return midplane_cell<SEM, GRID>(s, r, src, m - 1, ix, iy, iz, level - (m - 1));
    }
    return start;
}

/* Resumable traversal state of one ray (registers; the per-level pending
 * words and node indices live in the LDS Stack). */
struct TraceState {
    float4   entry; /* entry point of the current node */
    float4   start; /* root entry point */
    int      level, ix, iy, iz;
    int      ns, nd; /* current static / dynamic node */
    int      it;
    unsigned exps;
    uint32_t pend; /* single-tree compact steps: bit l set while level l has pending entries */
};

enum : int { ST_CONT = 2 };

/* Root entry (octree_fsh.c:157-211). Returns ST_CONT, ST_HIT (maxlevel 0)
 * or ST_DISCARD. */
template <int SEM>
__device__ __forceinline__ int trace_begin(const SceneDev& s, Ray& r, TraceState& t)
{
    r.ix = recip_or_nan(r.dx);
    r.iy = recip_or_nan(r.dy);
    r.iz = recip_or_nan(r.dz);
    float4 h0 = make_float4(0.f, 0.f, 0.f, 0.f), h1 = h0;
    int    hitc = 0;
    {
        const float tx = s.bx, ty = s.by, tz = s.bz, tw = s.bw;
        const float rx = tx + tw, ry = ty - tw, rz = tz - tw;
        float4      a;
#define QBT_CONSIDER(cond)                       \
    if ((SEM != kSemCpu || a.w < kFltMax) && (cond)) { \
        if (hitc == 0) h0 = a;                   \
        else if (hitc == 1) h1 = a;              \
        hitc++;                                  \
    }
        a = plane_z<SEM>(tz, r);
        QBT_CONSIDER(tx < a.x && a.x <= rx && ty > a.y && a.y >= ry)
        a = plane_z<SEM>(rz, r);
        QBT_CONSIDER(tx < a.x && a.x <= rx && ty > a.y && a.y >= ry)
        a = plane_x<SEM>(tx, r);
        QBT_CONSIDER(ty > a.y && a.y >= ry && tz > a.z && a.z >= rz)
        a = plane_x<SEM>(rx, r);
        QBT_CONSIDER(ty > a.y && a.y >= ry && tz > a.z && a.z >= rz)
        a = plane_y<SEM>(ty, r);
        QBT_CONSIDER(tx < a.x && a.x <= rx && tz > a.z && a.z >= rz)
        a = plane_y<SEM>(ry, r);
        QBT_CONSIDER(tx < a.x && a.x <= rx && tz > a.z && a.z >= rz)
#undef QBT_CONSIDER
    }
    t.exps = 0;
    if (hitc < 2) return ST_DISCARD;
    if (h0.w < 0.0f && h1.w < 0.0f) return ST_DISCARD;
    if (h1.w < h0.w) h0 = h1;
    if (h0.w < 0.0f) h0 = make_float4(r.px, r.py, r.pz, 0.0f);
    t.start = h0;
    t.entry = h0;
    t.level = t.ix = t.iy = t.iz = 0;
    t.ns = t.nd = 0;
    t.it   = 0;
    t.pend = 0;
    return s.maxlevel == 0 ? ST_HIT : ST_CONT;
}

/*
 * Single-tree steps on the compact layout (the hot configuration) keep a
 * different per-level encoding, so that backtracking never re-reads a node:
 *   ns word  first child index of the level's node (bits 0-29) | entries
 *            still pending (bits 30-31, 0..3)
 *   pw word  the entry being visited (bits 0-4), then up to three pending
 *            entries (bits 5-9, 10-14, 15-19); an entry is octet (3 bits) |
 *            source << 3 (2 bits); bits 24-31 hold the slots of the node's
 *            child block (alloc), so the child of octet o is first + (the
 *            number of those bits below o), counted when the entry is taken.
 * A per-lane bit mask marks the levels that still hold pending entries, so a
 * backtrack finds its level with one count-leading-zeros instead of a loop
 * over the stack.
 * Candidates are sorted as (t, key) pairs; each key carries the candidate's
 * octet and its tie-flipped octet, computed before the sort from the same
 * comparisons (octree_fsh.c:292-311), and its source; the entry point of the
 * first child is selected by source afterwards. Same results, fewer moves.
 */
/* candidate src (0 entry, 1 z, 2 x, 3 y) by value: a conditional over the
 * lvalues would select an address and keep the candidates in scratch memory */
__device__ __forceinline__ float4 pick_entry(int src, float4 e, float4 z, float4 x, float4 y)
{
    float4 r = e;
    if (src == 1) r = z;
    if (src == 2) r = x;
    if (src == 3) r = y;
    return r;
}

/* bit o of m (one bit-field extract) */
__device__ __forceinline__ bool bit_of(uint32_t m, int o) { return __builtin_amdgcn_ubfe(m, (uint32_t) o, 1u) != 0u; }

__device__ __forceinline__ void kswap(float& wa, int& ka, float& wb, int& kb)
{
    const bool sw = wb < wa; /* strict: ties keep order (octree_fsh.c:283); NaN never moves */
    const float tw = sw ? wb : wa;
    const int   tk = sw ? kb : ka;
    wb             = sw ? wa : wb;
    kb             = sw ? ka : kb;
    wa             = tw;
    ka             = tk;
}

/* The octet of a candidate point and its tie-flipped octet
 * (octree_fsh.c:292-311): the flip toggles the bit of the first axis, in
 * x, y, z order, on which the point lies on the mid-plane (the lowest set
 * bit of the on-plane mask). Key = octet | src << 3 | flipped << 5. */
__device__ __forceinline__ int octet_key(float x, float y, float z, float hx, float hy, float hz, int src)
{
    const int o  = (x > hx ? 1 : 0) | (y < hy ? 2 : 0) | (z < hz ? 4 : 0);
    const int on = (x == hx ? 1 : 0) | (y == hy ? 2 : 0) | (z == hz ? 4 : 0);
    return o | (src << 3) | ((o ^ (on & -on)) << 5);
}

/* The same keys for the three mid-plane candidates, whose coordinate on
 * their own plane is the mid-plane value itself (plane_k sets it) wherever
 * the candidate is valid, so that one comparison of each is known: z == hz
 * for the z-plane hit, x == hx for the x-plane hit, y == hy for the y-plane
 * hit. (Keys of invalid candidates are never read.) */
__device__ __forceinline__ int zplane_key(float x, float y, float hx, float hy)
{
    const int o  = (x > hx ? 1 : 0) | (y < hy ? 2 : 0);
    const int fl = x == hx ? 1 : (y == hy ? 2 : 4);
    return o | (1 << 3) | ((o ^ fl) << 5);
}
__device__ __forceinline__ int xplane_key(float y, float z, float hy, float hz)
{
    const int o = (y < hy ? 2 : 0) | (z < hz ? 4 : 0);
    return o | (2 << 3) | ((o ^ 1) << 5);
}
__device__ __forceinline__ int yplane_key(float x, float z, float hx, float hz)
{
    const int o = (x > hx ? 1 : 0) | (z < hz ? 4 : 0);
    return o | (3 << 3) | ((o ^ (x == hx ? 1 : 2)) << 5);
}

template <int SEM, bool GRID>
__device__ __forceinline__ int trace_step_c1(const SceneDev& s, const Ray& r, const Stack& st, TraceState& t)
{
    /* the iteration cap is checked at the end: a lane over it still takes
     * this step (in bounds, like any other), then reports the abort */
    const bool  over  = ++t.it > s.iter_cap;
    const int   level = t.level;
    const uint2 rec   = s.stat.rec[(uint32_t) t.ns]; /* node indices are non-negative: no sign extension */
    t.exps++;
    const uint32_t mask = rec.y & 0xFFu; /* live children; 0 for the root of an empty tree: no candidate taken */

    /* ---- expansion (octree_fsh.c:251-330), without branches ---- */
    float tx, ty, tz, tw, bxx, byy, bzz, hx, hy, hz;
    cube_at<GRID>(s, level, t.ix, t.iy, t.iz, tx, ty, tz, tw);
    cube_mid<GRID>(tx, ty, tz, tw, bxx, byy, bzz, hx, hy, hz);
    const float  qnan = __builtin_nanf("");
    const float4 az = plane_z<SEM>(hz, r);
    const float4 ax = plane_x<SEM>(hx, r);
    const float4 ay = plane_y<SEM>(hy, r);
    float        w0 = t.entry.w;
    float        w1 = (az.w > 0.0f && tx < az.x && az.x <= bxx && ty > az.y && az.y >= byy) ? az.w : qnan;
    float        w2 = (ax.w > 0.0f && ty > ax.y && ax.y >= byy && tz > ax.z && ax.z >= bzz) ? ax.w : qnan;
    float        w3 = (ay.w > 0.0f && tx < ay.x && ay.x <= bxx && tz > ay.z && ay.z >= bzz) ? ay.w : qnan;
    /* the entry's key without its tie-flipped octet: a key in slot 0 after
     * the sort never flips (there is no previous octet), and the entry only
     * leaves slot 0 in the branch below, which completes its key first */
    const int    oe = (t.entry.x > hx ? 1 : 0) | (t.entry.y < hy ? 2 : 0) | (t.entry.z < hz ? 4 : 0);
    int          k0 = oe | (oe << 5);
    /* the octree.c twin's defaults for d == 0 (0, 0, 0) can pass the
     * containment tests, so it keeps the general form */
    int k1 = SEM != kSemShader ? octet_key(az.x, az.y, az.z, hx, hy, hz, 1) : zplane_key(az.x, az.y, hx, hy);
    int k2 = SEM != kSemShader ? octet_key(ax.x, ax.y, ax.z, hx, hy, hz, 2) : xplane_key(ax.y, ax.z, hy, hz);
    int k3 = SEM != kSemShader ? octet_key(ay.x, ay.y, ay.z, hx, hy, hz, 3) : yplane_key(ay.x, ay.z, hx, hz);
    /* Sort network (octree_fsh.c:283, strict: ties keep their order). Slot
     * 0 holds the node's own entry point, which lies before every point of
     * the ray inside the node but for rounding near an edge, so the three
     * exchanges with slot 0 are no-ops for almost every lane; the branch
     * skips them where no lane of the wave needs them. NaN never moves, so
     * slot 0 stays valid. */
    if (w1 < w0 || w2 < w0 || w3 < w0) {
        k0 = octet_key(t.entry.x, t.entry.y, t.entry.z, hx, hy, hz, 0);
        kswap(w0, k0, w1, k1);
        kswap(w0, k0, w2, k2);
        kswap(w0, k0, w3, k3);
    }
    kswap(w1, k1, w2, k2);
    kswap(w1, k1, w3, k3);
    kswap(w2, k2, w3, k3);

    /* walk the sorted candidates: an octet equal to the previous valid one
     * is replaced by its flipped octet; those whose child exists are kept,
     * in order, as 5-bit entries (the word is assembled from the last one, so
     * that the first kept one ends in the low bits), the child mask on top */
    const bool v1 = w1 == w1, v2 = w2 == w2, v3 = w3 == w3;
    const int  o0 = k0 & 7;
    const int  o1 = (k1 & 7) == o0 ? ((k1 >> 5) & 7) : (k1 & 7);
    const int  p1 = v1 ? o1 : o0;
    const int  o2 = (k2 & 7) == p1 ? ((k2 >> 5) & 7) : (k2 & 7);
    const int  p2 = v2 ? o2 : p1;
    const int  o3 = (k3 & 7) == p2 ? ((k3 >> 5) & 7) : (k3 & 7);
    auto entry = [](int o, int k) -> uint32_t { return (uint32_t) o | ((uint32_t) k & 0x18u); };
    const bool t0 = bit_of(mask, o0);
    const bool t1 = v1 && bit_of(mask, o1);
    const bool t2 = v2 && bit_of(mask, o2);
    const bool t3 = v3 && bit_of(mask, o3);
    uint32_t   word = t3 ? entry(o3, k3) : 0u;
    word            = t2 ? (word << 5) | entry(o2, k2) : word;
    word            = t1 ? (word << 5) | entry(o1, k1) : word;
    word            = t0 ? (word << 5) | entry(o0, k0) : word;
    word |= (rec.y >> 8) << 24; /* the block's slots, for the rank of an entry's child */
    const int n     = (t0 ? 1 : 0) + (t1 ? 1 : 0) + (t2 ? 1 : 0) + (t3 ? 1 : 0);

    /* The entry to descend into: the first kept candidate, or, backtracking
     * (octree_fsh.c:368-375), the next pending entry of the nearest ancestor
     * that has one. Descending and backtracking lanes share one path (most
     * wave-steps hold both kinds): the level they work on is lv, whose two
     * stack words every lane reads and rewrites, and the new words, the
     * pending mask and the child block are selects. */
    const bool     desc  = n > 0;
    const uint32_t below = t.pend & ((1u << level) - 1u);
    if (!desc && below == 0u) return over ? ST_ABORT : ST_MISS;
    /* the deepest level with a pending entry, from the per-lane mask (levels
     * between it and this one have none: no loop over them) */
    const int      lv  = desc ? level : 31 - __clz((int) below);
    const uint32_t an  = (uint32_t) st.ns[lv * st.stride];
    const uint32_t pwv = st.pw[lv * st.stride];
    /* backtracking: the next entry of level lv becomes the current one */
    const uint32_t w   = desc ? word : ((pwv & 0xFFFFFu) >> 5) | (pwv & 0xFF000000u);
    st.pw[lv * st.stride] = w;
    st.ns[lv * st.stride] = (int) (desc ? rec.x | (uint32_t) (n - 1) << 30 : an - (1u << 30));
    /* descending: bit `level` set when more than one entry was kept;
     * backtracking: bit lv cleared when this was the level's last pending
     * entry (count field 1: bit 31 of the node word clear) */
    t.pend = desc ? t.pend | (n > 1 ? 1u : 0u) << level : t.pend ^ ((1u << lv) & ~(uint32_t) ((int) an >> 31));
    const uint32_t a = desc ? rec.x : an & 0x3FFFFFFFu;
    /* Its entry point. Source 0 (the node's own entry point): a descending
     * lane keeps its entry; a backtracking one takes the mid-plane hit of the
     * nearest node on the path entered through a mid-plane, else the root
     * entry. Sources 1-3: that mid-plane of the level-j node, recomputed from
     * the plane (same operations as the candidate, so the same bits). */
    int j = lv, src = (int) ((w >> 3) & 3u);
    if (!desc && src == 0) {
        while (src == 0 && j > 0) {
            j--;
            src = (int) ((st.pw[j * st.stride] >> 3) & 3u);
        }
    }
    if (!desc || src != 0)
        t.entry = src != 0 ? midplane_cell<SEM, GRID>(s, r, src, j, t.ix, t.iy, t.iz, level - j) : t.start;
    const int sh = level - lv, o = (int) (w & 7u);
    t.ns    = (int) (a + (uint32_t) __popc(__builtin_amdgcn_ubfe(w, 24u, (uint32_t) o))); /* first + rank */
    t.ix    = ((t.ix >> sh) << 1) | (o & 1);
    t.iy    = ((t.iy >> sh) << 1) | ((o >> 1) & 1);
    t.iz    = ((t.iz >> sh) << 1) | (o >> 2);
    t.level = lv + 1;
    return over ? ST_ABORT : (t.level == s.maxlevel ? ST_HIT : ST_CONT);
}

/*
 * One step: expand the node just entered and descend into its nearest
 * pending child, or backtrack and descend into the next pending child of an
 * ancestor. Returns ST_CONT, ST_HIT (t.level == maxlevel; the leaf is
 * t.ns / t.nd with entry point t.entry), ST_MISS or ST_ABORT.
 */
template <bool CMP, bool DUAL, int SEM, bool GRID>
__device__ __forceinline__ int trace_step(const SceneDev& s, const Ray& r, const Stack& st, TraceState& t)
{
    if constexpr (CMP && !DUAL) return trace_step_c1<SEM, GRID>(s, r, st, t);
    if (++t.it > s.iter_cap) return ST_ABORT;
    const int level = t.level;

    /* ---- expansion of the node just entered (octree_fsh.c:251-330) ---- */
    Block<CMP> sb, db;
    sb.clear();
    db.clear();
    if (t.ns != 0 || level == 0) {
        load_block<SEM>(sb, s.stat, t.ns);
        t.exps++;
    }
    if (DUAL && (t.nd != 0 || level == 0)) {
        db.load(s.dyn, t.nd);
        t.exps++;
    }
    unsigned mask = sb.mask();
    if (DUAL) mask |= db.mask();

    if (mask) {
        float tx, ty, tz, tw, bxx, byy, bzz, hx, hy, hz;
        cube_at<GRID>(s, level, t.ix, t.iy, t.iz, tx, ty, tz, tw);
        cube_mid<GRID>(tx, ty, tz, tw, bxx, byy, bzz, hx, hy, hz);

        /* candidates as (t, key) pairs, as in trace_step_c1 */
        const float  qnan = __builtin_nanf("");
        const float4 az   = plane_z<SEM>(hz, r);
        const float4 ax   = plane_x<SEM>(hx, r);
        const float4 ay   = plane_y<SEM>(hy, r);
        float        w0   = t.entry.w;
        float        w1 = (az.w > 0.0f && tx < az.x && az.x <= bxx && ty > az.y && az.y >= byy) ? az.w : qnan;
        float        w2 = (ax.w > 0.0f && ty > ax.y && ax.y >= byy && tz > ax.z && ax.z >= bzz) ? ax.w : qnan;
        float        w3 = (ay.w > 0.0f && tx < ay.x && ay.x <= bxx && tz > ay.z && ay.z >= bzz) ? ay.w : qnan;
        const int    oe = (t.entry.x > hx ? 1 : 0) | (t.entry.y < hy ? 2 : 0) | (t.entry.z < hz ? 4 : 0);
        int          k0 = oe | (oe << 5);
        int k1 = SEM != kSemShader ? octet_key(az.x, az.y, az.z, hx, hy, hz, 1) : zplane_key(az.x, az.y, hx, hy);
        int k2 = SEM != kSemShader ? octet_key(ax.x, ax.y, ax.z, hx, hy, hz, 2) : xplane_key(ax.y, ax.z, hy, hz);
        int k3 = SEM != kSemShader ? octet_key(ay.x, ay.y, ay.z, hx, hy, hz, 3) : yplane_key(ay.x, ay.z, hx, hz);
        if (w1 < w0 || w2 < w0 || w3 < w0) { /* selection-swap sort (octree_fsh.c:283); NaN never moves */
            k0 = octet_key(t.entry.x, t.entry.y, t.entry.z, hx, hy, hz, 0);
            kswap(w0, k0, w1, k1);
            kswap(w0, k0, w2, k2);
            kswap(w0, k0, w3, k3);
        }
        kswap(w1, k1, w2, k2);
        kswap(w1, k1, w3, k3);
        kswap(w2, k2, w3, k3);

        /* flip an octet equal to the previous valid one; keep those whose
         * child exists in either tree, in order, as 5-bit entries */
        const bool v1 = w1 == w1, v2 = w2 == w2, v3 = w3 == w3;
        const int  o0 = k0 & 7;
        const int  o1 = (k1 & 7) == o0 ? ((k1 >> 5) & 7) : (k1 & 7);
        const int  p1 = v1 ? o1 : o0;
        const int  o2 = (k2 & 7) == p1 ? ((k2 >> 5) & 7) : (k2 & 7);
        const int  p2 = v2 ? o2 : p1;
        const int  o3 = (k3 & 7) == p2 ? ((k3 >> 5) & 7) : (k3 & 7);
        const bool t0 = bit_of(mask, o0);
        const bool t1 = v1 && bit_of(mask, o1);
        const bool t2 = v2 && bit_of(mask, o2);
        const bool t3 = v3 && bit_of(mask, o3);
        auto ent = [](int o, int k) -> uint32_t { return (uint32_t) o | ((uint32_t) k & 0x18u); };
        uint32_t entries = t3 ? ent(o3, k3) : 0u;
        entries          = t2 ? (entries << 5) | ent(o2, k2) : entries;
        entries          = t1 ? (entries << 5) | ent(o1, k1) : entries;
        entries          = t0 ? (entries << 5) | ent(o0, k0) : entries;
        const int n      = (t0 ? 1 : 0) + (t1 ? 1 : 0) + (t2 ? 1 : 0) + (t3 ? 1 : 0);

        if (n > 0) {
            /* pop the first pending entry now (octree_fsh.c:336-367) */
            const int foct = (int) (entries & 7u);
            st.pw[level * st.stride] = (entries << 8) | (uint32_t) n | (1u << 3);
            st.ns[level * st.stride] = t.ns;
            if (DUAL) st.nd[level * st.stride] = t.nd;
            t.pend |= (n > 1 ? 1u : 0u) << level;
            t.entry = pick_entry((int) ((entries >> 3) & 3u), t.entry, az, ax, ay);
            t.ns    = sb.child(foct);
            t.nd    = DUAL ? db.child(foct) : 0;
            t.ix    = (t.ix << 1) | (foct & 1);
            t.iy    = (t.iy << 1) | ((foct >> 1) & 1);
            t.iz    = (t.iz << 1) | (foct >> 2);
            t.level = level + 1;
            return t.level == s.maxlevel ? ST_HIT : ST_CONT;
        }
    }

    /* ---- backtrack (octree_fsh.c:368-375) and pop the next pending ---- */
    /* the deepest level with a pending entry, from the per-lane mask (bit l
     * set while level l has one; the levels between have none) */
    const uint32_t below = t.pend & ((1u << level) - 1u);
    if (below == 0u) return ST_MISS;
    const int lv = 31 - __clz((int) below), sh = level - lv;
    const int ix = t.ix >> sh, iy = t.iy >> sh, iz = t.iz >> sh;
    {
        uint32_t w = st.pw[lv * st.stride];
        int      k = (int) ((w >> 3) & 7u), cnt = (int) (w & 7u);
        {
            int e   = (int) ((w >> (8 + 5 * k)) & 31u);
            int o   = e & 7, src = e >> 3;
            st.pw[lv * st.stride] = (w & ~(7u << 3)) | ((uint32_t) (k + 1) << 3);
            t.pend &= k + 1 == cnt ? ~(1u << lv) : ~0u;
            float4 isp;
            if (src == 0) {
                isp = entry_of<SEM, GRID>(s, r, st, lv, ix, iy, iz, t.start);
            } else {
                isp = midplane_cell<SEM, GRID>(s, r, src, lv, ix, iy, iz, 0);
            }
            t.ns = child_slot<CMP, SEM>(s.stat, st.ns[lv * st.stride], o, lv);
            t.nd = DUAL ? child_slot<CMP>(s.dyn, st.nd[lv * st.stride], o, lv) : 0;
            t.ix    = (ix << 1) | (o & 1);
            t.iy    = (iy << 1) | ((o >> 1) & 1);
            t.iz    = (iz << 1) | (o >> 2);
            t.entry = isp;
            t.level = lv + 1;
            return t.level == s.maxlevel ? ST_HIT : ST_CONT;
        }
    }
}

/*
 * Trace one ray to the end. Returns ST_HIT / ST_MISS / ST_DISCARD / ST_ABORT.
 * DUAL: the dynamic tree takes part. SEM: kSemCpu for octree.c plane-test
 * returns and root acceptance, kSemParticle for the particle shader's (both
 * single tree; callers pass DUAL = false).
 */
template <bool CMP, bool DUAL, int SEM, bool GRID>
__device__ int trace(const SceneDev& s, const Ray& r, const Stack& st, Hit& out)
{
    TraceState t;
    Ray        rr  = r;
    int        stt = trace_begin<SEM>(s, rr, t);
    while (stt == ST_CONT) stt = trace_step<CMP, DUAL, SEM, GRID>(s, rr, st, t);
    out.exps = t.exps;
    if (stt == ST_HIT) {
        out.isp = t.entry;
        out.ns  = t.ns;
        out.nd  = t.nd;
        out.ix  = t.ix;
        out.iy  = t.iy;
        out.iz  = t.iz;
    }
    return stt;
}

} // namespace qbt
