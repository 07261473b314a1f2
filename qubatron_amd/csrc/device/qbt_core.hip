/*
 * qbt_core.hip - context, buffer uploads, validation, rendering and batched
 * ray queries of libqbt.so (gfx950).
 *
 * Device layout (per tree): int4 child[2*n] (the 8 child indices of a node in
 * one 32-byte aligned block) and int model[n] (slot 8), split from the
 * reference's 48-byte nodes (octree.c:11-14) while uploading. Per point
 * model: float4 ca[2*n] = {colour rgb 1, normal xyz 1}, so the leaf fetch of
 * a primary hit reads one 32-byte block.
 */
#include <hip/hip_runtime.h>

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <new>
#include <vector>

#include "qbt.h"
#include "qbt_internal.hpp"
#include "trace.hpp"
#include "pixel.hpp"

#pragma clang fp contract(off)

using namespace qbt;

/* ------------------------------------------------------------------------ */
/* kernels: upload scatter, validation, counters                            */
/* ------------------------------------------------------------------------ */

/* octree texels [t0, t0+count) from staging into child/model arrays */
__global__ void k_scatter_octree(const int4* __restrict__ stage, size_t t0, size_t count, int4* __restrict__ child,
                                 int* __restrict__ model)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    size_t t    = t0 + i;
    size_t node = t / 3;
    int    k    = (int) (t - node * 3);
    int4   v    = stage[i];
    if (k < 2) child[2 * node + k] = v;
    else model[node] = v.x;
}

/* held octree texels: texel idx[i] = val[i] */
__global__ void k_scatter_texels(const int4* __restrict__ val, const uint64_t* __restrict__ idx, size_t count,
                                 int4* __restrict__ child, int* __restrict__ model)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t t    = idx[i];
    const uint64_t node = t / 3;
    const int      k    = (int) (t - node * 3);
    const int4     v    = val[i];
    if (k < 2) child[2 * node + k] = v;
    else model[node] = v.x;
}

/* float3 items [i0, i0+count) into the .xyz of ca[2*i + which] */
__global__ void k_scatter_attr(const float* __restrict__ stage, size_t i0, size_t count, float4* __restrict__ ca,
                               int which)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    ca[2 * (i0 + i) + which] = make_float4(stage[3 * i], stage[3 * i + 1], stage[3 * i + 2], 1.0f);
}

__global__ void k_scatter_points(const float* __restrict__ stage, size_t i0, size_t count, float* __restrict__ xyz)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * count) return;
    xyz[3 * i0 + i] = stage[i];
}

/* device -> host image of a tree: texels [t0, t0+count) */
__global__ void k_gather_octree(int4* __restrict__ stage, size_t t0, size_t count, const int4* __restrict__ child,
                                const int* __restrict__ model)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    size_t t    = t0 + i;
    size_t node = t / 3;
    int    k    = (int) (t - node * 3);
    stage[i]    = k < 2 ? child[2 * node + k] : make_int4(model[node], 0, 0, 0);
}

__global__ void k_gather_attr(float* __restrict__ stage, size_t i0, size_t count, const float4* __restrict__ ca,
                              int which)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 v          = ca[2 * (i0 + i) + which];
    stage[3 * i]      = v.x;
    stage[3 * i + 1]  = v.y;
    stage[3 * i + 2]  = v.z;
}

/* every child index of nodes [lo, hi) must lie in [0, n) */
__global__ void k_validate(const int4* __restrict__ child, size_t lo, size_t hi, int n,
                           unsigned int* __restrict__ bad)
{
    size_t i = lo + (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    int4 a = child[2 * i], b = child[2 * i + 1];
    bool ok = (unsigned) a.x < (unsigned) n && (unsigned) a.y < (unsigned) n && (unsigned) a.z < (unsigned) n &&
              (unsigned) a.w < (unsigned) n && (unsigned) b.x < (unsigned) n && (unsigned) b.y < (unsigned) n &&
              (unsigned) b.z < (unsigned) n && (unsigned) b.w < (unsigned) n;
    if (!ok) atomicAdd(bad, 1u);
}

/* the same over a list of node ranges (total nodes in all of them) */
__global__ void k_validate_ranges(const int4* __restrict__ child, const uint2* __restrict__ ranges,
                                  const uint32_t* __restrict__ roff, int nr, uint32_t total, int n,
                                  unsigned int* __restrict__ bad)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const size_t i = range_item(ranges, roff, nr, g);
    int4         a = child[2 * i], b = child[2 * i + 1];
    bool ok = (unsigned) a.x < (unsigned) n && (unsigned) a.y < (unsigned) n && (unsigned) a.z < (unsigned) n &&
              (unsigned) a.w < (unsigned) n && (unsigned) b.x < (unsigned) n && (unsigned) b.y < (unsigned) n &&
              (unsigned) b.z < (unsigned) n && (unsigned) b.w < (unsigned) n;
    if (!ok) atomicAdd(bad, 1u);
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

/* plain streaming copy, grid-stride over 16-byte elements */
__global__ void __launch_bounds__(256) k_copy16(const int4* __restrict__ src, int4* __restrict__ dst, size_t n)
{
    size_t stride = (size_t) gridDim.x * blockDim.x;
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

/* ------------------------------------------------------------------------ */
/* pixel program (octree_fsh.c:381-464)                                     */
/* ------------------------------------------------------------------------ */

/* per-view, per-light sprite trace (octree_fsh.c:452-459): the ray from the
 * camera to the light is the same for every pixel of a view. */
template <bool CMP, bool DUAL, bool GRID>
__device__ __forceinline__ int sprite_flags(const SceneDev& s, const ViewDev& v, int l, const Stack& st)
{
    Ray r = {v.cam[0], v.cam[1], v.cam[2], v.light[l][0] - v.cam[0], v.light[l][1] - v.cam[1],
             v.light[l][2] - v.cam[2]};
    Hit h;
    int stt   = trace<CMP, DUAL, false, GRID>(s, r, st, h);
    float ix  = stt == ST_HIT ? h.isp.x : 0.0f;
    float resv = ix - v.cam[0];
    int   f   = 0;
    if (stt == ST_DISCARD) f |= 1;
    if (resv / r.dx > 1.0f) f |= 2;
    if (stt == ST_ABORT) f |= 4;
    return f;
}

template <bool CMP, bool DUAL, bool GRID>
__global__ void k_sprite(SceneDev s, BatchDev b, int* __restrict__ sprite)
{
    extern __shared__ uint32_t lds[];
    const int                  l  = threadIdx.x;
    const Stack                st = make_stack(lds, s.maxlevel, DUAL);
    const ViewDev&             v  = b.views[blockIdx.x];
    if (l >= v.nlights) return;
    sprite[blockIdx.x * QBT_MAX_LIGHTS_RENDER + l] = sprite_flags<CMP, DUAL, GRID>(s, v, l, st);
}

/* the same for one view whose uniforms come by value (direct layout) */
template <bool DUAL, bool GRID>
__global__ void k_sprite_view(SceneDev s, ViewDev v, int* __restrict__ sprite)
{
    extern __shared__ uint32_t lds[];
    const int   l  = threadIdx.x;
    const Stack st = make_stack(lds, s.maxlevel, DUAL);
    if (l >= v.nlights) return;
    sprite[l] = sprite_flags<false, DUAL, GRID>(s, v, l, st);
}

template <bool DUAL, bool GRID, bool COUNT>
__global__ void __launch_bounds__(256) k_render(SceneDev s, ViewDev v, const int* __restrict__ sprite, FrameDev f,
                                                unsigned mask, unsigned long long* __restrict__ counters)
{
    extern __shared__ uint32_t lds[];
    const Stack st = make_stack(lds, s.maxlevel, DUAL);

    /* tiles of 16 x (blockDim/16) pixels */
    const int tw = 16, th = blockDim.x / 16;
    const int tiles_x = (v.w + tw - 1) / tw;
    const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int px = tx * tw + (threadIdx.x % tw), py = ty * th + (threadIdx.x / tw);
    const bool valid = px < v.w && py < v.h;

    unsigned n_shadow = 0, n_hit = 0, n_exp = 0, n_exp_p = 0, n_abort = 0;
    if (valid) {
        const size_t o = (size_t) py * v.w + px;
        float        csv[3];
        pixel_dir(v, px, py, csv);
        Ray r = {v.cam[0], v.cam[1], v.cam[2], csv[0], csv[1], csv[2]};
        Hit h;
        int stt = trace<false, DUAL, false, GRID>(s, r, st, h); /* fused path: direct layout */
        if (COUNT) n_exp += h.exps, n_exp_p += h.exps;

        int32_t  hit = -1, leaf = -1, dyn = -1;
        uint32_t flg = 0, sbits = 0;
        float    col[4] = {0.f, 0.f, 0.f, 0.f};
        float4   isp    = make_float4(0.f, 0.f, 0.f, 0.f);
        if (stt == ST_ABORT) n_abort++;
        if (stt == ST_DISCARD) {
            flg |= QBT_PX_DISCARD;
        } else {
            float nrm[4] = {0.f, 0.f, 0.f, 0.f};
            if (stt == ST_HIT) {
                const int L  = s.maxlevel;
                int       ms = (h.ns == 0 && L > 0) ? 0 : s.stat.model[h.ns];
                int       md = 0;
                if (DUAL) md = (h.nd == 0 && L > 0) ? 0 : s.dyn.model[h.nd];
                fetch_attr(s.astat, ms, col, nrm);
                if (DUAL && md > 0) fetch_attr(s.adyn, md, col, nrm);
                flg |= QBT_PX_LEAF;
                hit  = ms;
                leaf = h.ns;
                dyn  = md;
                isp  = h.isp;
            }
            if (stt == ST_HIT && h.isp.w > 0.0f) {
                flg |= QBT_PX_SHADED;
                n_hit++;
                float nn[3], mc[3] = {-csv[0], -csv[1], -csv[2]}, nmc[3];
                normalize3(nrm, nn);
                normalize3(mc, nmc);
                float camv = gmax(dot3(nmc, nn), 0.0f);
                float fac  = 0.1f + 0.2f * camv;
                for (int l = 0; l < v.nlights; l++) {
                    Ray lr = {v.light[l][0], v.light[l][1], v.light[l][2], h.isp.x - v.light[l][0],
                              h.isp.y - v.light[l][1], h.isp.z - v.light[l][2]};
                    Hit lh;
                    int ls = trace<false, DUAL, false, GRID>(s, lr, st, lh);
                    n_shadow++;
                    if (COUNT) n_exp += lh.exps;
                    if (ls == ST_ABORT) n_abort++;
                    if (ls == ST_DISCARD) {
                        flg |= QBT_PX_DISCARD;
                        break;
                    }
                    float4 li   = ls == ST_HIT ? lh.isp : make_float4(0.f, 0.f, 0.f, 0.f);
                    float  ix   = li.x - h.isp.x, iy = li.y - h.isp.y, iz = li.z - h.isp.z;
                    float  sqr  = ix * ix + iy * iy + iz * iz;
                    float  bit  = (15.0f < sqr) ? 0.0f : 1.0f; /* step(sqr, 15.0) */
                    if (bit != 0.0f) sbits |= 1u << l;
                    float ml[3] = {-lr.dx, -lr.dy, -lr.dz}, nml[3];
                    normalize3(ml, nml);
                    float lam = gmax(dot3(nml, nn), 0.0f);
                    fac       = fac + lam * bit * 0.7f;
                }
                if (!(flg & QBT_PX_DISCARD)) {
                    col[0] = col[0] * fac;
                    col[1] = col[1] * fac;
                    col[2] = col[2] * fac;
                    col[2] *= 0.7f;
                    float g = (float) v.shoot * camv * 0.1f;
                    col[0] += g;
                    col[1] += g;
                    col[2] += g;
                }
            }
            if (!(flg & QBT_PX_DISCARD) && v.nlights > 0) {
                float ncsv[3];
                normalize3(csv, ncsv);
                for (int l = 0; l < v.nlights; l++) {
                    if (dot3(v.cl[l], ncsv) > v.cos_thr) {
                        int sf = sprite[l];
                        if (sf & 1) {
                            flg |= QBT_PX_DISCARD;
                            break;
                        }
                        if (sf & 2) {
                            col[0] = col[1] = col[2] = col[3] = 1.0f;
                            flg |= QBT_PX_SPRITE;
                        }
                    }
                }
            }
        }
        if (flg & QBT_PX_DISCARD) {
            col[0] = col[1] = col[2] = col[3] = 0.0f;
            sbits                             = 0;
        }
        if (mask & QBT_OUT_HIT) f.hit[o] = hit;
        if (mask & QBT_OUT_LEAF) f.leaf[o] = leaf;
        if (mask & QBT_OUT_DYN) f.dyn[o] = dyn;
        if (mask & QBT_OUT_SHADOW) f.shadow[o] = (uint8_t) sbits;
        if (mask & QBT_OUT_FLAGS) f.flags[o] = (uint8_t) flg;
        if (mask & QBT_OUT_RGBA8) f.rgba8[o] = pack_rgba8(col[0], col[1], col[2], col[3]);
        if (mask & QBT_OUT_RGBAF) f.rgbaf[o] = make_float4(col[0], col[1], col[2], col[3]);
        if (mask & QBT_OUT_ISP) f.isp[o] = isp;
        if (COUNT && (mask & QBT_OUT_EXP)) f.exp[o] = (int) n_exp;
    }
    /* per-wave totals: shadow traces, shaded pixels, expansions, aborts */
    n_shadow = wave_sum(n_shadow);
    n_hit    = wave_sum(n_hit);
    n_abort  = wave_sum(n_abort);
    if (COUNT) n_exp = wave_sum(n_exp), n_exp_p = wave_sum(n_exp_p);
    if ((threadIdx.x & 63) == 0) {
        if (n_shadow) atomicAdd(&counters[CNT_SHADOW], (unsigned long long) n_shadow);
        if (COUNT && n_exp_p) atomicAdd(&counters[CNT_EXP_PRIM], (unsigned long long) n_exp_p);
        if (n_hit) atomicAdd(&counters[CNT_HITS], (unsigned long long) n_hit);
        if (n_abort) atomicAdd(&counters[CNT_ABORT], (unsigned long long) n_abort);
        if (COUNT && n_exp) atomicAdd(&counters[CNT_EXP], (unsigned long long) n_exp);
    }
}

/* batched octree_trace_line (SEM: kSemShader, kSemCpu or kSemParticle) */
template <bool CMP, bool DUAL, int SEM, bool GRID>
__global__ void __launch_bounds__(256) k_trace_rays(SceneDev s, const float* __restrict__ org,
                                                     const float* __restrict__ dir, size_t n, int* __restrict__ status,
                                                     int* __restrict__ model, int* __restrict__ leaf,
                                                     float4* __restrict__ isp, float4* __restrict__ tlf,
                                                     unsigned long long* __restrict__ counters)
{
    extern __shared__ uint32_t lds[];
    const Stack st = make_stack(lds, s.maxlevel, DUAL);
    size_t      i  = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    unsigned    ab = 0;
    if (i < n) {
        Ray r = {org[3 * i], org[3 * i + 1], org[3 * i + 2], dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
        Hit h;
        int stt = trace<CMP, DUAL, SEM, GRID>(s, r, st, h);
        if (stt == ST_ABORT) ab = 1;
        int    m = -1, lf = -1;
        float4 ip = make_float4(0.f, 0.f, 0.f, 0.f), tc = ip;
        if (stt == ST_HIT) {
            const int L = s.maxlevel;
            m           = (h.ns == 0 && L > 0) ? 0 : s.stat.model[h.ns];
            if (SEM == kSemParticle && particle_model_lost(h.ns)) m = 0; /* its slot-8 address */
            lf          = CMP ? s.stat.ref[h.ns] : h.ns;
            ip          = h.isp;
            float x, y, z, w;
            cube_at<GRID>(s, L, h.ix, h.iy, h.iz, x, y, z, w);
            tc = make_float4(x, y, z, w);
        }
        if (status) status[i] = stt == ST_ABORT ? -2 : stt;
        if (model) model[i] = m;
        if (leaf) leaf[i] = lf;
        if (isp) isp[i] = ip;
        if (tlf) tlf[i] = tc;
    }
    ab = wave_sum(ab);
    if ((threadIdx.x & 63) == 0 && ab) atomicAdd(&counters[CNT_ABORT], (unsigned long long) ab);
}

/* The particle shader's main (particle_vsh.c:345-373), one particle per
 * thread: gravity step, a trace of the static tree (direct layout: the
 * shader's slot address is a function of the reference node index) with
 * that shader's semantics, stop at the leaf corner when it is closer than
 * 10 units, else move. pos/spd may alias pos_out/spd_out. */
template <bool GRID>
__global__ void __launch_bounds__(256) k_particles(SceneDev s, const float* pos, const float* spd, size_t n,
                                                   float* pos_out, float* spd_out,
                                                   unsigned long long* __restrict__ counters)
{
    extern __shared__ uint32_t lds[];
    const Stack  st = make_stack(lds, s.maxlevel, false);
    const size_t i  = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    unsigned     ab = 0;
    if (i < n) {
        const float p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        float       v[3] = {spd[3 * i], spd[3 * i + 1], spd[3 * i + 2]};
        float       o[3] = {p[0], p[1], p[2]};
        if (v[0] > -90000.0f) {
            v[1] -= 0.4f;
            Ray r = {p[0], p[1], p[2], v[0], v[1], v[2]};
            Hit h;
            const int stt  = trace<false, false, kSemParticle, GRID>(s, r, st, h);
            bool      stop = false;
            if (stt == ST_ABORT) ab = 1;
            if (stt == ST_HIT && h.isp.w > 0.0f) {
                float x, y, z, w;
                cube_at<GRID>(s, s.maxlevel, h.ix, h.iy, h.iz, x, y, z, w);
                const float dx = x - p[0], dy = y - p[1], dz = z - p[2];
                if (sqrtf(dx * dx + dy * dy + dz * dz) < 10.0f) {
                    o[0] = x, o[1] = y, o[2] = z;
                    v[0] = -100000.0f;
                    stop = true;
                }
            }
            if (!stop) {
                o[0] = p[0] + v[0], o[1] = p[1] + v[1], o[2] = p[2] + v[2];
                if (o[1] < -10.0f) v[0] = -100000.0f;
            }
        }
        for (int a = 0; a < 3; a++) {
            pos_out[3 * i + a] = o[a];
            spd_out[3 * i + a] = v[a];
        }
    }
    ab = wave_sum(ab);
    if ((threadIdx.x & 63) == 0 && ab) atomicAdd(&counters[CNT_ABORT], (unsigned long long) ab);
}

/* ------------------------------------------------------------------------ */
/* host side                                                                */
/* ------------------------------------------------------------------------ */

#define HIPCHK(x) QBT_HIPCHK(x)

static char g_last_hip[256];

/* ------------------------------------------------------------------------ */
/* scratch of qbt_render_noalloc                                            */
/* ------------------------------------------------------------------------ */

/* A static array on every GPU the library runs on holds what the fused
 * render path needs besides its inputs and outputs: the counters (CNT_N x
 * u64), the word of the node-array check and the sprite flags of up to
 * QBT_MAX_VIEWS views (QBT_MAX_LIGHTS_RENDER ints each). It lives until the
 * process ends, and every qbt_render_noalloc call on a GPU shares that GPU's
 * copy: calls that overlap on one GPU race. */
constexpr size_t kScrCounters  = 0;
constexpr size_t kScrValFlag   = 256;
constexpr size_t kScrSprite    = 512;
constexpr size_t kScratchBytes = kScrSprite + sizeof(int) * QBT_MAX_LIGHTS_RENDER * QBT_MAX_VIEWS;
static_assert(sizeof(unsigned long long) * CNT_N <= kScrValFlag, "counters fit before the check word");
static_assert(kScratchBytes == QBT_NOALLOC_SCRATCH_BYTES, "qbt.h states the scratch size");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

/* Device address of `scratch` on device `dev`, which must be current: each
 * GPU holds its own copy at its own address, looked up once per device. On
 * the host the array's name is only a stand-in, never a device address. */
static std::mutex         g_scratch_mu;
static std::vector<void*> g_scratch_ptr; /* by device ordinal */

static int scratch_address(int dev, void** out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int) g_scratch_ptr.size() <= dev) g_scratch_ptr.resize((size_t) dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void*      p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            qbt_set_hip_error(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
            (void) hipGetLastError();
            return QBT_EDEVICE;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return QBT_OK;
}

void qbt_set_hip_error(hipError_t e, const char* what, const char* file, int line)
{
    snprintf(g_last_hip, sizeof(g_last_hip), "%s: %s (%s:%d)", hipGetErrorString(e), what, file, line);
}

extern "C" const char* qbt_last_hip_error(void) { return g_last_hip; }

int qbt_ctx::dev_alloc(DevBuf& b, size_t bytes)
{
    if (bytes <= b.cap) return QBT_OK;
    size_t grow = bytes - b.cap;
    /* over the cap: first release the builder's idle kept blocks */
    if (max_bytes && held + grow > max_bytes && !(builder_trim() && held + grow <= max_bytes)) return QBT_ENOMEM;
    if (b.p) {
        HIPCHK(hipFree(b.p));
        held -= b.cap;
        b.p   = nullptr;
        b.cap = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        (void) hipGetLastError();
        e = builder_trim() ? hipMalloc(&b.p, bytes) : e; /* the device is full: the same, then once more */
        if (e != hipSuccess) {
            (void) hipGetLastError();
            b.p = nullptr;
            return QBT_ENOMEM;
        }
    }
    b.cap = bytes;
    held += bytes;
    return QBT_OK;
}

void qbt_ctx::dev_free(DevBuf& b)
{
    if (b.p) {
        (void) hipFree(b.p);
        held -= b.cap;
    }
    b.p   = nullptr;
    b.cap = 0;
}

static const size_t kStageBytes = 64u << 20;

extern "C" int qbt_create(const qbt_opts* opts, qbt_ctx** out)
{
    if (!out) return QBT_EINVAL;
    *out    = nullptr;
    int dev = opts ? opts->device : 0;
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0) {
        (void) hipGetLastError();
        return QBT_EDEVICE;
    }
    if (dev < 0 || dev >= cnt) return QBT_EINVAL;
    qbt_ctx* c = new (std::nothrow) qbt_ctx();
    if (!c) return QBT_ENOMEM;
    c->device    = dev;
    c->max_bytes = opts ? opts->max_device_bytes : 0;
    if (hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        c->num_cus <= 0) {
        (void) hipGetLastError();
        c->num_cus = 256;
    }
    if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev2) != hipSuccess) {
        (void) hipGetLastError();
        delete c;
        return QBT_EDEVICE;
    }
    if (scratch_address(dev, &c->scratch)) {
        qbt_destroy(c);
        return QBT_EDEVICE;
    }
    if (c->dev_alloc(c->stage, kStageBytes) || c->dev_alloc(c->counters, sizeof(unsigned long long) * CNT_N) ||
        c->dev_alloc(c->sprite, sizeof(int) * QBT_MAX_LIGHTS_RENDER * QBT_MAX_VIEWS)) {
        qbt_destroy(c);
        return QBT_ENOMEM;
    }
    *out = c;
    return QBT_OK;
}

extern "C" void qbt_destroy(qbt_ctx* c)
{
    if (!c) return;
    (void) hipSetDevice(c->device);
    if (c->stream) (void) hipStreamSynchronize(c->stream);
    for (int t = 0; t < 2; t++) {
        c->dev_free(c->tree[t].child);
        c->dev_free(c->tree[t].model);
        c->dev_free(c->tree[t].cmp.rec);
        c->dev_free(c->tree[t].cmp.model);
        c->dev_free(c->tree[t].cmp.ref);
        c->dev_free(c->tree[t].cmp.r2c);
        c->dev_free(c->tree[t].cmp.scratch);
        c->dev_free(c->attr[t].ca);
        c->dev_free(c->pts[t].xyz);
    }
    for (auto& fr : c->frames)
        for (int k = 0; k < FR_N; k++) c->dev_free(fr.buf[k]);
    c->dev_free(c->stage);
    c->dev_free(c->counters);
    c->dev_free(c->sprite);
    c->dev_free(c->rays);
    c->graph_free();
    for (DevBuf* b : {&c->wb_isp, &c->wb_ms, &c->wb_md, &c->wb_leaf, &c->wb_flags, &c->wb_disc, &c->wb_lstat,
                      &c->wb_exp, &c->wb_cand, &c->wb_items, &c->wb_q, &c->wb_lists, &c->wb_fb, &c->wb_gstack,
                      &c->wb_bcnt, &c->wb_lexp, &c->wb_slots,
                      &c->bt_tab, &c->valflag, &c->valrg})
        c->dev_free(*b);
    for (int k = 0; k < qbt_ctx::kTabSlots; k++) {
        if (c->tab_host[k]) (void) hipHostFree(c->tab_host[k]);
        if (c->tab_ev[k]) (void) hipEventDestroy(c->tab_ev[k]);
    }
    c->mlp_free();
    c->builder_free();
    for (hipEvent_t e : c->prof_ev) (void) hipEventDestroy(e);
    if (c->ev0) (void) hipEventDestroy(c->ev0);
    if (c->ev1) (void) hipEventDestroy(c->ev1);
    if (c->ev2) (void) hipEventDestroy(c->ev2);
    if (c->stream) (void) hipStreamDestroy(c->stream);
    delete c;
}

extern "C" size_t qbt_device_bytes(const qbt_ctx* c) { return c ? c->held : 0; }

extern "C" const char* qbt_strerror(int code)
{
    switch (code) {
    case QBT_OK: return "ok";
    case QBT_EINVAL: return "invalid argument";
    case QBT_ENOMEM: return "out of memory (or device allocation cap reached)";
    case QBT_ERANGE: return "value out of range";
    case QBT_EIO: return "file missing, short or unreadable";
    case QBT_EDEVICE: return "no usable GPU or HIP runtime error";
    case QBT_ETREE: return "octree references a node out of range";
    case QBT_EABORT: return "a ray exceeded the traversal iteration cap";
    case QBT_ENOTSUP: return "not supported";
    default: return "unknown error";
    }
}

extern "C" const char* qbt_version(void) { return "qubatron-mi355x 0.1 (gfx950)"; }

static inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned) ((n + bs - 1) / bs); }

/* octree range uploads of at most this many texels are held on the host
 * and applied together (the reference uploads one 48-byte node per edited
 * node, modelutil.c:425-500: hundreds of calls per edit) */
static const size_t kSmallTexels = 96;
static const size_t kValRanges   = 65536; /* written node ranges checked one by one */
static const size_t kPendMax     = kStageBytes / 24; /* 16-byte value + 8-byte index each */

int qbt_ctx::flush_pending(int t, bool discard)
{
    TreeBuf&     tb = tree[t];
    const size_t n  = tb.pend_idx.size();
    if (n && !discard) {
        /* values, then indices, through the staging buffer in one copy */
        std::vector<uint8_t> pack(n * 24);
        memcpy(pack.data(), tb.pend_val.data(), n * 16);
        memcpy(pack.data() + n * 16, tb.pend_idx.data(), n * 8);
        HIPCHK(hipMemcpyAsync(stage.p, pack.data(), n * 24, hipMemcpyHostToDevice, stream));
        k_scatter_texels<<<blocks_for(n, 256), 256, 0, stream>>>((const int4*) stage.p,
                                                                 (const uint64_t*) ((char*) stage.p + n * 16), n,
                                                                 (int4*) tb.child.p, (int*) tb.model.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream)); /* the staging buffer is free again */
    }
    tb.pend_idx.clear();
    tb.pend_val.clear();
    tb.pend_pos.clear();
    return QBT_OK;
}

/* logical/capacity bookkeeping of one upload target */
extern "C" int qbt_upload(qbt_ctx* c, const void* data, int type, size_t size, size_t itemsize, size_t start,
                          size_t end, int buftype)
{
    if (!c || buftype < 0 || buftype > QBT_BUFFER_DYNAMIC_POINTS) return QBT_EINVAL;
    const bool octree = buftype == QBT_BUFFER_STATIC_OCTREE || buftype == QBT_BUFFER_DYNAMIC_OCTREE;
    if (octree ? (type != QBT_TYPE_INT || itemsize != 16) : (type != QBT_TYPE_FLOAT || itemsize != 12))
        return QBT_EINVAL;
    if (start > end || end > size || start % itemsize || end % itemsize || size % itemsize) return QBT_EINVAL;
    if (octree && size % 48) return QBT_EINVAL; /* whole 48-byte nodes */
    if (!data && end > start) return QBT_EINVAL;
    if (size / itemsize > (size_t) INT32_MAX * 3) return QBT_ENOTSUP;
    HIPCHK(hipSetDevice(c->device));

    const int t     = (buftype == QBT_BUFFER_STATIC_COLOR || buftype == QBT_BUFFER_STATIC_NORMAL ||
                   buftype == QBT_BUFFER_STATIC_OCTREE || buftype == QBT_BUFFER_STATIC_POINTS)
                          ? 0
                          : 1;
    int       rc    = QBT_OK;
    size_t    items = size / itemsize;
    if (octree) {
        TreeBuf& tb    = c->tree[t];
        size_t   nodes = items / 3;
        if (nodes > INT32_MAX) return QBT_ENOTSUP;
        const size_t t0r = start / 16, t1r = end / 16;
        const bool   small = nodes <= tb.cap_nodes && t1r > t0r && t1r - t0r <= kSmallTexels;
        /* held uploads go before a direct copy (call order); a regrowth's full
         * copy supersedes them */
        if (!small && (rc = c->flush_pending(t, nodes > tb.cap_nodes))) return rc;
        if (nodes > tb.cap_nodes) { /* regrowth: reallocate and copy everything (octree_glc.c:412-428) */
            size_t cap = nodes + nodes / 4 + 1024;
            c->dev_free(tb.child);
            c->dev_free(tb.model);
            if ((rc = c->dev_alloc(tb.child, cap * 32)) || (rc = c->dev_alloc(tb.model, cap * 4))) {
                tb.cap_nodes = 0;
                tb.n         = 0;
                return rc;
            }
            tb.cap_nodes = cap;
            start        = 0;
            end          = size;
            tb.full_check = true;
            tb.cmp_stale  = true;
        }
        if (nodes < tb.n) tb.full_check = tb.cmp_stale = true;
        tb.n = nodes;
        /* copy texels [start/16, end/16) through the staging buffer */
        size_t t0 = start / 16, t1 = end / 16;
        if (t1 > t0) {
            size_t lo = t0 / 3, hi = (t1 + 2) / 3;
            if (tb.dirty_hi <= tb.dirty_lo) tb.dirty_lo = lo, tb.dirty_hi = hi;
            else tb.dirty_lo = lo < tb.dirty_lo ? lo : tb.dirty_lo, tb.dirty_hi = hi > tb.dirty_hi ? hi : tb.dirty_hi;
            if (tb.vdirty.size() < kValRanges) tb.vdirty.emplace_back(lo, hi);
            else tb.vdirty_over = true;
        }
        tb.validated = false;
        /* the compact layout: relayout after a whole-array upload, else patch
         * the written node range in place at the next render */
        if (t1 > t0 && !tb.cmp_stale) {
            const size_t lo = t0 / 3, hi = (t1 + 2) / 3;
            if (lo == 0 && hi >= nodes) tb.cmp_stale = true;
            else if (tb.patch.size() >= 65536) tb.cmp_stale = true; /* as much bookkeeping as a relayout */
            else tb.patch.emplace_back(lo, hi);
        }
        if (small) { /* hold the texels (copied now: the caller may change its array) */
            const int4* src = (const int4*) ((const char*) data + t0 * 16);
            for (size_t a = t0; a < t1; a++) {
                auto it = tb.pend_pos.find(a);
                if (it != tb.pend_pos.end()) {
                    tb.pend_val[it->second] = src[a - t0];
                } else {
                    tb.pend_pos.emplace(a, (uint32_t) tb.pend_idx.size());
                    tb.pend_idx.push_back(a);
                    tb.pend_val.push_back(src[a - t0]);
                }
            }
            if (tb.pend_idx.size() >= kPendMax && (rc = c->flush_pending(t))) return rc;
            return QBT_OK;
        }
        const size_t per = kStageBytes / 16;
        for (size_t a = t0; a < t1; a += per) {
            size_t cnt = t1 - a < per ? t1 - a : per;
            HIPCHK(hipMemcpyAsync(c->stage.p, (const char*) data + a * 16, cnt * 16, hipMemcpyHostToDevice, c->stream));
            k_scatter_octree<<<blocks_for(cnt, 256), 256, 0, c->stream>>>((const int4*) c->stage.p, a, cnt,
                                                                          (int4*) tb.child.p, (int*) tb.model.p);
            HIPCHK(hipGetLastError());
        }
    } else if (buftype == QBT_BUFFER_STATIC_POINTS || buftype == QBT_BUFFER_DYNAMIC_POINTS) {
        PointsBuf& pb = c->pts[t];
        if (items > pb.cap) {
            size_t cap = items + items / 4 + 1024;
            c->dev_free(pb.xyz);
            if ((rc = c->dev_alloc(pb.xyz, cap * 12))) {
                pb.cap = pb.n = 0;
                return rc;
            }
            pb.cap = cap;
            start  = 0;
            end    = size;
        }
        pb.n             = items;
        size_t       i0  = start / 12, i1 = end / 12;
        const size_t per = kStageBytes / 12;
        for (size_t a = i0; a < i1; a += per) {
            size_t cnt = i1 - a < per ? i1 - a : per;
            HIPCHK(hipMemcpyAsync(c->stage.p, (const char*) data + a * 12, cnt * 12, hipMemcpyHostToDevice, c->stream));
            k_scatter_points<<<blocks_for(3 * cnt, 256), 256, 0, c->stream>>>((const float*) c->stage.p, a, cnt,
                                                                             (float*) pb.xyz.p);
            HIPCHK(hipGetLastError());
        }
    } else {
        AttrBuf& ab    = c->attr[t];
        int      which = (buftype == QBT_BUFFER_STATIC_NORMAL || buftype == QBT_BUFFER_DYNAMIC_NORMAL) ? 1 : 0;
        if (items > INT32_MAX) return QBT_ENOTSUP;
        if (items > ab.cap) {
            size_t cap = items + items / 4 + 1024;
            /* the other attribute of the pair lives in the same array: keep it */
            DevBuf nb;
            if ((rc = c->dev_alloc(nb, cap * 32))) return rc;
            k_fill_u32<<<blocks_for(cap * 8, 256), 256, 0, c->stream>>>((uint32_t*) nb.p, cap * 8, 0u);
            HIPCHK(hipGetLastError());
            if (ab.ca.p && ab.n) HIPCHK(hipMemcpyAsync(nb.p, ab.ca.p, ab.n * 32, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            c->dev_free(ab.ca);
            ab.ca  = nb;
            ab.cap = cap;
            start  = 0;
            end    = size;
        }
        size_t& cnt_which = which ? ab.n_nrm : ab.n_col;
        cnt_which         = items;
        ab.n              = ab.n_col > ab.n_nrm ? ab.n_col : ab.n_nrm;
        size_t       i0   = start / 12, i1 = end / 12;
        const size_t per  = kStageBytes / 12;
        for (size_t a = i0; a < i1; a += per) {
            size_t cnt = i1 - a < per ? i1 - a : per;
            HIPCHK(hipMemcpyAsync(c->stage.p, (const char*) data + a * 12, cnt * 12, hipMemcpyHostToDevice, c->stream));
            k_scatter_attr<<<blocks_for(cnt, 256), 256, 0, c->stream>>>((const float*) c->stage.p, a, cnt,
                                                                        (float4*) ab.ca.p, which);
            HIPCHK(hipGetLastError());
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return QBT_OK;
}

extern "C" int qbt_download(qbt_ctx* c, int buftype, void* host, size_t start, size_t end)
{
    if (!c || !host || start > end || buftype < 0 || buftype > QBT_BUFFER_DYNAMIC_POINTS) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    const bool octree = buftype == QBT_BUFFER_STATIC_OCTREE || buftype == QBT_BUFFER_DYNAMIC_OCTREE;
    const int  t      = (buftype == QBT_BUFFER_STATIC_COLOR || buftype == QBT_BUFFER_STATIC_NORMAL ||
                   buftype == QBT_BUFFER_STATIC_OCTREE || buftype == QBT_BUFFER_STATIC_POINTS)
                            ? 0
                            : 1;
    if (octree) {
        const TreeBuf& tb = c->tree[t];
        if (start % 16 || end % 16 || end > tb.n * 48) return QBT_EINVAL;
        int rc;
        if ((rc = c->flush_pending(t))) return rc;
        size_t       t0 = start / 16, t1 = end / 16;
        const size_t per = kStageBytes / 16;
        for (size_t a = t0; a < t1; a += per) {
            size_t cnt = t1 - a < per ? t1 - a : per;
            k_gather_octree<<<blocks_for(cnt, 256), 256, 0, c->stream>>>((int4*) c->stage.p, a, cnt,
                                                                         (const int4*) tb.child.p, (const int*) tb.model.p);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync((char*) host + (a - t0) * 16, c->stage.p, cnt * 16, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        return QBT_OK;
    }
    if (start % 12 || end % 12) return QBT_EINVAL;
    size_t i0 = start / 12, i1 = end / 12;
    if (buftype == QBT_BUFFER_STATIC_POINTS || buftype == QBT_BUFFER_DYNAMIC_POINTS) {
        if (i1 > c->pts[t].n) return QBT_EINVAL;
        if (i1 > i0)
            HIPCHK(hipMemcpy(host, (const char*) c->pts[t].xyz.p + i0 * 12, (i1 - i0) * 12, hipMemcpyDeviceToHost));
        return QBT_OK;
    }
    const AttrBuf& ab    = c->attr[t];
    int            which = (buftype == QBT_BUFFER_STATIC_NORMAL || buftype == QBT_BUFFER_DYNAMIC_NORMAL) ? 1 : 0;
    if (i1 > (which ? ab.n_nrm : ab.n_col)) return QBT_EINVAL;
    const size_t per = kStageBytes / 12;
    for (size_t a = i0; a < i1; a += per) {
        size_t cnt = i1 - a < per ? i1 - a : per;
        k_gather_attr<<<blocks_for(cnt, 256), 256, 0, c->stream>>>((float*) c->stage.p, a, cnt,
                                                                   (const float4*) ab.ca.p, which);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync((char*) host + (a - i0) * 12, c->stage.p, cnt * 12, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return QBT_OK;
}

/* Check that the child indices of dirty node ranges stay inside the tree. */
int qbt_ctx::validate_tree(int t, unsigned* flag)
{
    TreeBuf& tb = tree[t];
    if (tb.validated) return QBT_OK;
    if (tb.n == 0) {
        tb.validated = true;
        return QBT_OK;
    }
    size_t lo = tb.full_check ? 0 : tb.dirty_lo, hi = tb.full_check ? tb.n : (tb.dirty_hi < tb.n ? tb.dirty_hi : tb.n);
    const bool own_flag_given = flag != nullptr;
    if (!flag) {
        if (!valflag.p) {
            int rc = dev_alloc(valflag, 256);
            if (rc) return rc;
        }
        flag = (unsigned*) valflag.p;
    }
    k_fill_u32<<<1, 64, 0, stream>>>((uint32_t*) flag, 1, 0u);
    /* scattered writes (an edit touches nodes all over the array): check the
     * written ranges only, when that is much less than their bounding range
     * (not for a caller's flag word: that path allocates nothing) */
    std::vector<uint2>    rg;
    std::vector<uint32_t> roff;
    uint64_t              total = 0;
    if (!tb.full_check && !tb.vdirty_over && !own_flag_given && tb.vdirty.size() > 1) {
        std::sort(tb.vdirty.begin(), tb.vdirty.end());
        size_t a = 0, b = 0;
        for (size_t k = 0; k <= tb.vdirty.size(); k++) {
            const bool flush = k == tb.vdirty.size() || tb.vdirty[k].first > b;
            if (flush && b > a) {
                const size_t bb = b < tb.n ? b : tb.n;
                if (bb > a) {
                    rg.push_back(make_uint2((uint32_t) a, (uint32_t) bb));
                    roff.push_back((uint32_t) total);
                    total += bb - a;
                }
            }
            if (k == tb.vdirty.size()) break;
            if (flush) a = tb.vdirty[k].first, b = tb.vdirty[k].second;
            else b = tb.vdirty[k].second > b ? tb.vdirty[k].second : b;
        }
        if (total * 4 > hi - lo) rg.clear(); /* not worth a second kernel shape */
    }
    if (!rg.empty()) {
        const size_t bytes = rg.size() * 12;
        if (int rc = dev_alloc(valrg, bytes)) return rc;
        std::vector<uint8_t> pack(bytes);
        memcpy(pack.data(), rg.data(), rg.size() * 8);
        memcpy(pack.data() + rg.size() * 8, roff.data(), rg.size() * 4);
        HIPCHK(hipMemcpyAsync(valrg.p, pack.data(), bytes, hipMemcpyHostToDevice, stream));
        k_validate_ranges<<<blocks_for(total, 256), 256, 0, stream>>>(
            (const int4*) tb.child.p, (const uint2*) valrg.p, (const uint32_t*) ((char*) valrg.p + rg.size() * 8),
            (int) rg.size(), (uint32_t) total, (int) tb.n, (unsigned int*) flag);
    } else if (hi > lo) {
        k_validate<<<blocks_for(hi - lo, 256), 256, 0, stream>>>((const int4*) tb.child.p, lo, hi, (int) tb.n,
                                                                 (unsigned int*) flag);
    }
    HIPCHK(hipGetLastError());
    unsigned int nbad = 0;
    HIPCHK(hipMemcpyAsync(&nbad, flag, 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream)); /* also frees the host copy of the ranges */
    if (nbad) return QBT_ETREE;
    tb.validated  = true;
    tb.full_check = false;
    tb.dirty_lo = tb.dirty_hi = 0;
    tb.vdirty.clear();
    tb.vdirty_over = false;
    return QBT_OK;
}

/* every subdivision coordinate is an exact fp32 multiple of the leaf edge */
bool qbt_grid_exact(const float b[4], int L)
{
    float u = ldexpf(b[3], -L);
    if (!(u >= FLT_MIN) || !(b[3] > 0.0f) || !isfinite(b[3])) return false;
    const double lim = 16777216.0;
    const double sgn[3] = {1.0, -1.0, -1.0};
    for (int k = 0; k < 3; k++) {
        double q  = (double) b[k] / (double) u;
        double qi = nearbyint(q);
        if (qi != q || qi * (double) u != (double) b[k]) return false;
        double far = qi + sgn[k] * ldexp(1.0, L);
        if (fabs(qi) > lim || fabs(far) > lim) return false;
    }
    return true;
}

int qbt_ctx::scene_for(const float basecube[4], int maxlevel, SceneDev& s, bool& dual, bool& grid, bool& cmp,
                       bool allow_compact, unsigned* valflag_dev)
{
    if (maxlevel < 0 || maxlevel > kMaxLevel) return QBT_EINVAL;
    if (!(basecube[3] > 0.0f) || !isfinite(basecube[0]) || !isfinite(basecube[1]) || !isfinite(basecube[2]) ||
        !isfinite(basecube[3]))
        return QBT_EINVAL;
    if (!(ldexpf(basecube[3], -maxlevel) >= FLT_MIN)) return QBT_EINVAL;
    if (tree[0].n == 0) return QBT_EINVAL; /* no static octree uploaded */
    int rc;
    if ((rc = flush_pending(0)) || (rc = flush_pending(1))) return rc;
    if ((rc = validate_tree(0, valflag_dev))) return rc;
    if ((rc = validate_tree(1, valflag_dev))) return rc;
    dual = tree[1].n > 1 || (maxlevel == 0 && tree[1].n >= 1);
    grid = qbt_grid_exact(basecube, maxlevel);
    /* compact layout: rebuilt lazily after the node arrays change */
    cmp = false;
    if (allow_compact && use_compact) {
        for (int t = 0; t < (dual ? 2 : 1); t++) {
            /* range uploads since the layout: patched in place; a relayout
             * when the array was replaced or the patch cannot represent them */
            TreeBuf& tb   = tree[t];
            bool     need = tb.cmp_stale;
            if (!need && !tb.patch.empty()) {
                bool done = false;
                if ((rc = patch_layout(t, &done))) return rc;
                need = !done;
            }
            if (need) {
                if ((rc = relayout(t))) return rc;
                tb.cmp_stale = false;
            }
        }
        cmp = tree[0].cmp.ok && (!dual || tree[1].cmp.ok);
    }
    memset(&s, 0, sizeof(s));
    for (int t = 0; t < 2; t++) {
        TreeDev& d = t ? s.dyn : s.stat;
        if (cmp) {
            const CompactTree& cm = tree[t].cmp;
            d = {nullptr, (const int*) cm.model.p, (int) cm.n, (const uint2*) cm.rec.p, (const int*) cm.ref.p};
        } else {
            d = {(const int4*) tree[t].child.p, (const int*) tree[t].model.p, (int) tree[t].n, nullptr, nullptr};
        }
    }
    s.astat = {(const float4*) attr[0].ca.p, (int) attr[0].n};
    s.adyn  = {(const float4*) attr[1].ca.p, (int) attr[1].n};
    s.bx = basecube[0], s.by = basecube[1], s.bz = basecube[2], s.bw = basecube[3];
    s.maxlevel = maxlevel;
    s.iter_cap = iter_cap;
    return QBT_OK;
}

size_t qbt_stack_bytes(int L, bool dual, int threads) { return (size_t) threads * (size_t) (L > 0 ? L : 1) * (dual ? 12 : 8); }

/* view uniforms: the camera quaternions are computed once per view with the
 * host's sinf/cosf (octree_fsh.c:406-408) */
void qbt_make_view(const qbt_view& q, const float* lights, int nlights, ViewDev& v)
{
    memset(&v, 0, sizeof(v));
    v.cam[0] = q.camfp[0], v.cam[1] = q.camfp[1], v.cam[2] = q.camfp[2];
    const float ay[3] = {0.0f, 1.0f, 0.0f}, mx[3] = {-1.0f, 0.0f, 0.0f};
    auto axis = [](const float a[3], float angle, float out[4]) {
        float h = angle * 0.5f;
        out[0]  = a[0] * sinf(h);
        out[1]  = a[1] * sinf(h);
        out[2]  = a[2] * sinf(h);
        out[3]  = cosf(h);
    };
    auto rot = [](const float qq[4], const float vv[3], float r[3]) {
        float c1[3] = {qq[1] * vv[2] - qq[2] * vv[1], qq[2] * vv[0] - qq[0] * vv[2], qq[0] * vv[1] - qq[1] * vv[0]};
        float s[3]  = {c1[0] + qq[3] * vv[0], c1[1] + qq[3] * vv[1], c1[2] + qq[3] * vv[2]};
        float c2[3] = {qq[1] * s[2] - qq[2] * s[1], qq[2] * s[0] - qq[0] * s[2], qq[0] * s[1] - qq[1] * s[0]};
        r[0]        = vv[0] + 2.0f * c2[0];
        r[1]        = vv[1] + 2.0f * c2[1];
        r[2]        = vv[2] + 2.0f * c2[2];
    };
    float vx[3];
    axis(ay, -q.angle[0], v.qz);
    rot(v.qz, mx, vx);
    axis(vx, -q.angle[1], v.qx);
    v.W       = q.width;
    v.H       = q.height;
    v.w       = (int) q.width;
    v.h       = (int) q.height;
    v.shoot   = q.shoot;
    v.nlights = nlights;
    v.cos_thr = cosf(0.02f);
    for (int l = 0; l < nlights; l++) {
        float lv[3] = {lights[3 * l] - q.camfp[0], lights[3 * l + 1] - q.camfp[1], lights[3 * l + 2] - q.camfp[2]};
        float d     = lv[0] * lv[0] + lv[1] * lv[1] + lv[2] * lv[2];
        float sc    = 1.0f / sqrtf(d);
        for (int k = 0; k < 3; k++) {
            v.light[l][k] = lights[3 * l + k];
            v.cl[l][k]    = lv[k] * sc;
        }
    }
}

int qbt_ctx::ensure_frame(int view, int w, int h, unsigned mask)
{
    if (view >= (int) frames.size()) frames.resize(view + 1);
    Frame&       fr  = frames[view];
    const size_t n   = (size_t) w * (size_t) h;
    const size_t esz[FR_N] = {4, 4, 1, 4, 16, 16, 1, 4, 4};
    for (int k = 0; k < FR_N; k++) {
        if (!(mask & (1u << k))) continue;
        int rc = dev_alloc(fr.buf[k], n * esz[k] > 0 ? n * esz[k] : 16);
        if (rc) return rc;
    }
    fr.w    = w;
    fr.h    = h;
    fr.mask = mask & QBT_OUT_ALL;
    return QBT_OK;
}

FrameDev qbt_ctx::frame_dev(int view)
{
    Frame&   fr = frames[view];
    FrameDev f;
    f.hit    = (int*) fr.buf[FR_HIT].p;
    f.leaf   = (int*) fr.buf[FR_LEAF].p;
    f.shadow = (uint8_t*) fr.buf[FR_SHADOW].p;
    f.rgba8  = (uint32_t*) fr.buf[FR_RGBA8].p;
    f.rgbaf  = (float4*) fr.buf[FR_RGBAF].p;
    f.isp    = (float4*) fr.buf[FR_ISP].p;
    f.flags  = (uint8_t*) fr.buf[FR_FLAGS].p;
    f.dyn    = (int*) fr.buf[FR_DYN].p;
    f.exp    = (int*) fr.buf[FR_EXP].p;
    return f;
}

template <bool DUAL, bool GRID, bool COUNT>
static void launch_render_t(qbt_ctx* c, const SceneDev& s, const ViewDev& v, const int* spr, FrameDev f, unsigned mask)
{
    const int    bs    = c->block_threads(s.maxlevel, DUAL);
    const int    th    = bs / 16;
    const int    tiles = ((v.w + 15) / 16) * ((v.h + th - 1) / th);
    const size_t lds   = qbt_stack_bytes(s.maxlevel, DUAL, bs);
    c->prof_mark_begin(QBT_PROF_PRIMARY);
    k_render<DUAL, GRID, COUNT><<<tiles, bs, lds, c->stream>>>(s, v, spr, f, mask, (unsigned long long*) c->counters.p);
    c->prof_mark_end();
}

int qbt_ctx::block_threads(int L, bool dual) const
{
    int bs = tune_block > 0 ? tune_block : 256;
    while (bs > 64 && qbt_stack_bytes(L, dual, bs) > 64 * 1024) bs /= 2;
    return bs;
}

int qbt_ctx::launch_render(const SceneDev& s, bool dual, bool grid, bool count, const ViewDev& v, const int* spr,
                           FrameDev f, unsigned mask)
{
    if (v.w <= 0 || v.h <= 0) return QBT_OK;
    if (dual) {
        if (grid) count ? launch_render_t<true, true, true>(this, s, v, spr, f, mask) : launch_render_t<true, true, false>(this, s, v, spr, f, mask);
        else count ? launch_render_t<true, false, true>(this, s, v, spr, f, mask) : launch_render_t<true, false, false>(this, s, v, spr, f, mask);
    } else {
        if (grid) count ? launch_render_t<false, true, true>(this, s, v, spr, f, mask) : launch_render_t<false, true, false>(this, s, v, spr, f, mask);
        else count ? launch_render_t<false, false, true>(this, s, v, spr, f, mask) : launch_render_t<false, false, false>(this, s, v, spr, f, mask);
    }
    HIPCHK(hipGetLastError());
    return QBT_OK;
}

int qbt_ctx::launch_sprite(const SceneDev& s, bool cmp, bool dual, bool grid, const BatchDev& b, hipStream_t st)
{
    const size_t lds = qbt_stack_bytes(s.maxlevel, dual, 64);
    qbt_dispatch3(cmp, dual, grid, [&](auto C_, auto D_, auto G_) {
        k_sprite<decltype(C_)::value, decltype(D_)::value, decltype(G_)::value><<<b.n, 64, lds, st>>>(
            s, b, (int*) sprite.p);
        return 0;
    });
    HIPCHK(hipGetLastError());
    return QBT_OK;
}

/* The view tables of a batch go to the device through a ring of pinned
 * staging slots; a slot is reused only after its previous copy completed. */
static int upload_batch(qbt_ctx* c, const BatchHost& bh, BatchDev& b)
{
    const int    n     = (int) bh.views.size();
    const size_t o_it  = sizeof(ViewDev) * QBT_MAX_VIEWS;
    const size_t o_px  = o_it + 4 * (QBT_MAX_VIEWS + 1);
    const size_t o_fr  = (o_px + 4 * (QBT_MAX_VIEWS + 1) + 15) & ~(size_t) 15;
    const size_t bytes = o_fr + sizeof(FrameDev) * QBT_MAX_VIEWS;
    int          rc;
    if ((rc = c->dev_alloc(c->bt_tab, bytes))) return rc;
    if (!c->tab_host[0]) {
        for (int k = 0; k < qbt_ctx::kTabSlots; k++) {
            HIPCHK(hipHostMalloc(&c->tab_host[k], bytes, hipHostMallocDefault));
            HIPCHK(hipEventCreateWithFlags(&c->tab_ev[k], hipEventDisableTiming));
            HIPCHK(hipEventRecord(c->tab_ev[k], c->stream));
        }
    }
    const int k = c->tab_slot;
    c->tab_slot = (c->tab_slot + 1) % qbt_ctx::kTabSlots;
    HIPCHK(hipEventSynchronize(c->tab_ev[k]));
    uint8_t* h = (uint8_t*) c->tab_host[k];
    memcpy(h, bh.views.data(), sizeof(ViewDev) * n);
    memcpy(h + o_it, bh.item_off.data(), 4 * (size_t) (n + 1));
    memcpy(h + o_px, bh.pix_off.data(), 4 * (size_t) (n + 1));
    memcpy(h + o_fr, bh.frames.data(), sizeof(FrameDev) * n);
    HIPCHK(hipMemcpyAsync(c->bt_tab.p, h, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->tab_ev[k], c->stream));
    uint8_t* d = (uint8_t*) c->bt_tab.p;
    b.views    = (const ViewDev*) d;
    b.item_off = (const uint32_t*) (d + o_it);
    b.pix_off  = (const uint32_t*) (d + o_px);
    b.frames   = (const FrameDev*) (d + o_fr);
    b.sprite   = (const int*) c->sprite.p;
    b.n        = n;
    return QBT_OK;
}

extern "C" int qbt_render(qbt_ctx* c, const qbt_view* views, int nviews, const float* lights, int nlights,
                          unsigned out_mask)
{
    if (!c || !views || nviews <= 0 || nviews > QBT_MAX_VIEWS) return QBT_EINVAL;
    if (nlights < 0 || nlights > QBT_MAX_LIGHTS_RENDER || (nlights > 0 && !lights)) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    const bool count = (out_mask & QBT_RENDER_COUNT) != 0;
    unsigned   mask  = out_mask & QBT_OUT_ALL;
    if (!count) mask &= ~QBT_OUT_EXP;
    for (int i = 0; i < nviews; i++) {
        const qbt_view& q = views[i];
        if (!(q.width >= 1.0f) || !(q.height >= 1.0f) || q.width > 65536.0f || q.height > 65536.0f) return QBT_EINVAL;
        if ((double) (int) q.width * (double) (int) q.height > 2147483647.0) return QBT_EINVAL;
    }
    const unsigned flags = out_mask & (QBT_RENDER_SURROGATE | QBT_RENDER_COLLECT);
    /* the surrogate and sample collection are stages of the wavefront path */
    const int path = flags ? 1 : c->path;
    if ((flags & QBT_RENDER_SURROGATE) && !qbt_surrogate_ready(c)) return QBT_EINVAL;

    c->last = qbt_stats();
    k_zero_u64<<<1, 64, 0, c->stream>>>((unsigned long long*) c->counters.p, CNT_N);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    uint64_t  prim = 0;
    LightsDev L;
    memset(&L, 0, sizeof(L));
    L.n = nlights;
    if (nlights > 0) memcpy(L.p, lights, sizeof(float) * 3 * (size_t) nlights);
    /* batches: runs of consecutive views with the same basecube and maxlevel */
    for (int i = 0; i < nviews;) {
        int j = i + 1;
        while (j < nviews && memcmp(views[j].basecube, views[i].basecube, sizeof(views[i].basecube)) == 0 &&
               views[j].maxlevel == views[i].maxlevel)
            j++;
        SceneDev s;
        bool     dual, grid, cmp;
        /* the fused kernel walks the direct layout; the wavefront kernels either */
        int rc = c->scene_for(views[i].basecube, views[i].maxlevel, s, dual, grid, cmp, path == 1);
        if (rc) return rc;
        BatchHost bh;
        bh.first       = i;
        uint64_t items = 0, pix = 0;
        for (int k = i; k < j; k++) {
            ViewDev v;
            qbt_make_view(views[k], lights, nlights, v);
            if ((rc = c->ensure_frame(k, v.w, v.h, mask))) return rc;
            bh.item_off.push_back((uint32_t) items);
            bh.pix_off.push_back((uint32_t) pix);
            items += (uint64_t) ((v.w + 7) >> 3) * (uint64_t) ((v.h + 7) >> 3) * 64u;
            pix += (uint64_t) v.w * (uint64_t) v.h;
            if (items > 0xFFFF0000ull) return QBT_ERANGE; /* batch indices are 32-bit */
            bh.views.push_back(v);
            bh.frames.push_back(c->frame_dev(k));
        }
        bh.item_off.push_back((uint32_t) items);
        bh.pix_off.push_back((uint32_t) pix);
        bh.nitems = (uint32_t) items;
        bh.npix   = (uint32_t) pix;
        BatchDev b;
        if ((rc = upload_batch(c, bh, b))) return rc;
        if (path == 0) {
            if (nlights > 0 && (rc = c->launch_sprite(s, cmp, dual, grid, b, c->stream))) return rc;
            for (int k = i; k < j; k++) {
                const int* spr = (const int*) c->sprite.p + (k - i) * QBT_MAX_LIGHTS_RENDER;
                if ((rc = c->launch_render(s, dual, grid, count, bh.views[k - i], spr, c->frame_dev(k), mask)))
                    return rc;
                c->last.launches += 1;
            }
        } else if ((rc = c->render_batch_wavefront(s, cmp, dual, grid, count, bh, b, L, mask, flags))) {
            return rc;
        }
        c->last.compact    = cmp;
        c->last.grid_exact = grid;
        c->last.dual       = dual;
        c->last.launches += 2;
        prim += pix;
        i = j;
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    c->last.primary_rays = prim;
    c->pending_stats     = true;
    c->cnt_src           = nullptr;
    c->stats_has_trace   = false;
    return QBT_OK;
}

extern "C" int qbt_update(qbt_ctx* c, float width, float height, const float position[3], const float angle[3],
                          float lighta, uint8_t quality, int maxlevel, float basesize, int shoot)
{
    if (!c || !position || !angle) return QBT_EINVAL;
    /* octree_glc.c:263-270: the arithmetic of the reference, double where it is double */
    float    light[3] = {420.0f, (float) (200.0f - (double) sinf(lighta) * 20.0),
                      (float) (680.0f - (double) sinf(lighta) * 200.0)};
    qbt_view v;
    memset(&v, 0, sizeof(v));
    v.camfp[0]    = position[0];
    v.camfp[1]    = position[1];
    v.camfp[2]    = position[2];
    v.angle[0]    = angle[0];
    v.angle[1]    = angle[1];
    v.width       = (float) ((double) width / (6.0 - (double) ((float) quality / 2.0)));
    v.height      = (float) ((double) height / (6.0 - (double) ((float) quality / 2.0)));
    v.shoot       = shoot;
    v.basecube[0] = 0.0f;
    v.basecube[1] = basesize;
    v.basecube[2] = basesize;
    v.basecube[3] = basesize;
    v.maxlevel    = maxlevel;
    return qbt_render(c, &v, 1, light, 1, QBT_OUT_DEFAULT);
}

/* qbt_render's fused path into caller-owned device frames, with every
 * buffer beyond the inputs and outputs in the static scratch array: no
 * allocation of any kind. */
extern "C" int qbt_render_noalloc(qbt_ctx* c, const qbt_view* views, int nviews, const float* lights, int nlights,
                                  unsigned out_mask, const qbt_frame* frames)
{
    if (!c || !views || !frames || nviews <= 0) return QBT_EINVAL;
    if (nviews > QBT_MAX_VIEWS) return QBT_ERANGE; /* sprite flags of QBT_MAX_VIEWS views fit the scratch */
    if (nlights < 0 || nlights > QBT_MAX_LIGHTS_RENDER || (nlights > 0 && !lights)) return QBT_EINVAL;
    if (out_mask & (QBT_RENDER_SURROGATE | QBT_RENDER_COLLECT)) return QBT_ENOTSUP; /* wavefront stages only */
    if (!c->scratch) return QBT_EDEVICE;
    const bool count = (out_mask & QBT_RENDER_COUNT) != 0;
    unsigned   mask  = out_mask & QBT_OUT_ALL;
    if (!count) mask &= ~QBT_OUT_EXP;
    for (int i = 0; i < nviews; i++) {
        const qbt_view&  q = views[i];
        const qbt_frame& f = frames[i];
        if (!(q.width >= 1.0f) || !(q.height >= 1.0f) || q.width > 65536.0f || q.height > 65536.0f) return QBT_EINVAL;
        if ((double) (int) q.width * (double) (int) q.height > 2147483647.0) return QBT_EINVAL;
        if (f.w != (int) q.width || f.h != (int) q.height) return QBT_EINVAL;
        const void* need[FR_N] = {f.hit, f.leaf, f.shadow, f.rgba8, f.rgbaf, f.isp, f.flags, f.dyn, f.exp};
        for (int k = 0; k < FR_N; k++)
            if ((mask & (1u << k)) && !need[k]) return QBT_EINVAL; /* a selected output without a buffer */
        if (((uintptr_t) f.rgbaf | (uintptr_t) f.isp) & 15u) return QBT_EINVAL; /* float4 stores */
    }
    HIPCHK(hipSetDevice(c->device));
    uint8_t* const            scr = (uint8_t*) c->scratch; /* this device's copy of `scratch` */
    unsigned long long* const cnt = (unsigned long long*) (scr + kScrCounters);
    unsigned* const           vfl = (unsigned*) (scr + kScrValFlag);
    int* const                spr = (int*) (scr + kScrSprite);
    c->last = qbt_stats();
    k_zero_u64<<<1, 64, 0, c->stream>>>(cnt, CNT_N);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    uint64_t prim = 0;
    for (int i = 0; i < nviews; i++) {
        SceneDev s;
        bool     dual, grid, cmp;
        /* the fused kernel walks the direct layout: no relayout, no allocation */
        int rc = c->scene_for(views[i].basecube, views[i].maxlevel, s, dual, grid, cmp, false, vfl);
        if (rc) return rc;
        ViewDev v;
        qbt_make_view(views[i], lights, nlights, v);
        const qbt_frame& q = frames[i];
        FrameDev f = {q.hit, q.leaf, q.shadow, q.rgba8, (float4*) q.rgbaf, (float4*) q.isp, q.flags, q.dyn, q.exp};
        int* const   sp    = spr + i * QBT_MAX_LIGHTS_RENDER;
        const int    bs    = c->block_threads(s.maxlevel, dual);
        const int    th    = bs / 16;
        const int    tiles = ((v.w + 15) / 16) * ((v.h + th - 1) / th);
        const size_t lds   = qbt_stack_bytes(s.maxlevel, dual, bs);
        const size_t slds  = qbt_stack_bytes(s.maxlevel, dual, 64);
        qbt_dispatch3(dual, grid, count, [&](auto D_, auto G_, auto N_) {
            constexpr bool D = decltype(D_)::value, G = decltype(G_)::value, N = decltype(N_)::value;
            if (nlights > 0) k_sprite_view<D, G><<<1, 64, slds, c->stream>>>(s, v, sp);
            k_render<D, G, N><<<tiles, bs, lds, c->stream>>>(s, v, sp, f, mask, cnt);
            return 0;
        });
        HIPCHK(hipGetLastError());
        c->last.grid_exact = grid;
        c->last.dual       = dual;
        c->last.launches += nlights > 0 ? 2 : 1;
        prim += (uint64_t) v.w * (uint64_t) v.h;
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    c->last.primary_rays = prim;
    c->pending_stats     = true;
    c->stats_has_trace   = false;
    c->cnt_src           = cnt;
    return QBT_OK;
}

extern "C" int qbt_get_stats(qbt_ctx* c, qbt_stats* out)
{
    if (!c || !out) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    if (c->pending_stats) {
        HIPCHK(hipStreamSynchronize(c->stream));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last.kernel_ms = ms;
        c->last.trace_ms  = ms;
        if (c->stats_has_trace) {
            float tms = 0.0f;
            HIPCHK(hipEventElapsedTime(&tms, c->ev0, c->ev2));
            c->last.trace_ms = tms;
        }
        unsigned long long cnt[CNT_N];
        HIPCHK(hipMemcpy(cnt, c->cnt_src ? c->cnt_src : c->counters.p, sizeof(cnt), hipMemcpyDeviceToHost));
        c->last.shadow_rays = cnt[CNT_SHADOW];
        c->last.hits        = cnt[CNT_HITS];
        c->last.expansions  = cnt[CNT_EXP];
        c->last.expansions_primary = cnt[CNT_EXP_PRIM];
        c->last.shadow_deferred    = cnt[CNT_DEFER];
        c->last.predicted   = cnt[CNT_PRED];
        c->aborted          = cnt[CNT_ABORT];
        c->pending_stats    = false;
    }
    *out = c->last;
    return c->aborted ? QBT_EABORT : QBT_OK;
}

extern "C" int qbt_layout_stats(qbt_ctx* c, int tree, uint64_t* relayouts, uint64_t* patches, uint64_t* records)
{
    if (!c || (tree != QBT_BUFFER_STATIC_OCTREE && tree != QBT_BUFFER_DYNAMIC_OCTREE)) return QBT_EINVAL;
    const TreeBuf& tb = c->tree[tree == QBT_BUFFER_STATIC_OCTREE ? 0 : 1];
    if (relayouts) *relayouts = tb.relayouts;
    if (patches) *patches = tb.patches;
    if (records) *records = tb.cmp.ok ? tb.cmp.n : 0;
    return QBT_OK;
}

extern "C" int qbt_synchronize(qbt_ctx* c)
{
    if (!c) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return QBT_OK;
}

extern "C" int qbt_profile_begin(qbt_ctx* c)
{
    if (!c) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->profiling = true;
    c->prof_n    = 0;
    return QBT_OK;
}

extern "C" int qbt_profile_end(qbt_ctx* c, qbt_profile* out)
{
    if (!c || !out) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->profiling = false;
    qbt_profile p{};
    p.min_ms = 0.0;
    for (uint32_t i = 0; i < c->prof_n; i++) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, c->prof_ev[2 * i], c->prof_ev[2 * i + 1]));
        p.total_ms += ms;
        const int k = i < c->prof_kind.size() ? c->prof_kind[i] : 0;
        p.kind_launches[k]++;
        p.kind_ms[k] += ms;
        if (i == 0 || ms < p.min_ms) p.min_ms = ms;
        if (ms > p.max_ms) p.max_ms = ms;
    }
    p.launches = c->prof_n;
    p.avg_ms   = c->prof_n ? p.total_ms / c->prof_n : 0.0;
    *out       = p;
    return QBT_OK;
}

extern "C" int qbt_get_tuning(qbt_ctx* c, qbt_tuning* out)
{
    if (!c || !out) return QBT_EINVAL;
    memset(out, 0, sizeof(*out));
    out->path          = c->path;
    out->block_threads = c->tune_block;
    out->refill        = c->tune_refill;
    out->layout        = c->use_compact ? 0 : 1;
    out->stack         = c->tune_stack;
    out->reorder       = c->tune_reorder;
    out->graph         = c->tune_graph;
    out->strip         = c->tune_strip;
    out->merge         = c->tune_merge;
    return QBT_OK;
}

extern "C" int qbt_set_tuning(qbt_ctx* c, const qbt_tuning* in)
{
    if (!c || !in) return QBT_EINVAL;
    if (in->path < 0 || in->path > 1) return QBT_EINVAL;
    if (in->block_threads != 0 && in->block_threads != 64 && in->block_threads != 128 && in->block_threads != 256)
        return QBT_EINVAL;
    if (in->refill < 0 || in->refill > 64) return QBT_EINVAL;
    if (in->layout < 0 || in->layout > 1) return QBT_EINVAL;
    if (in->stack < 0 || in->stack > 1 || in->reorder < 0 || in->reorder > 1 || in->graph < 0 || in->graph > 1 ||
        in->strip < 0 || in->strip > 1024 || in->merge < 0 || in->merge > 2)
        return QBT_EINVAL;
    c->tune_strip   = in->strip;
    c->tune_merge   = in->merge;
    c->path         = in->path;
    c->tune_stack   = in->stack;
    c->tune_reorder = in->reorder;
    c->tune_graph   = in->graph;
    c->tune_block  = in->block_threads;
    c->tune_refill = in->refill;
    c->use_compact = in->layout == 0;
    return QBT_OK;
}

/* time `reps` renders of the batch under the context's current tuning (one
 * untimed render first: relayout, allocation, graph capture) */
static int time_renders(qbt_ctx* c, const qbt_view* views, int nviews, const float* lights, int nlights,
                        unsigned mask, int reps, double* ms)
{
    int rc = qbt_render(c, views, nviews, lights, nlights, mask);
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->ev2, c->stream));
    HIPCHK(hipEventSynchronize(c->ev2));
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void) hipEventDestroy(a);
        return QBT_EDEVICE;
    }
    (void) hipEventRecord(a, c->stream);
    for (int r = 0; r < reps && !rc; r++) rc = qbt_render(c, views, nviews, lights, nlights, mask);
    (void) hipEventRecord(b, c->stream);
    hipError_t e  = hipEventSynchronize(b);
    float      el = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&el, a, b);
    (void) hipEventDestroy(a);
    (void) hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return QBT_EDEVICE;
    *ms = el / reps;
    c->pending_stats = false; /* the stats events were overwritten by the timing renders */
    return QBT_OK;
}

extern "C" int qbt_autotune(qbt_ctx* c, const qbt_view* views, int nviews, const float* lights, int nlights,
                            unsigned mask, int reps, qbt_tune_result* results, int max_results, int* nresults)
{
    if (!c || !views || nviews <= 0 || reps <= 0 || reps > 100 || (max_results > 0 && !results) || !nresults)
        return QBT_EINVAL;
    mask &= ~(QBT_RENDER_COUNT | QBT_RENDER_COLLECT); /* tune the production render */
    qbt_tuning best;
    int        rc = qbt_get_tuning(c, &best);
    if (rc) return rc;
    best.path          = 1;
    best.block_threads = best.block_threads ? best.block_threads : 256; /* the defaults, made explicit */
    double best_ms     = 0.0;
    int    n       = 0;
    auto   run     = [&](const qbt_tuning& t, double* ms) -> int {
        int e = qbt_set_tuning(c, &t);
        if (!e) e = time_renders(c, views, nviews, lights, nlights, mask, reps, ms);
        if (!e && n < max_results) results[n].tuning = t, results[n].ms = *ms;
        if (!e) n++;
        return e;
    };
    if ((rc = run(best, &best_ms))) return rc;
    /* coordinate descent over the knobs, two sweeps */
    static const int blocks[]  = {64, 128, 256};
    static const int refills[] = {0, 8, 16, 24, 32, 48}; /* 0: the per-launch defaults */
    for (int sweep = 0; sweep < 2; sweep++) {
        bool improved = false;
        for (int knob = 0; knob < 5; knob++) {
            const int* vals = knob == 0 ? blocks : knob == 1 ? refills : nullptr;
            const int  nv   = knob == 0 ? 3 : knob == 1 ? 6 : knob == 4 ? 3 : 2;
            for (int k = 0; k < nv; k++) {
                qbt_tuning t = best;
                int*       f = knob == 0   ? &t.block_threads
                               : knob == 1 ? &t.refill
                               : knob == 2 ? &t.stack
                               : knob == 3 ? &t.reorder
                                           : &t.merge;
                const int  v = vals ? vals[k] : k;
                if (v == *f) continue;
                *f = v;
                double ms = 0.0;
                if ((rc = run(t, &ms))) return rc;
                if (ms < best_ms * 0.99) { /* keep only a clear gain */
                    best     = t;
                    best_ms  = ms;
                    improved = true;
                }
            }
        }
        if (!improved) break;
    }
    *nresults = n;
    return qbt_set_tuning(c, &best);
}

extern "C" int qbt_copy_bandwidth(qbt_ctx* c, size_t bytes, int reps, double* avg_ms, double* gbps)
{
    if (!c || bytes == 0 || bytes % 16 || reps <= 0 || reps > 1000) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    DevBuf a, b;
    int    rc = c->dev_alloc(a, bytes);
    if (!rc) rc = c->dev_alloc(b, bytes);
    if (rc) {
        c->dev_free(a);
        c->dev_free(b);
        return rc;
    }
    const size_t n = bytes / 16;
    k_fill_u32<<<blocks_for(n * 4, 256), 256, 0, c->stream>>>((uint32_t*) a.p, n * 4, 1u);
    k_copy16<<<4096, 256, 0, c->stream>>>((const int4*) a.p, (int4*) b.p, n); /* warm */
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; r++) k_copy16<<<4096, 256, 0, c->stream>>>((const int4*) a.p, (int4*) b.p, n);
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->dev_free(a);
    c->dev_free(b);
    double avg = ms / reps;
    if (avg_ms) *avg_ms = avg;
    if (gbps) *gbps = 2.0 * (double) bytes / (avg * 1e-3) / 1e9;
    return QBT_OK;
}

static int frame_fields(qbt_ctx* c, int view, qbt_frame* out, bool device)
{
    if (!c || !out || view < 0 || view >= (int) c->frames.size()) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    Frame& fr = c->frames[view];
    if (device) {
        FrameDev f = c->frame_dev(view);
        out->w     = fr.w;
        out->h     = fr.h;
        out->hit   = (fr.mask & QBT_OUT_HIT) ? f.hit : nullptr;
        out->leaf  = (fr.mask & QBT_OUT_LEAF) ? f.leaf : nullptr;
        out->shadow = (fr.mask & QBT_OUT_SHADOW) ? f.shadow : nullptr;
        out->rgba8  = (fr.mask & QBT_OUT_RGBA8) ? f.rgba8 : nullptr;
        out->rgbaf  = (fr.mask & QBT_OUT_RGBAF) ? (float*) f.rgbaf : nullptr;
        out->isp    = (fr.mask & QBT_OUT_ISP) ? (float*) f.isp : nullptr;
        out->flags  = (fr.mask & QBT_OUT_FLAGS) ? f.flags : nullptr;
        out->dyn    = (fr.mask & QBT_OUT_DYN) ? f.dyn : nullptr;
        out->exp    = (fr.mask & QBT_OUT_EXP) ? f.exp : nullptr;
        return QBT_OK;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t n = (size_t) fr.w * fr.h;
    struct {
        void*    dst;
        unsigned bit;
        int      k;
        size_t   esz;
    } m[] = {{out->hit, QBT_OUT_HIT, FR_HIT, 4},       {out->leaf, QBT_OUT_LEAF, FR_LEAF, 4},
             {out->shadow, QBT_OUT_SHADOW, FR_SHADOW, 1}, {out->rgba8, QBT_OUT_RGBA8, FR_RGBA8, 4},
             {out->rgbaf, QBT_OUT_RGBAF, FR_RGBAF, 16},   {out->isp, QBT_OUT_ISP, FR_ISP, 16},
             {out->flags, QBT_OUT_FLAGS, FR_FLAGS, 1},    {out->dyn, QBT_OUT_DYN, FR_DYN, 4},
             {out->exp, QBT_OUT_EXP, FR_EXP, 4}};
    out->w = fr.w;
    out->h = fr.h;
    for (auto& e : m) {
        if (!e.dst) continue;
        if (!(fr.mask & e.bit)) return QBT_EINVAL; /* not rendered */
        HIPCHK(hipMemcpy(e.dst, fr.buf[e.k].p, n * e.esz, hipMemcpyDeviceToHost));
    }
    return QBT_OK;
}

extern "C" int qbt_read_frame(qbt_ctx* c, int view, qbt_frame* out) { return frame_fields(c, view, out, false); }
extern "C" int qbt_frame_device(qbt_ctx* c, int view, qbt_frame* out) { return frame_fields(c, view, out, true); }

extern "C" int qbt_trace_rays(qbt_ctx* c, const float* origins, const float* dirs, size_t n, int semantics,
                              const float basecube[4], int maxlevel, int32_t* status, int32_t* model, int32_t* leaf,
                              float* isp, float* tlf)
{
    if (!c || !basecube || (n && (!origins || !dirs)) || semantics < 0 || semantics > 2) return QBT_EINVAL;
    if (n == 0) return QBT_OK;
    if (n > (size_t) 1 << 31) return QBT_ENOTSUP;
    HIPCHK(hipSetDevice(c->device));
    SceneDev s;
    bool     dual, grid, cmp;
    /* the particle semantics addresses reference node indices: direct layout */
    int rc = c->scene_for(basecube, maxlevel, s, dual, grid, cmp, semantics != 2);
    if (rc) return rc;
    if (semantics != 0) dual = false; /* octree_trace_line and the particle copy walk one tree */
    /* ray scratch: org, dir (n x 12 B each), status/model/leaf (n x 4 B each),
     * then isp and tlf (n x 16 B each) at a 16-byte aligned offset */
    const size_t off_isp = ((n * 36 + 15) / 16) * 16;
    if ((rc = c->dev_alloc(c->rays, off_isp + n * 32))) return rc;
    char*   base = (char*) c->rays.p;
    float*  dorg = (float*) base;
    float*  ddir = (float*) (base + n * 12);
    int*    dst  = (int*) (base + n * 24);
    int*    dmod = (int*) (base + n * 28);
    int*    dlf  = (int*) (base + n * 32);
    float4* disp = (float4*) (base + off_isp);
    float4* dtlf = (float4*) (base + off_isp + n * 16);
    HIPCHK(hipMemcpyAsync(dorg, origins, n * 12, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(ddir, dirs, n * 12, hipMemcpyHostToDevice, c->stream));
    k_zero_u64<<<1, 64, 0, c->stream>>>((unsigned long long*) c->counters.p, CNT_N);
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    const int    bs  = c->block_threads(maxlevel, dual);
    const size_t lds = qbt_stack_bytes(maxlevel, dual, bs);
    unsigned     nb  = blocks_for(n, bs);
    unsigned long long* cnt = (unsigned long long*) c->counters.p;
    c->prof_mark_begin(QBT_PROF_PRIMARY);
    /* (DUAL, SEM) in {(0,shader), (1,shader), (0,cpu)}: the CPU twin walks the
     * static tree only; the particle copy: static tree, direct layout */
    if (semantics == 2) {
        auto k = grid ? k_trace_rays<false, false, kSemParticle, true> : k_trace_rays<false, false, kSemParticle, false>;
        k<<<nb, bs, lds, c->stream>>>(s, dorg, ddir, n, dst, dmod, dlf, disp, dtlf, cnt);
    } else {
        qbt_dispatch4(cmp, dual, semantics == 1, grid, [&](auto C_, auto D_, auto P_, auto G_) {
            constexpr bool kD = decltype(D_)::value && !decltype(P_)::value;
            k_trace_rays<decltype(C_)::value, kD, decltype(P_)::value ? kSemCpu : kSemShader, decltype(G_)::value>
                <<<nb, bs, lds, c->stream>>>(s, dorg, ddir, n, dst, dmod, dlf, disp, dtlf, cnt);
            return 0;
        });
    }
    HIPCHK(hipGetLastError());
    c->prof_mark_end();
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    if (status) HIPCHK(hipMemcpyAsync(status, dst, n * 4, hipMemcpyDeviceToHost, c->stream));
    if (model) HIPCHK(hipMemcpyAsync(model, dmod, n * 4, hipMemcpyDeviceToHost, c->stream));
    if (leaf) HIPCHK(hipMemcpyAsync(leaf, dlf, n * 4, hipMemcpyDeviceToHost, c->stream));
    if (isp) HIPCHK(hipMemcpyAsync(isp, disp, n * 16, hipMemcpyDeviceToHost, c->stream));
    if (tlf) HIPCHK(hipMemcpyAsync(tlf, dtlf, n * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->last            = qbt_stats();
    c->last.primary_rays = n;
    c->last.grid_exact   = grid;
    c->last.dual         = dual;
    c->last.launches     = 1;
    c->pending_stats     = true;
    c->cnt_src           = nullptr;
    c->stats_has_trace   = false;
    qbt_stats st;
    return qbt_get_stats(c, &st);
}

/* particle_glc_alloc_in + particle_glc_update + read-back (particle_glc.c:118-167) */
extern "C" int qbt_particle_update(qbt_ctx* c, const float* pos, const float* spd, size_t n, int maxlevel,
                                   float basesize, float* pos_out, float* spd_out)
{
    if (!c || (n && (!pos || !spd || !pos_out || !spd_out))) return QBT_EINVAL;
    if (n == 0) return QBT_OK;
    if (n > (size_t) 1 << 31) return QBT_ENOTSUP;
    HIPCHK(hipSetDevice(c->device));
    const float basecube[4] = {0.0f, basesize, basesize, basesize}; /* particle_glc.c:130 */
    SceneDev    s;
    bool        dual, grid, cmp;
    int         rc = c->scene_for(basecube, maxlevel, s, dual, grid, cmp, false); /* direct layout */
    if (rc) return rc;
    /* particle scratch: pos, spd in place (n x 12 B each) */
    if ((rc = c->dev_alloc(c->rays, n * 24))) return rc;
    float* dpos = (float*) c->rays.p;
    float* dspd = dpos + 3 * n;
    HIPCHK(hipMemcpyAsync(dpos, pos, n * 12, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dspd, spd, n * 12, hipMemcpyHostToDevice, c->stream));
    k_zero_u64<<<1, 64, 0, c->stream>>>((unsigned long long*) c->counters.p, CNT_N);
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    const int    bs  = c->block_threads(maxlevel, false);
    const size_t lds = qbt_stack_bytes(maxlevel, false, bs);
    const unsigned nb = blocks_for(n, bs);
    unsigned long long* cnt = (unsigned long long*) c->counters.p;
    c->prof_mark_begin(QBT_PROF_PRIMARY);
    if (grid) k_particles<true><<<nb, bs, lds, c->stream>>>(s, dpos, dspd, n, dpos, dspd, cnt);
    else k_particles<false><<<nb, bs, lds, c->stream>>>(s, dpos, dspd, n, dpos, dspd, cnt);
    HIPCHK(hipGetLastError());
    c->prof_mark_end();
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipMemcpyAsync(pos_out, dpos, n * 12, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(spd_out, dspd, n * 12, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->last              = qbt_stats();
    c->last.primary_rays = n;
    c->last.grid_exact   = grid;
    c->last.launches     = 1;
    c->pending_stats     = true;
    c->cnt_src           = nullptr;
    c->stats_has_trace   = false;
    qbt_stats st;
    return qbt_get_stats(c, &st);
}
