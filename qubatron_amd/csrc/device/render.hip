/*
 * render.hip - the wavefront render path of qbt_render.
 *
 * The pixel program of octree_fsh.c:399-464 is split into stages so that
 * every traversal kernel runs one kind of ray and keeps its SIMD lanes busy:
 *
 *   k_primary  camera rays, one per pixel, pixels in 8x8-tile order;
 *              writes the hit record (entry point, model indices, leaf, flags)
 *   select     stable compaction (rocprim) of the shaded pixels' work items
 *   k_shadow   one launch per light: the ray from the light to each shaded
 *              pixel's entry point (octree_fsh.c:431-434); writes the per-light
 *              status (lit / occluded / discarded); pixels discarded by an
 *              earlier light are skipped, as the shader stops there
 *   k_shade    per pixel: attributes, the shading formula
 *              (octree_fsh.c:436-449), the light sprite (:452-459), outputs
 *
 * k_primary and k_shadow are persistent: each wave keeps up to 64 rays in
 * flight and refills idle lanes from a work counter with one atomic per
 * refill, so a lane that finishes early starts the next ray instead of
 * idling until the slowest ray of its wave ends. Work items are split into
 * eight contiguous segments, one per XCD (workgroups are dispatched to the
 * eight XCDs round-robin, so block b starts on segment b % 8 and moves to
 * the next segment when its own runs dry): neighbouring pixels, which share
 * most of their node path, stay on one XCD's L2.
 *
 * All traversal arithmetic is trace.hpp's; results are identical to the
 * fused kernel and to the reference algorithm.
 */
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/rocprim.hpp>

#include "pixel.hpp"
#include "qbt.h"
#include "qbt_internal.hpp"
#include "trace.hpp"

#pragma clang fp contract(off)

using namespace qbt;

#define HIPCHK(x) QBT_HIPCHK(x)

namespace {

constexpr int kSeg = 8;

/* per-view device counters (uint32), zeroed before each view: refill
 * counters of the primary and per-light shadow launches, the number of
 * shaded pixels, the per-light exact-trace counts of the surrogate mode */
enum {
    Q_PRIM   = 0,
    Q_SHADOW = kSeg,
    Q_NSEL   = kSeg + kSeg * QBT_MAX_LIGHTS_RENDER,
    Q_FB     = Q_NSEL + 1,
    Q_N      = Q_FB + QBT_MAX_LIGHTS_RENDER
};

struct PrimOut {
    float4*  isp;    /* entry point of the primary leaf, 0 when none */
    int*     ms;     /* static model index, -1 none */
    int*     md;     /* dynamic model index, -1 none */
    int*     leaf;   /* static leaf node, -1 none */
    uint8_t* flags;  /* QBT_PX_LEAF | QBT_PX_SHADED | QBT_PX_DISCARD */
    uint8_t* sel;    /* per work item (tile order): shaded */
    int*     exp;    /* per-pixel expansions (counting renders) */
};

__device__ __forceinline__ uint32_t seg_bound(uint32_t n, int s) { return (uint32_t) (((uint64_t) n * s) / kSeg); }

/* Wave-uniform refill: hands consecutive work items of the wave's current
 * segment to its idle lanes. Returns the item of this lane or UINT32_MAX. */
struct Refill {
    int      seg, left;
    uint32_t n;
    __device__ Refill(uint32_t total) : seg((int) (blockIdx.x % kSeg)), left(kSeg), n(total) {}

    __device__ __forceinline__ uint32_t fetch(unsigned* ctr, uint64_t idle, int lane)
    {
        const uint32_t nidle  = (uint32_t) __popcll(idle);
        const int      leader = __ffsll((unsigned long long) idle) - 1;
        const uint32_t lo = seg_bound(n, seg), hi = seg_bound(n, seg + 1);
        uint32_t       base = 0;
        if (lane == leader) base = atomicAdd(&ctr[seg], nidle);
        base = __shfl(base, leader, 64);
        uint32_t item = UINT32_MAX;
        if ((idle >> lane) & 1ull) {
            uint32_t rank = (uint32_t) __popcll(idle & ((1ull << lane) - 1ull));
            uint64_t it   = (uint64_t) lo + base + rank;
            if (it < hi) item = (uint32_t) it;
        }
        if ((uint64_t) lo + base + nidle >= hi) { /* this segment is done: move on */
            seg = (seg + 1) % kSeg;
            left--;
        }
        return item;
    }
};

template <bool CMP, bool DUAL, bool GRID, bool COUNT>
__global__ void __launch_bounds__(256) k_primary(SceneDev s, ViewDev v, PrimOut o, unsigned* __restrict__ q,
                                                 unsigned refill, unsigned long long* __restrict__ counters)
{
    extern __shared__ uint32_t lds[];
    const Stack    st      = make_stack(lds, s.maxlevel, DUAL);
    const int      lane    = threadIdx.x & 63;
    const int      tiles_x = (v.w + 7) >> 3, tiles_y = (v.h + 7) >> 3;
    Refill         rf((uint32_t) tiles_x * (uint32_t) tiles_y * 64u);
    bool           active = false;
    uint32_t       item = 0, pix = 0;
    Ray            r = {v.cam[0], v.cam[1], v.cam[2], 0.f, 0.f, 0.f};
    TraceState     t;
    unsigned       n_exp = 0, n_abort = 0, n_hit = 0;

    auto finish = [&](int stt) {
        float4  isp = make_float4(0.f, 0.f, 0.f, 0.f);
        int     ms = -1, md = -1, leaf = -1;
        uint8_t f = 0, sel = 0;
        if (stt == ST_DISCARD) f = QBT_PX_DISCARD;
        if (stt == ST_ABORT) n_abort++;
        if (stt == ST_HIT) {
            const int L = s.maxlevel;
            ms          = (t.ns == 0 && L > 0) ? 0 : s.stat.model[t.ns];
            md          = 0;
            if (DUAL) md = (t.nd == 0 && L > 0) ? 0 : s.dyn.model[t.nd];
            leaf = CMP ? s.stat.ref[t.ns] : t.ns; /* reference node index */
            isp  = t.entry;
            f    = QBT_PX_LEAF;
            if (isp.w > 0.0f) {
                f |= QBT_PX_SHADED;
                sel = 1;
                n_hit++;
            }
        }
        o.isp[pix]   = isp;
        o.ms[pix]    = ms;
        o.md[pix]    = md;
        o.leaf[pix]  = leaf;
        o.flags[pix] = f;
        o.sel[item]  = sel;
        if (COUNT) {
            o.exp[pix] = (int) t.exps;
            n_exp += t.exps;
        }
    };

    for (;;) {
        const uint64_t idle = __ballot(!active);
        if (rf.left > 0 && (__popcll(idle) >= (int) refill || idle == ~0ull)) {
            uint32_t it = rf.fetch(&q[Q_PRIM], idle, lane);
            if (it != UINT32_MAX) {
                int px, py;
                item = it;
                if (item_pixel(it, tiles_x, v.w, v.h, px, py)) {
                    pix = (uint32_t) py * (uint32_t) v.w + (uint32_t) px;
                    float csv[3];
                    pixel_dir(v, px, py, csv);
                    r.dx    = csv[0];
                    r.dy    = csv[1];
                    r.dz    = csv[2];
                    int stt = trace_begin<false>(s, r, t);
                    if (stt == ST_CONT) active = true;
                    else finish(stt);
                } else {
                    o.sel[it] = 0;
                }
            }
        }
        if (__ballot(active) == 0ull) {
            if (rf.left <= 0) break;
            continue;
        }
        if (active) {
            int stt = trace_step<CMP, DUAL, false, GRID>(s, r, st, t);
            if (stt != ST_CONT) {
                finish(stt);
                active = false;
            }
        }
    }
    n_hit   = wave_sum(n_hit);
    n_abort = wave_sum(n_abort);
    if (COUNT) n_exp = wave_sum(n_exp);
    if (lane == 0) {
        if (n_hit) atomicAdd(&counters[CNT_HITS], (unsigned long long) n_hit);
        if (n_abort) atomicAdd(&counters[CNT_ABORT], (unsigned long long) n_abort);
        if (COUNT && n_exp) atomicAdd(&counters[CNT_EXP], (unsigned long long) n_exp);
        if (COUNT && n_exp) atomicAdd(&counters[CNT_EXP_PRIM], (unsigned long long) n_exp);
    }
}

template <bool CMP, bool DUAL, bool GRID, bool COUNT>
__global__ void __launch_bounds__(256) k_shadow(SceneDev s, ViewDev v, int l, const uint32_t* __restrict__ sel_items,
                                                const unsigned* __restrict__ nsel,
                                                const float4* __restrict__ isp, uint8_t* __restrict__ lstat,
                                                uint8_t* __restrict__ disc, unsigned* __restrict__ q,
                                                unsigned refill, int* __restrict__ exp,
                                                unsigned long long* __restrict__ counters)
{
    extern __shared__ uint32_t lds[];
    const Stack st      = make_stack(lds, s.maxlevel, DUAL);
    const int   lane    = threadIdx.x & 63;
    const int   tiles_x = (v.w + 7) >> 3;
    Refill      rf(*nsel);
    bool        active = false;
    uint32_t    pix = 0;
    const float Lx = v.light[l][0], Ly = v.light[l][1], Lz = v.light[l][2];
    Ray         r = {Lx, Ly, Lz, 0.f, 0.f, 0.f};
    float4      ip = make_float4(0.f, 0.f, 0.f, 0.f);
    TraceState  t;
    unsigned    n_exp = 0, n_abort = 0, n_ray = 0;

    auto finish = [&](int stt) {
        uint8_t ls;
        if (stt == ST_DISCARD) {
            ls        = LS_DISCARD;
            disc[pix] = 1;
        } else {
            if (stt == ST_ABORT) n_abort++;
            float4 li  = stt == ST_HIT ? t.entry : make_float4(0.f, 0.f, 0.f, 0.f);
            float  ix  = li.x - ip.x, iy = li.y - ip.y, iz = li.z - ip.z;
            float  sqr = ix * ix + iy * iy + iz * iz;
            ls         = (15.0f < sqr) ? LS_OCCLUDED : LS_LIT; /* step(sqr, 15.0) */
        }
        lstat[pix] = ls;
        if (COUNT) {
            atomicAdd(&exp[pix], (int) t.exps);
            n_exp += t.exps;
        }
    };

    for (;;) {
        const uint64_t idle = __ballot(!active);
        if (rf.left > 0 && (__popcll(idle) >= (int) refill || idle == ~0ull)) {
            uint32_t k = rf.fetch(&q[Q_SHADOW + kSeg * l], idle, lane);
            if (k != UINT32_MAX) {
                int px, py;
                item_pixel(sel_items[k], tiles_x, v.w, v.h, px, py);
                pix = (uint32_t) py * (uint32_t) v.w + (uint32_t) px;
                if (l > 0 && disc[pix]) {
                    lstat[pix] = LS_SKIPPED; /* an earlier light discarded the fragment */
                } else {
                    ip      = isp[pix];
                    r.dx    = ip.x - Lx;
                    r.dy    = ip.y - Ly;
                    r.dz    = ip.z - Lz;
                    n_ray++;
                    int stt = trace_begin<false>(s, r, t);
                    if (stt == ST_CONT) active = true;
                    else finish(stt);
                }
            }
        }
        if (__ballot(active) == 0ull) {
            if (rf.left <= 0) break;
            continue;
        }
        if (active) {
            int stt = trace_step<CMP, DUAL, false, GRID>(s, r, st, t);
            if (stt != ST_CONT) {
                finish(stt);
                active = false;
            }
        }
    }
    n_ray   = wave_sum(n_ray);
    n_abort = wave_sum(n_abort);
    if (COUNT) n_exp = wave_sum(n_exp);
    if (lane == 0) {
        if (n_ray) atomicAdd(&counters[CNT_SHADOW], (unsigned long long) n_ray);
        if (n_abort) atomicAdd(&counters[CNT_ABORT], (unsigned long long) n_abort);
        if (COUNT && n_exp) atomicAdd(&counters[CNT_EXP], (unsigned long long) n_exp);
    }
}

/* per pixel: attributes, shading, sprite, outputs (octree_fsh.c:421-463) */
template <bool DUAL>
__global__ void __launch_bounds__(256) k_shade(SceneDev s, ViewDev v, const int* __restrict__ sprite, PrimOut in,
                                               const uint8_t* __restrict__ lstat, size_t lstride, FrameDev f,
                                               unsigned mask)
{
    const size_t npix = (size_t) v.w * v.h;
    const size_t o    = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= npix) return;
    const int px = (int) (o % (size_t) v.w), py = (int) (o / (size_t) v.w);
    uint32_t  flg   = in.flags[o];
    uint32_t  sbits = 0;
    float     col[4] = {0.f, 0.f, 0.f, 0.f};
    float     nrm[4] = {0.f, 0.f, 0.f, 0.f};
    const int ms = in.ms[o], md = in.md[o];
    float4    isp = in.isp[o];
    if (!(flg & QBT_PX_DISCARD)) {
        float csv[3];
        pixel_dir(v, px, py, csv);
        if (flg & QBT_PX_LEAF) {
            fetch_attr(s.astat, ms, col, nrm);
            if (DUAL && md > 0) fetch_attr(s.adyn, md, col, nrm);
        }
        if (flg & QBT_PX_SHADED) {
            float nn[3], mc[3] = {-csv[0], -csv[1], -csv[2]}, nmc[3];
            normalize3(nrm, nn);
            normalize3(mc, nmc);
            float camv = gmax(dot3(nmc, nn), 0.0f);
            float fac  = 0.1f + 0.2f * camv;
            for (int l = 0; l < v.nlights; l++) {
                uint8_t ls = lstat[l * lstride + o];
                if (ls == LS_DISCARD) {
                    flg |= QBT_PX_DISCARD;
                    break;
                }
                if (ls >= LS_PRED_OCC) flg |= QBT_PX_PREDICTED;
                float bit = (ls == LS_LIT || ls == LS_PRED_LIT) ? 1.0f : 0.0f;
                if (bit != 0.0f) sbits |= 1u << l;
                /* normalize(-lghtv) with lghtv = isp - light (octree_fsh.c:431, 436) */
                float lv[3] = {isp.x - v.light[l][0], isp.y - v.light[l][1], isp.z - v.light[l][2]};
                float ml[3] = {-lv[0], -lv[1], -lv[2]};
                float nml[3];
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
    if (mask & QBT_OUT_HIT) f.hit[o] = ms;
    if (mask & QBT_OUT_LEAF) f.leaf[o] = in.leaf[o];
    if (mask & QBT_OUT_DYN) f.dyn[o] = md;
    if (mask & QBT_OUT_SHADOW) f.shadow[o] = (uint8_t) sbits;
    if (mask & QBT_OUT_FLAGS) f.flags[o] = (uint8_t) flg;
    if (mask & QBT_OUT_RGBA8) f.rgba8[o] = pack_rgba8(col[0], col[1], col[2], col[3]);
    if (mask & QBT_OUT_RGBAF) f.rgbaf[o] = make_float4(col[0], col[1], col[2], col[3]);
    if (mask & QBT_OUT_ISP) f.isp[o] = isp;
}

__global__ void k_zero_u32(unsigned* p, int n)
{
    int i = threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_zero_u8(uint8_t* p, size_t n)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

inline unsigned nblk(size_t n, unsigned bs = 256) { return (unsigned) ((n + bs - 1) / bs); }

template <class K>
int persistent_grid(qbt_ctx* c, K kernel, int bs, size_t lds)
{
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, bs, lds) != hipSuccess || per <= 0) {
        (void) hipGetLastError();
        per = 1;
    }
    return per * c->num_cus;
}

} // namespace

/* internal per-view buffers, sized for the largest frame so far */
int qbt_ctx::ensure_wave_bufs(size_t npix, int nlights)
{
    int rc;
    if ((rc = dev_alloc(wb_isp, npix * 16)) || (rc = dev_alloc(wb_ms, npix * 4)) || (rc = dev_alloc(wb_md, npix * 4)) ||
        (rc = dev_alloc(wb_leaf, npix * 4)) || (rc = dev_alloc(wb_flags, npix)) || (rc = dev_alloc(wb_disc, npix)) ||
        (rc = dev_alloc(wb_lstat, npix * (size_t) (nlights > 0 ? nlights : 1))) || (rc = dev_alloc(wb_exp, npix * 4)))
        return rc;
    return QBT_OK;
}

template <bool CMP, bool DUAL, bool GRID, bool COUNT>
static int render_view_t(qbt_ctx* c, const SceneDev& s, const ViewDev& v, const int* spr, FrameDev f, unsigned mask,
                         unsigned flags)
{
    hipStream_t  st      = c->stream;
    const size_t npix    = (size_t) v.w * v.h;
    const int    tiles_x = (v.w + 7) >> 3, tiles_y = (v.h + 7) >> 3;
    const size_t items   = (size_t) tiles_x * tiles_y * 64;
    int          rc;
    if ((rc = c->ensure_wave_bufs(npix, v.nlights))) return rc;
    const bool surrogate = (flags & QBT_RENDER_SURROGATE) && v.nlights > 0;
    if ((rc = c->dev_alloc(c->wb_sel, items)) || (rc = c->dev_alloc(c->wb_items, items * 4)) ||
        (rc = c->dev_alloc(c->wb_q, sizeof(unsigned) * Q_N)))
        return rc;
    if (surrogate && (rc = c->dev_alloc(c->wb_fb, items * 4 * (size_t) v.nlights))) return rc;
    size_t tmp = 0;
    HIPCHK(rocprim::select(nullptr, tmp, rocprim::counting_iterator<uint32_t>(0u), (const uint8_t*) c->wb_sel.p,
                           (uint32_t*) c->wb_items.p, (unsigned*) c->wb_q.p + Q_NSEL, items, st));
    if ((rc = c->dev_alloc(c->wb_tmp, tmp))) return rc;

    PrimOut po;
    po.isp   = (float4*) c->wb_isp.p;
    po.ms    = (int*) c->wb_ms.p;
    po.md    = (int*) c->wb_md.p;
    po.leaf  = (int*) c->wb_leaf.p;
    po.flags = (uint8_t*) c->wb_flags.p;
    po.sel   = (uint8_t*) c->wb_sel.p;
    po.exp   = (COUNT && (mask & QBT_OUT_EXP)) ? f.exp : (int*) c->wb_exp.p;
    unsigned*           q   = (unsigned*) c->wb_q.p;
    unsigned long long* cnt = (unsigned long long*) c->counters.p;

    k_zero_u32<<<1, 128, 0, st>>>(q, Q_N);
    k_zero_u8<<<nblk(npix), 256, 0, st>>>((uint8_t*) c->wb_disc.p, npix);
    HIPCHK(hipGetLastError());

    const int    bs  = c->block_threads(s.maxlevel, DUAL);
    const size_t lds = qbt_stack_bytes(s.maxlevel, DUAL, bs);
    const unsigned refill = (unsigned) (c->tune_refill > 0 && c->tune_refill <= 64 ? c->tune_refill : 32);
    {
        auto kp = k_primary<CMP, DUAL, GRID, COUNT>;
        int  g  = persistent_grid(c, kp, bs, lds);
        c->prof_mark_begin(QBT_PROF_PRIMARY);
        kp<<<g, bs, lds, st>>>(s, v, po, q, refill, cnt);
        c->prof_mark_end();
        HIPCHK(hipGetLastError());
    }
    size_t tb = tmp;
    HIPCHK(rocprim::select(c->wb_tmp.p, tb, rocprim::counting_iterator<uint32_t>(0u), (const uint8_t*) c->wb_sel.p,
                           (uint32_t*) c->wb_items.p, q + Q_NSEL, items, st));
    if (surrogate && (rc = qbt_surrogate_decide(c, s, DUAL, v, (const uint32_t*) c->wb_items.p, q, Q_NSEL, npix,
                                                (uint32_t*) c->wb_fb.p, items, q + Q_FB)))
        return rc;
    for (int l = 0; l < v.nlights; l++) {
        auto ks = k_shadow<CMP, DUAL, GRID, COUNT>;
        int  g  = persistent_grid(c, ks, bs, lds);
        /* exact mode: every shaded pixel; surrogate mode: this light's unsure ones */
        const uint32_t* li = surrogate ? (const uint32_t*) c->wb_fb.p + (size_t) l * items : (const uint32_t*) c->wb_items.p;
        const unsigned* ln = surrogate ? q + Q_FB + l : q + Q_NSEL;
        c->prof_mark_begin(QBT_PROF_SHADOW);
        ks<<<g, bs, lds, st>>>(s, v, l, li, ln, (const float4*) c->wb_isp.p,
                               (uint8_t*) c->wb_lstat.p + (size_t) l * npix, (uint8_t*) c->wb_disc.p, q, refill,
                               po.exp, cnt);
        c->prof_mark_end();
        HIPCHK(hipGetLastError());
    }
    if ((flags & QBT_RENDER_COLLECT) && !surrogate && v.nlights > 0 &&
        (rc = qbt_collect_view(c, s, DUAL, v, (const uint32_t*) c->wb_items.p, q, Q_NSEL, npix)))
        return rc;
    k_shade<DUAL><<<nblk(npix), 256, 0, st>>>(s, v, spr, po, (const uint8_t*) c->wb_lstat.p, npix, f, mask);
    HIPCHK(hipGetLastError());
    c->last.launches += 3 + v.nlights + (surrogate ? 1 : 0);
    return QBT_OK;
}

int qbt_ctx::render_view_wavefront(const SceneDev& s, bool cmp, bool dual, bool grid, bool count, const ViewDev& v,
                                   const int* spr, FrameDev f, unsigned mask, unsigned flags)
{
    if (v.w <= 0 || v.h <= 0) return QBT_OK;
    return qbt_dispatch4(cmp, dual, grid, count, [&](auto C_, auto D_, auto G_, auto N_) {
        return render_view_t<decltype(C_)::value, decltype(D_)::value, decltype(G_)::value, decltype(N_)::value>(
            this, s, v, spr, f, mask, flags);
    });
}
