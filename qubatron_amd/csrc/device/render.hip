/*
 * render.hip - the wavefront render path of qbt_render, one batch of views
 * (same scene, basecube, maxlevel and lights) per set of launches.
 *
 * The pixel program of octree_fsh.c:399-464 is split into stages so that
 * every traversal kernel runs one kind of ray and keeps its SIMD lanes busy:
 *
 *   k_primary  first the sprite traces (:452-459; one per view and light,
 *              the same for every pixel of a view), then the camera rays of
 *              every view of the batch, one per pixel, pixels in 8x8-tile
 *              order, views in order; writes the hit record (entry point,
 *              model indices, leaf, flags) and, per work item, the batch
 *              pixel index when the pixel is shaded
 *   k_sel_*    list of the shaded pixels in item order: per-block counts, a
 *              scan of the counts, scatter (no same-address atomics)
 *   k_oct_*    per light, the list partitioned by the direction octant of
 *              the shadow ray, list order kept within an octant (tuning
 *              reorder = 1)
 *   (k_mlp_shadow, surrogate.hip: QBT_RENDER_SURROGATE decisions)
 *   k_shadow   one launch for all lights: the ray from each light to each
 *              shaded pixel's entry point (:431-434), light after light in
 *              one work list; writes the per-light status; pixels that an
 *              earlier light discarded are skipped (the shader stops there)
 *              once that discard is visible
 *   k_shade    per pixel of the batch: attributes, shading (:436-449),
 *              sprite (:452-459), outputs into each view's frame
 *
 * With tuning merge (exact renders without reordering), k_frame replaces
 * k_primary, the list and k_shadow: one persistent launch whose waves take
 * camera rays first and then the shadow rays of finished pixels, handed
 * over through per-(light, work item) slots (see k_frame).
 *
 * k_primary and k_shadow are persistent: each wave keeps up to 64 rays in
 * flight and refills idle lanes from a work counter with one atomic per
 * refill, so a lane that finishes early starts the next ray instead of
 * idling until the slowest ray of its wave ends. Work is split into eight
 * contiguous segments, one per XCD (workgroups are dispatched to the eight
 * XCDs round-robin, so block b starts on segment b % 8 and moves on when its
 * own runs dry): neighbouring pixels, which share most of their node path,
 * stay on one XCD's L2. The segment counters sit on separate 256-byte lines,
 * since device-scope atomics on one line queue behind each other.
 *
 * The whole sequence contains kernels only (no memset, copy or
 * allocation), so it is captured once into a hipGraph and replayed for later
 * batches of the same shape; the view tables it reads are copied in before
 * each replay.
 *
 * All traversal arithmetic is trace.hpp's; results are identical to the
 * fused kernel and to the reference algorithm, whatever the batch
 * composition, ray order or stack placement.
 */
#include <hip/hip_runtime.h>

#include <cstring>

#include "pixel.hpp"
#include "qbt.h"
#include "qbt_internal.hpp"
#include "trace.hpp"

#pragma clang fp contract(off)

using namespace qbt;

#define HIPCHK(x) QBT_HIPCHK(x)

namespace {

constexpr int kSeg     = 8;  /* work segments per launch (block b starts in segment b % kSeg: one XCD each) */
constexpr int kCtrPad  = 64; /* uint32 between refill counters: each on its own 256-byte line */
constexpr int kNL      = QBT_MAX_LIGHTS_RENDER;

/* batch counters (uint32), zeroed before each batch: the number of shaded
 * pixels, the per-light exact-trace counts of the surrogate mode, the
 * per-(light, octant) counts and cursors of the reordering, then the refill
 * counters of the primary and shadow launches, kCtrPad apart
 * (every refill is a device-scope atomic on one of them; counters that
 * share a line would queue behind each other) */
enum {
    Q_NSEL   = 0,
    Q_FB     = Q_NSEL + 1,
    Q_OCT    = Q_FB + kNL,
    Q_CUR    = Q_OCT + 8 * kNL,
    Q_PRIM   = ((Q_CUR + 8 * kNL + kCtrPad - 1) / kCtrPad) * kCtrPad,
    Q_SHADOW = Q_PRIM + kSeg * kCtrPad,
    Q_N      = Q_SHADOW + kSeg * kCtrPad
};

struct PrimOut {
    float4*   isp;   /* entry point of the primary leaf, 0 when none */
    int*      ms;    /* static model index, -1 none */
    int*      md;    /* dynamic model index, -1 none */
    int*      leaf;  /* static leaf node (reference index), -1 none */
    uint8_t*  flags; /* QBT_PX_LEAF | QBT_PX_SHADED | QBT_PX_DISCARD */
    uint32_t* cand;  /* per work item: batch pixel index if shaded, else ~0 */
    int*      exp;   /* per pixel expansions (counting renders) */
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
        if (lane == leader) base = atomicAdd(&ctr[seg * kCtrPad], nidle);
        /* readlane of a uniform lane: a scalar, so seg and left stay in
         * scalar registers */
        base          = __builtin_amdgcn_readlane(base, leader);
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

/* Lanes with a ray step until it ends; the wave leaves the loop together
 * once fewer than `keep` lanes are left (wave-uniform), so that the refill
 * test, the ballots and the merge of the ray state run once per refill,
 * not once per step. A lane whose ray ends keeps its status in `fin` and
 * its state as it is: the caller writes the results of all such lanes
 * together at the next refill point (a lane that ends would otherwise run
 * the output code with one or two lanes of its wave, in most steps). */
__device__ __forceinline__ int keep_active(const Refill& rf, unsigned refill)
{
    return rf.left > 0 ? 65 - (int) refill : 1; /* refill when idle lanes >= refill */
}

template <bool CMP, bool DUAL, bool GRID>
__device__ __forceinline__ void step_run(const SceneDev& s, const Ray& r, const Stack& st, TraceState& t, bool& active,
                                         int keep, int& fin)
{
    for (;;) {
        const int stt = trace_step<CMP, DUAL, false, GRID>(s, r, st, t);
        if (stt != ST_CONT) {
            fin    = stt;
            active = false;
            break;
        }
        if (__popcll(__ballot(1)) < keep) break;
    }
}

/* the calling lane's traversal stack: LDS, or global memory (GSTK) */
template <bool DUAL, bool GSTK>
__device__ __forceinline__ Stack lane_stack(uint32_t* lds, uint32_t* gstack, int L)
{
    return GSTK ? make_stack_global(gstack, L, DUAL) : make_stack(lds, L, DUAL);
}

/* Camera rays of the batch. Work items 0 .. nspr-1 are the sprite traces
 * (view i / nl, light i % nl: the ray from the camera to the light,
 * octree_fsh.c:452-459, one per view and light; result flags to
 * b.sprite), then nitems pixel items. */
template <bool CMP, bool DUAL, bool GRID, bool COUNT, bool GSTK>
__global__ void __launch_bounds__(256) k_primary(SceneDev s, BatchDev b, PrimOut o, uint32_t nitems, uint32_t nspr,
                                                 int nl, int strip, int* __restrict__ sprite, unsigned* __restrict__ q,
                                                 unsigned refill, uint32_t* gstack,
                                                 unsigned long long* __restrict__ counters)
{
    extern __shared__ uint32_t lds[];
    const Stack st   = lane_stack<DUAL, GSTK>(lds, gstack, s.maxlevel);
    const int   lane = threadIdx.x & 63;
    Refill      rf(nspr + nitems);
    bool        active = false;
    uint32_t    item = 0, gp = 0;
    int         spr  = -1; /* sprite slot of a sprite lane */
    Ray         r = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    TraceState  t;
    unsigned    n_exp = 0, n_abort = 0, n_hit = 0;

    auto finish = [&](int stt) {
        if (spr >= 0) { /* white if the light is in front of the first surface (resv / lghtv.x > 1) */
            const float ix = stt == ST_HIT ? t.entry.x : 0.0f;
            int         f  = 0;
            if (stt == ST_DISCARD) f |= 1;
            if ((ix - r.px) / r.dx > 1.0f) f |= 2;
            if (stt == ST_ABORT) f |= 4;
            sprite[spr] = f;
            return;
        }
        float4   isp = make_float4(0.f, 0.f, 0.f, 0.f);
        int      ms = -1, md = -1, leaf = -1;
        uint8_t  f    = 0;
        uint32_t cand = UINT32_MAX;
        if (stt == ST_DISCARD) f = QBT_PX_DISCARD;
        /* counters by value: increments under different branches would be
         * merged into a select of their addresses (scratch memory) */
        n_abort += stt == ST_ABORT ? 1u : 0u;
        n_hit += (stt == ST_HIT && t.entry.w > 0.0f) ? 1u : 0u;
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
                cand = gp;
            }
        }
        o.isp[gp]    = isp;
        o.ms[gp]     = ms;
        o.md[gp]     = md;
        o.leaf[gp]   = leaf;
        o.flags[gp]  = f;
        o.cand[item] = cand;
        if (COUNT) {
            o.exp[gp] = (int) t.exps;
            n_exp += t.exps;
        }
    };

    int fin = ST_CONT; /* status of a ray that ended and awaits its outputs */
    for (;;) {
        if (fin != ST_CONT) { /* rays that ended since the last refill point, together */
            finish(fin);
            fin = ST_CONT;
        }
        const uint64_t idle = __ballot(!active);
        if (rf.left > 0 && (__popcll(idle) >= (int) refill || idle == ~0ull)) {
            const uint32_t fetched = rf.fetch(&q[Q_PRIM], idle, lane);
            if (fetched != UINT32_MAX && fetched < nspr) {
                const int      v = (int) (fetched / (uint32_t) nl), l = (int) (fetched % (uint32_t) nl);
                const ViewDev& V = b.views[v];
                spr              = v * kNL + l;
                r.px             = V.cam[0];
                r.py             = V.cam[1];
                r.pz             = V.cam[2];
                r.dx             = V.light[l][0] - V.cam[0];
                r.dy             = V.light[l][1] - V.cam[1];
                r.dz             = V.light[l][2] - V.cam[2];
                int stt          = trace_begin<false>(s, r, t);
                if (stt == ST_CONT) active = true;
                else finish(stt);
            } else if (fetched != UINT32_MAX) {
                const uint32_t it = fetched - nspr;
                const int      v  = batch_view(b.item_off, b.n, it);
                const ViewDev& V  = b.views[v];
                int            px, py;
                item = it;
                spr  = -1;
                if (item_pixel(it - b.item_off[v], V.w, V.h, strip, px, py)) {
                    gp = b.pix_off[v] + (uint32_t) py * (uint32_t) V.w + (uint32_t) px;
                    float csv[3];
                    pixel_dir(V, px, py, csv);
                    r.px    = V.cam[0];
                    r.py    = V.cam[1];
                    r.pz    = V.cam[2];
                    r.dx    = csv[0];
                    r.dy    = csv[1];
                    r.dz    = csv[2];
                    int stt = trace_begin<false>(s, r, t);
                    if (stt == ST_CONT) active = true;
                    else finish(stt);
                } else {
                    o.cand[it] = UINT32_MAX; /* padding item of an edge tile */
                }
            }
        }
        if (__ballot(active) == 0ull) {
            if (rf.left <= 0) break;
            continue;
        }
        if (active) step_run<CMP, DUAL, GRID>(s, r, st, t, active, keep_active(rf, refill), fin);
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

/* ---- stream compaction: per-block counts, a scan of the counts, scatter ---- */

/* the shaded-pixel list: work items whose candidate is a pixel index */
struct PredCand {
    const uint32_t* cand;
    uint32_t        n;
    __device__ __forceinline__ bool operator()(uint32_t i, int, uint32_t& v) const
    {
        if (i >= n) return false;
        v = cand[i];
        return v != UINT32_MAX;
    }
};

/* the surrogate's undecided queries of light `row` */
struct PredUnsure {
    const uint32_t* lists;
    size_t          lstride;
    const unsigned* nsel;
    const uint8_t*  lstat;
    size_t          npix;
    __device__ __forceinline__ bool operator()(uint32_t i, int row, uint32_t& v) const
    {
        if (i >= *nsel) return false;
        v = lists[(size_t) row * lstride + i];
        return lstat[(size_t) row * npix + v] == LS_UNSURE;
    }
};

template <class P>
__global__ void __launch_bounds__(256) k_sel_count(P p, uint32_t* __restrict__ bcnt, uint32_t nb)
{
    __shared__ uint32_t ws[4];
    const int row = blockIdx.y;
    uint32_t  c   = 0;
    for (int k = 0; k < kSelItems / 256; k++) {
        uint32_t v;
        c += p(blockIdx.x * kSelItems + k * 256 + threadIdx.x, row, v) ? 1u : 0u;
    }
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) bcnt[(size_t) row * nb + blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

/* in place: exclusive scan of the nb counts of row blockIdx.x; the row's total
 * to total[row * tstride] */
__global__ void __launch_bounds__(1024) k_sel_scan(uint32_t* __restrict__ bcnt, uint32_t nb,
                                                   unsigned* __restrict__ total, int tstride)
{
    __shared__ uint32_t ws[16];
    __shared__ uint32_t carry;
    uint32_t*           b    = bcnt + (size_t) blockIdx.x * nb;
    const int           lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? b[i] : 0u;
        uint32_t       x = v; /* inclusive scan within the wave */
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) ws[w] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t a = carry;
            for (int k = 0; k < 16; k++) {
                const uint32_t t = ws[k];
                ws[k]            = a;
                a += t;
            }
            carry = a;
        }
        __syncthreads();
        if (i < nb) b[i] = ws[w] + x - v;
        __syncthreads();
    }
    if (threadIdx.x == 0) total[(size_t) blockIdx.x * tstride] = carry;
}

/* values of the kept indices, in index order, at out + row * ostride */
template <class P>
__global__ void __launch_bounds__(256) k_sel_scatter(P p, const uint32_t* __restrict__ boff, uint32_t nb,
                                                     uint32_t* __restrict__ out, size_t ostride)
{
    __shared__ uint32_t ws[4];
    const int           row  = blockIdx.y;
    const int           lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t            base = boff[(size_t) row * nb + blockIdx.x];
    for (int k = 0; k < kSelItems / 256; k++) {
        uint32_t       v    = 0;
        const bool     keep = p(blockIdx.x * kSelItems + k * 256 + threadIdx.x, row, v);
        const uint64_t m    = __ballot(keep);
        if (lane == 0) ws[w] = (uint32_t) __popcll(m);
        __syncthreads();
        uint32_t pre = 0;
        for (int j = 0; j < w; j++) pre += ws[j];
        const uint32_t tot = ws[0] + ws[1] + ws[2] + ws[3];
        if (keep) out[(size_t) row * ostride + base + pre + (uint32_t) __popcll(m & ((1ull << lane) - 1ull))] = v;
        base += tot;
        __syncthreads();
    }
}

template <class P>
void compact(hipStream_t st, P p, uint32_t n, int rows, uint32_t* bcnt, unsigned* total, int tstride, uint32_t* out,
             size_t ostride)
{
    const uint32_t nb = (n + kSelItems - 1) / kSelItems + 1;
    k_sel_count<P><<<dim3(nb, rows), 256, 0, st>>>(p, bcnt, nb);
    k_sel_scan<<<rows, 1024, 0, st>>>(bcnt, nb, total, tstride);
    k_sel_scatter<P><<<dim3(nb, rows), 256, 0, st>>>(p, bcnt, nb, out, ostride);
}

/* direction octant of the shadow ray light -> isp */
__device__ __forceinline__ int shadow_octant(const float4& ip, const float* Lp)
{
    return (ip.x - Lp[0] < 0.0f ? 1 : 0) | (ip.y - Lp[1] < 0.0f ? 2 : 0) | (ip.z - Lp[2] < 0.0f ? 4 : 0);
}

/* Octant reordering: per light, the shaded-pixel list partitioned by the
 * direction octant of the shadow ray, each octant's pixels in list order (a
 * stable partition: neighbouring pixels of one octant stay neighbours).
 * Pass 1 counts, per block of kSelItems list entries, the entries of each
 * (light, octant) row into bcnt[row * nb + block]; k_sel_scan turns those
 * counts into block offsets and row totals, k_oct_scan the totals into each
 * octant's start in its light's list; pass 2 scatters. */
__global__ void __launch_bounds__(256) k_oct_count(const uint32_t* __restrict__ list, const unsigned* __restrict__ nsel,
                                                   const float4* __restrict__ isp, LightsDev L,
                                                   uint32_t* __restrict__ bcnt, uint32_t nb)
{
    __shared__ uint32_t h[8 * kNL];
    const int lane = threadIdx.x & 63;
    if (threadIdx.x < 8 * kNL) h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = *nsel;
    for (int c = 0; c < kSelItems / 256; c++) {
        const uint32_t i     = blockIdx.x * kSelItems + c * 256 + threadIdx.x;
        const bool     valid = i < n;
        const float4   ip    = valid ? isp[list[i]] : make_float4(0.f, 0.f, 0.f, 0.f);
        for (int l = 0; l < L.n; l++) {
            const int oct = shadow_octant(ip, L.p[l]);
            for (int k = 0; k < 8; k++) {
                const uint64_t m = __ballot(valid && oct == k);
                if (lane == 0 && m) atomicAdd(&h[8 * l + k], (uint32_t) __popcll(m));
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 8 * L.n) bcnt[(size_t) threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

/* pass 2: entry i to lists[l * stride + start of its octant + block offset
 * + its rank among the block's entries of that octant before it] */
__global__ void __launch_bounds__(256) k_oct_scatter(const uint32_t* __restrict__ list,
                                                     const unsigned* __restrict__ nsel, const float4* __restrict__ isp,
                                                     LightsDev L, const uint32_t* __restrict__ boff, uint32_t nb,
                                                     const unsigned* __restrict__ cur, uint32_t* __restrict__ lists,
                                                     size_t stride)
{
    __shared__ uint32_t base[8 * kNL];
    __shared__ uint32_t ws[4][8];
    const int           lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x < 8 * L.n) base[threadIdx.x] = cur[threadIdx.x] + boff[(size_t) threadIdx.x * nb + blockIdx.x];
    __syncthreads();
    const uint32_t n = *nsel;
    for (int c = 0; c < kSelItems / 256; c++) {
        const uint32_t i     = blockIdx.x * kSelItems + c * 256 + threadIdx.x;
        const bool     valid = i < n;
        const uint32_t gp    = valid ? list[i] : 0u;
        const float4   ip    = valid ? isp[gp] : make_float4(0.f, 0.f, 0.f, 0.f);
        for (int l = 0; l < L.n; l++) {
            const int oct  = shadow_octant(ip, L.p[l]);
            uint32_t  rank = 0;
            for (int k = 0; k < 8; k++) {
                const uint64_t m = __ballot(valid && oct == k);
                if (lane == 0) ws[w][k] = (uint32_t) __popcll(m);
                if (valid && oct == k) rank = (uint32_t) __popcll(m & ((1ull << lane) - 1ull));
            }
            __syncthreads();
            if (valid) {
                uint32_t pre = base[8 * l + oct];
                for (int v = 0; v < w; v++) pre += ws[v][oct];
                lists[(size_t) l * stride + pre + rank] = gp;
            }
            __syncthreads();
            if (threadIdx.x < 8) base[8 * l + threadIdx.x] += ws[0][threadIdx.x] + ws[1][threadIdx.x] +
                                                              ws[2][threadIdx.x] + ws[3][threadIdx.x];
            __syncthreads();
        }
    }
}

/* per light: exclusive scan of the 8 octant counts into the cursors */
__global__ void k_oct_scan(const unsigned* __restrict__ cnt, unsigned* __restrict__ cur, int nl)
{
    const int l = threadIdx.x;
    if (l >= nl) return;
    unsigned acc = 0;
    for (int k = 0; k < 8; k++) {
        cur[8 * l + k] = acc;
        acc += cnt[8 * l + k];
    }
}

/* Shadow rays of every light in one persistent launch: the work items are
 * light 0's list, then light 1's, ... (light l: lists + l * lstride, count
 * counts[l * cstride]), so the tail of one light's rays overlaps the next
 * light's instead of draining the GPU once per light. Each lane carries its
 * ray's light. disc[pixel] holds the lowest light whose ray discarded the
 * fragment (atomicMin; ~0 none). A light's ray is skipped when a lower light
 * has already discarded; otherwise it is traced as well: k_shade stops at the
 * first discarding light, for the colour, the bits and the expansion count
 * (lexp: per light, counting renders), so the frame does not depend on the
 * order in which rays end (the sample collector applies the same rule). */
template <bool CMP, bool DUAL, bool GRID, bool COUNT, bool GSTK>
__global__ void __launch_bounds__(256) k_shadow(SceneDev s, LightsDev L, const uint32_t* __restrict__ lists,
                                                size_t lstride, const unsigned* __restrict__ counts, int cstride,
                                                const float4* __restrict__ isp, uint8_t* __restrict__ lstat,
                                                size_t npix, unsigned* __restrict__ disc, unsigned* __restrict__ q,
                                                unsigned refill, uint32_t* gstack, int* __restrict__ lexp,
                                                unsigned long long* __restrict__ counters)
{
    extern __shared__ uint32_t lds[];
    const Stack st   = lane_stack<DUAL, GSTK>(lds, gstack, s.maxlevel);
    const int   lane = threadIdx.x & 63;
    uint32_t    total = 0;
    for (int j = 0; j < L.n; j++) total += counts[j * cstride];
    Refill      rf(total);
    bool        active = false;
    uint32_t    gp     = 0;
    int         lt     = 0; /* the light of this lane's ray */
    Ray         r      = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float4      ip     = make_float4(0.f, 0.f, 0.f, 0.f);
    TraceState  t;
    unsigned    n_abort = 0, n_ray = 0;

    auto finish = [&](int stt) {
        uint8_t ls;
        if (stt == ST_DISCARD) {
            ls = LS_DISCARD;
            atomicMin(&disc[gp], (unsigned) lt);
        } else {
            if (stt == ST_ABORT) n_abort++;
            float4 li  = stt == ST_HIT ? t.entry : make_float4(0.f, 0.f, 0.f, 0.f);
            float  ix  = li.x - ip.x, iy = li.y - ip.y, iz = li.z - ip.z;
            float  sqr = ix * ix + iy * iy + iz * iz;
            ls         = (15.0f < sqr) ? LS_OCCLUDED : LS_LIT; /* step(sqr, 15.0) */
        }
        lstat[(size_t) lt * npix + gp] = ls;
        if (COUNT) lexp[(size_t) lt * npix + gp] = (int) t.exps;
    };

    int fin = ST_CONT; /* status of a ray that ended and awaits its outputs */
    for (;;) {
        if (fin != ST_CONT) { /* rays that ended since the last refill point, together */
            finish(fin);
            fin = ST_CONT;
        }
        const uint64_t idle = __ballot(!active);
        if (rf.left > 0 && (__popcll(idle) >= (int) refill || idle == ~0ull)) {
            const uint32_t k = rf.fetch(&q[Q_SHADOW], idle, lane);
            if (k != UINT32_MAX) {
                /* item k -> (light, index in its list); the light's position by
                 * selects (a per-lane index into the lights would go through
                 * scratch memory) */
                uint32_t kk = k;
                int      l  = 0;
                float    lx = L.p[0][0], ly = L.p[0][1], lz = L.p[0][2];
                for (int j = 0; j + 1 < L.n; j++) {
                    const uint32_t c    = counts[j * cstride];
                    const bool     past = l == j && kk >= c;
                    kk -= past ? c : 0u;
                    l += past ? 1 : 0;
                    lx = past ? L.p[j + 1][0] : lx;
                    ly = past ? L.p[j + 1][1] : ly;
                    lz = past ? L.p[j + 1][2] : lz;
                }
                lt = l;
                gp = lists[(size_t) l * lstride + kk];
                if (disc[gp] < (unsigned) l) { /* a lower light discarded the fragment */
                    lstat[(size_t) l * npix + gp] = LS_SKIPPED;
                    if (COUNT) lexp[(size_t) l * npix + gp] = 0;
                } else {
                    ip      = isp[gp];
                    r.px    = lx;
                    r.py    = ly;
                    r.pz    = lz;
                    r.dx    = ip.x - lx;
                    r.dy    = ip.y - ly;
                    r.dz    = ip.z - lz;
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
        if (active) step_run<CMP, DUAL, GRID>(s, r, st, t, active, keep_active(rf, refill), fin);
    }
    n_ray   = wave_sum(n_ray);
    n_abort = wave_sum(n_abort);
    if (lane == 0) {
        if (n_ray) atomicAdd(&counters[CNT_SHADOW], (unsigned long long) n_ray);
        if (n_abort) atomicAdd(&counters[CNT_ABORT], (unsigned long long) n_abort);
    }
}

/* ---- camera and shadow rays in one persistent launch ------------------------
 *
 * k_frame runs k_primary's work items, then the shadow rays of k_shadow, in
 * one launch: a wave whose camera work has run dry takes shadow rays while
 * the last camera rays of other waves are still being traced, so the end of
 * the camera pass no longer leaves the SIMDs with a few lanes each, and no
 * compaction runs between the passes.
 *
 * Hand-over: per (light l, camera work item i) one slot of four words at
 * slots + 4 (l * nitems + i): the state word and the entry point (x, y, z).
 * All four hold kSlotFree before the launch. The camera lane that finishes
 * item i writes the entry point, then exchanges the state word for the
 * pixel index (kSlotNone when the pixel is not shaded); the shadow lane
 * that claims the slot exchanges it for kSlotDefer. Whichever comes second
 * sees the other's value: a shadow lane that finds a pixel index traces the
 * ray (waiting, within a bound, until the entry point's words are visible:
 * they are written before the exchange), and a camera lane that finds
 * kSlotDefer traces that light's ray itself after its camera ray. No lane
 * waits for another lane's ray. Slots are claimed in light-major item order,
 * the order of k_shadow's lists. Slot words are read and written with
 * device-scope atomics, so lanes on different XCDs see each other's values. */
constexpr uint32_t kSlotFree  = 0xFFFFFFFFu; /* also the entry-point words not written yet (a NaN) */
constexpr uint32_t kSlotDefer = 0xFFFFFFFEu;
constexpr uint32_t kSlotNone  = 0xFFFFFFFDu;
constexpr int      kSlotSpin  = 1 << 16; /* bound on the wait for an entry point; past it the ray counts as aborted */

__device__ __forceinline__ uint32_t slot_load(uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void slot_store(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t slot_xchg(uint32_t* p, uint32_t v)
{
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* lane state of k_frame, packed into three words so that the kernel needs
 * few more registers than the separate launches (88-89 against 78-79, 5
 * waves per SIMD): gp = the ray's batch pixel; k = the camera ray's work
 * item, or the shadow ray's slot; tag = kind bits, the sprite slot of a
 * sprite trace, and the lights still deferred to this lane (bits 8..15) */
constexpr uint32_t kTagShadow = 0x80000000u; /* the lane's ray is a shadow ray */
constexpr uint32_t kTagSprite = 0x40000000u; /* camera-side sprite trace; low bits: sprite slot */
constexpr uint32_t kTagFresh  = 0x20000000u; /* k still holds the finished camera item */
constexpr uint32_t kTagPend   = 0x0000ff00u; /* deferred lights, bit 8 + l */

/* light and camera item of slot k = l * nitems + item */
__device__ __forceinline__ int slot_light(uint32_t k, uint32_t nitems, int nl, uint32_t& item)
{
    uint32_t it = k;
    int      l  = 0;
    for (int j = 1; j < nl; j++) {
        const bool past = it >= nitems;
        it -= past ? nitems : 0u;
        l += past ? 1 : 0;
    }
    item = it;
    return l;
}

template <bool CMP, bool DUAL, bool GRID, bool COUNT, bool GSTK>
__global__ void __launch_bounds__(256)
    k_frame(SceneDev s, BatchDev b, PrimOut o, LightsDev L, uint32_t nitems, uint32_t nspr, int strip,
            int* __restrict__ sprite, uint32_t* __restrict__ slots, uint8_t* __restrict__ lstat, size_t npix,
            unsigned* __restrict__ disc, unsigned* __restrict__ q, unsigned refill_p, unsigned refill_s,
            uint32_t* gstack, int* __restrict__ lexp, unsigned long long* __restrict__ counters)
{
    extern __shared__ uint32_t lds[];
    const Stack st   = lane_stack<DUAL, GSTK>(lds, gstack, s.maxlevel);
    const int   lane = threadIdx.x & 63;
    const int   nl   = L.n;
    Refill      rp(nspr + nitems), rs(nitems * (uint32_t) nl);
    bool        active = false;
    uint32_t    gp = 0, k = 0, tag = 0;
    Ray         r = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    TraceState  t;
    unsigned    n_exp = 0;
    int         fin   = ST_CONT; /* status of a ray that ended and awaits its outputs */
    /* shaded camera hits, shadow rays traced and of those the deferred ones,
     * per wave: ballots taken in wave-uniform code, so the sums sit in scalar
     * registers */
    uint32_t    w_hit = 0, w_ray = 0, w_def = 0;

    auto abort_count = [&]() { atomicAdd(&counters[CNT_ABORT], 1ull); }; /* rare: no register kept for it */

    /* the shadow ray's result: lit when the light's first leaf starts within
     * sqrt(15) of the entry point it aims at (octree_fsh.c:431-441) */
    auto finish_shadow = [&](int stt) {
        uint32_t  item;
        const int l = slot_light(k, nitems, nl, item);
        uint8_t   ls;
        if (stt == ST_DISCARD) {
            ls = LS_DISCARD;
            atomicMin(&disc[gp], (unsigned) l);
        } else {
            if (stt == ST_ABORT) abort_count();
            uint32_t*   w   = slots + 4 * (size_t) k;
            const float px  = __uint_as_float(slot_load(w + 1));
            const float py  = __uint_as_float(slot_load(w + 2));
            const float pz  = __uint_as_float(slot_load(w + 3));
            float4      li  = stt == ST_HIT ? t.entry : make_float4(0.f, 0.f, 0.f, 0.f);
            float       ix  = li.x - px, iy = li.y - py, iz = li.z - pz;
            float       sqr = ix * ix + iy * iy + iz * iz;
            ls              = (15.0f < sqr) ? LS_OCCLUDED : LS_LIT; /* step(sqr, 15.0) */
        }
        lstat[(size_t) l * npix + gp] = ls;
        if (COUNT) lexp[(size_t) l * npix + gp] = (int) t.exps;
    };

    /* the ray of slot k (light l) to the entry point (px, py, pz) of pixel gp */
    auto start_shadow = [&](int l, float px, float py, float pz) -> bool {
        if (disc[gp] < (unsigned) l) { /* a lower light discarded the fragment */
            lstat[(size_t) l * npix + gp] = LS_SKIPPED;
            if (COUNT) lexp[(size_t) l * npix + gp] = 0;
            return false;
        }
        /* the light's position by selects (a per-lane index into the lights
         * would go through scratch memory) */
        float lx = L.p[0][0], ly = L.p[0][1], lz = L.p[0][2];
        for (int j = 1; j < nl; j++) {
            lx = l == j ? L.p[j][0] : lx;
            ly = l == j ? L.p[j][1] : ly;
            lz = l == j ? L.p[j][2] : lz;
        }
        r.px          = lx;
        r.py          = ly;
        r.pz          = lz;
        r.dx          = px - lx;
        r.dy          = py - ly;
        r.dz          = pz - lz;
        const int stt = trace_begin<false>(s, r, t);
        if (stt == ST_CONT) active = true;
        else fin = stt; /* outputs at the next refill point */
        return true;
    };

    /* the camera ray's outputs (k_primary's), then its slots */
    auto finish_camera = [&](int stt) -> bool {
        if (tag & kTagSprite) { /* white if the light is in front of the first surface (resv / lghtv.x > 1) */
            const float ix = stt == ST_HIT ? t.entry.x : 0.0f;
            int         f  = 0;
            if (stt == ST_DISCARD) f |= 1;
            if ((ix - r.px) / r.dx > 1.0f) f |= 2;
            if (stt == ST_ABORT) f |= 4;
            sprite[tag & ~kTagSprite] = f;
            tag                       = 0;
            return false;
        }
        float4  isp = make_float4(0.f, 0.f, 0.f, 0.f);
        int     ms = -1, md = -1, leaf = -1;
        uint8_t f  = 0;
        bool    sh = false;
        if (stt == ST_DISCARD) f = QBT_PX_DISCARD;
        if (stt == ST_ABORT) abort_count();
        if (stt == ST_HIT) {
            const int L_ = s.maxlevel;
            ms           = (t.ns == 0 && L_ > 0) ? 0 : s.stat.model[t.ns];
            md           = 0;
            if (DUAL) md = (t.nd == 0 && L_ > 0) ? 0 : s.dyn.model[t.nd];
            leaf = CMP ? s.stat.ref[t.ns] : t.ns; /* reference node index */
            isp  = t.entry;
            f    = QBT_PX_LEAF;
            if (isp.w > 0.0f) {
                f |= QBT_PX_SHADED;
                sh = true;
            }
        }
        o.isp[gp]   = isp;
        o.ms[gp]    = ms;
        o.md[gp]    = md;
        o.leaf[gp]  = leaf;
        o.flags[gp] = f;
        if (COUNT) {
            o.exp[gp] = (int) t.exps;
            n_exp += t.exps;
        }
        uint32_t pend = 0;
        for (int l = 0; l < nl; l++) {
            uint32_t* w = slots + 4 * ((size_t) l * nitems + k);
            if (sh) {
                slot_store(w + 1, __float_as_uint(isp.x));
                slot_store(w + 2, __float_as_uint(isp.y));
                slot_store(w + 3, __float_as_uint(isp.z));
            }
            const uint32_t old = slot_xchg(w, sh ? gp : kSlotNone);
            if (sh && old == kSlotDefer) pend |= 1u << l;
        }
        tag = pend ? (kTagShadow | kTagFresh | pend << 8) : 0u;
        return sh;
    };

    for (;;) {
        bool hit_now = false, ray_now = false, def_now = false; /* this lane's events of this refill point */
        if (fin != ST_CONT) { /* rays that ended since the last refill point, together */
            if (tag & kTagShadow) finish_shadow(fin);
            else hit_now = finish_camera(fin);
            fin = ST_CONT;
        }
        if (!active && (tag & kTagPend)) { /* the next deferred light of this lane's pixel */
            uint32_t item = k;
            if (!(tag & kTagFresh)) (void) slot_light(k, nitems, nl, item);
            const uint32_t pend = (tag & kTagPend) >> 8;
            const int      l    = __ffs(pend) - 1;
            tag                 = kTagShadow | (pend & (pend - 1u)) << 8;
            k                   = (uint32_t) l * nitems + item;
            uint32_t* w         = slots + 4 * (size_t) k; /* written by this lane */
            ray_now = start_shadow(l, __uint_as_float(slot_load(w + 1)), __uint_as_float(slot_load(w + 2)),
                                   __uint_as_float(slot_load(w + 3)));
            def_now = ray_now;
        }
        /* free lanes: no ray, no deferred light, no outputs still to write */
        const uint64_t idle = __ballot(!active && !(tag & kTagPend) && fin == ST_CONT);
        if (rp.left > 0) {
            if (__popcll(idle) >= (int) refill_p || idle == ~0ull) {
                const uint32_t fetched = rp.fetch(&q[Q_PRIM], idle, lane);
                if (fetched != UINT32_MAX && fetched < nspr) {
                    const int      v = (int) (fetched / (uint32_t) nl), l = (int) (fetched % (uint32_t) nl);
                    const ViewDev& V = b.views[v];
                    tag              = kTagSprite | (uint32_t) (v * kNL + l);
                    r.px             = V.cam[0];
                    r.py             = V.cam[1];
                    r.pz             = V.cam[2];
                    r.dx             = V.light[l][0] - V.cam[0];
                    r.dy             = V.light[l][1] - V.cam[1];
                    r.dz             = V.light[l][2] - V.cam[2];
                    const int stt    = trace_begin<false>(s, r, t);
                    if (stt == ST_CONT) active = true;
                    else fin = stt;
                } else if (fetched != UINT32_MAX) {
                    const uint32_t it = fetched - nspr;
                    const int      v  = batch_view(b.item_off, b.n, it);
                    const ViewDev& V  = b.views[v];
                    int            px, py;
                    tag = 0;
                    k   = it;
                    if (item_pixel(it - b.item_off[v], V.w, V.h, strip, px, py)) {
                        gp = b.pix_off[v] + (uint32_t) py * (uint32_t) V.w + (uint32_t) px;
                        float csv[3];
                        pixel_dir(V, px, py, csv);
                        r.px          = V.cam[0];
                        r.py          = V.cam[1];
                        r.pz          = V.cam[2];
                        r.dx          = csv[0];
                        r.dy          = csv[1];
                        r.dz          = csv[2];
                        const int stt = trace_begin<false>(s, r, t);
                        if (stt == ST_CONT) active = true;
                        else fin = stt;
                    } else { /* padding item of an edge tile: no shadow rays */
                        for (int l = 0; l < nl; l++) (void) slot_xchg(slots + 4 * ((size_t) l * nitems + it), kSlotNone);
                    }
                }
            }
        } else if (rs.left > 0) {
            if (__popcll(idle) >= (int) refill_s || idle == ~0ull) {
                const uint32_t ks = rs.fetch(&q[Q_SHADOW], idle, lane);
                if (ks != UINT32_MAX) {
                    uint32_t*      w   = slots + 4 * (size_t) ks;
                    const uint32_t old = slot_xchg(w, kSlotDefer);
                    if (old < kSlotNone) { /* the camera lane has published its pixel */
                        uint32_t x = slot_load(w + 1), y = slot_load(w + 2), z = slot_load(w + 3);
                        for (int i = 0; (x == kSlotFree || y == kSlotFree || z == kSlotFree) && i < kSlotSpin; i++) {
                            __builtin_amdgcn_s_sleep(1);
                            x = slot_load(w + 1);
                            y = slot_load(w + 2);
                            z = slot_load(w + 3);
                        }
                        uint32_t  item;
                        const int l = slot_light(ks, nitems, nl, item);
                        gp          = old;
                        k           = ks;
                        tag         = kTagShadow;
                        if (x == kSlotFree || y == kSlotFree || z == kSlotFree) {
                            abort_count(); /* reported to the caller as an aborted render */
                            lstat[(size_t) l * npix + gp] = LS_OCCLUDED;
                        } else {
                            ray_now = start_shadow(l, __uint_as_float(x), __uint_as_float(y), __uint_as_float(z));
                        }
                    }
                }
            }
        }
        w_hit += (uint32_t) __popcll(__ballot(hit_now));
        w_ray += (uint32_t) __popcll(__ballot(ray_now));
        w_def += (uint32_t) __popcll(__ballot(def_now));
        if (__ballot(active) == 0ull) {
            if (rp.left <= 0 && rs.left <= 0 && __ballot((tag & kTagPend) != 0u || fin != ST_CONT) == 0ull) break;
            continue;
        }
        /* step until `keep` lanes are left: refill thresholds as in the two
         * separate launches; a wave with deferred lights returns as soon as
         * one ray ends, so that lane starts its next light */
        const int keep = __ballot((tag & kTagPend) != 0u) ? 64
                         : rp.left > 0                    ? 65 - (int) refill_p
                         : rs.left > 0                    ? 65 - (int) refill_s
                                                          : 1;
        if (active) step_run<CMP, DUAL, GRID>(s, r, st, t, active, keep, fin);
    }
    if (COUNT) n_exp = wave_sum(n_exp);
    if (lane == 0) {
        if (w_hit) atomicAdd(&counters[CNT_HITS], (unsigned long long) w_hit);
        if (w_ray) atomicAdd(&counters[CNT_SHADOW], (unsigned long long) w_ray);
        if (w_def) atomicAdd(&counters[CNT_DEFER], (unsigned long long) w_def);
        if (COUNT && n_exp) atomicAdd(&counters[CNT_EXP], (unsigned long long) n_exp);
        if (COUNT && n_exp) atomicAdd(&counters[CNT_EXP_PRIM], (unsigned long long) n_exp);
    }
}

/* per batch pixel: attributes, shading, sprite, outputs (octree_fsh.c:421-463);
 * counting renders add the shadow expansions of the lights up to the first
 * discarding one (lexp) to the pixel's count and to the counters */
template <bool DUAL, bool COUNT>
__global__ void __launch_bounds__(256) k_shade(SceneDev s, BatchDev b, PrimOut in, const uint8_t* __restrict__ lstat,
                                               const int* __restrict__ lexp, uint32_t npix, unsigned mask,
                                               unsigned long long* __restrict__ counters)
{
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned       n_exp = 0;
    if (COUNT) {
        if (o < npix && (in.flags[o] & QBT_PX_SHADED) && !(in.flags[o] & QBT_PX_DISCARD)) {
            const int nl = b.views[batch_view(b.pix_off, b.n, o)].nlights;
            for (int l = 0; l < nl; l++) {
                if (lstat[(size_t) l * npix + o] == LS_DISCARD) break;
                n_exp += (unsigned) lexp[(size_t) l * npix + o];
            }
        }
        const unsigned w = wave_sum(n_exp);
        if ((threadIdx.x & 63) == 0 && w) atomicAdd(&counters[CNT_EXP], (unsigned long long) w);
    }
    if (o >= npix) return;
    const int      vi     = batch_view(b.pix_off, b.n, o);
    const ViewDev& v      = b.views[vi];
    const uint32_t lo     = o - b.pix_off[vi];
    const int      px     = (int) (lo % (uint32_t) v.w), py = (int) (lo / (uint32_t) v.w);
    const int*     spr    = b.sprite + vi * kNL;
    uint32_t       flg    = in.flags[o];
    uint32_t       sbits  = 0;
    float          col[4] = {0.f, 0.f, 0.f, 0.f};
    float          nrm[4] = {0.f, 0.f, 0.f, 0.f};
    const int      ms = in.ms[o], md = in.md[o];
    const float4   isp = in.isp[o];
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
                uint8_t ls = lstat[(size_t) l * npix + o];
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
                    int sf = spr[l];
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
    const FrameDev& f = b.frames[vi];
    if (mask & QBT_OUT_HIT) f.hit[lo] = ms;
    if (mask & QBT_OUT_LEAF) f.leaf[lo] = in.leaf[o];
    if (mask & QBT_OUT_DYN) f.dyn[lo] = md;
    if (mask & QBT_OUT_SHADOW) f.shadow[lo] = (uint8_t) sbits;
    if (mask & QBT_OUT_FLAGS) f.flags[lo] = (uint8_t) flg;
    if (mask & QBT_OUT_RGBA8) f.rgba8[lo] = pack_rgba8(col[0], col[1], col[2], col[3]);
    if (mask & QBT_OUT_RGBAF) f.rgbaf[lo] = make_float4(col[0], col[1], col[2], col[3]);
    if (mask & QBT_OUT_ISP) f.isp[lo] = isp;
    if (COUNT && (mask & QBT_OUT_EXP)) f.exp[lo] = in.exp[o] + (int) n_exp;
}

__global__ void k_zero_u32(unsigned* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_fill_u32n(unsigned* p, size_t n, unsigned v)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned nblk(size_t n, unsigned bs = 256) { return (unsigned) ((n + bs - 1) / bs); }

/* Blocks of a persistent traversal launch over `items` rays. All the blocks
 * that fit (6 waves per SIMD for the traversal kernels: VGPRs and LDS
 * stacks) when there are at least kFullGridRays rays per lane of that grid;
 * fewer, 4 waves per SIMD, for shorter launches, where the end of a launch
 * (each wave stepping its last few long rays with most lanes idle once the
 * queue runs dry) is a large share: there each wave serves half as many rays
 * again before the queue runs dry. Measured: the 1080p camera launch of the
 * 90M-point scene (5 rays per lane) -2.6% with 4, the one-light shadow launch
 * of the 10M-point scene (1.3 rays per lane) -8%; 3840x2160 frames (20-30
 * rays per lane) and 64-view batches 5-8% slower with 4. */
constexpr int kSmallGridWavesPerSimd = 4;
constexpr int kFullGridRays          = 16;

template <class K>
int resident_blocks(K kernel, int bs, size_t lds)
{
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, bs, lds) != hipSuccess || per <= 0) {
        (void) hipGetLastError();
        per = 1;
    }
    return per;
}

/* lanes of the full resident grid of `kernel` */
template <class K>
size_t grid_lanes(qbt_ctx* c, K kernel, int bs, size_t lds)
{
    return (size_t) resident_blocks(kernel, bs, lds) * (size_t) c->num_cus * (size_t) bs;
}

/* whether `items` rays fill the full resident grid of `kernel` (at least
 * kFullGridRays per lane) */
template <class K>
bool fills_grid(qbt_ctx* c, K kernel, int bs, size_t lds, size_t items)
{
    return items >= (size_t) kFullGridRays * grid_lanes(c, kernel, bs, lds);
}

template <class K>
int persistent_grid(qbt_ctx* c, K kernel, int bs, size_t lds, size_t items)
{
    const int per = resident_blocks(kernel, bs, lds);
    if (fills_grid(c, kernel, bs, lds, items)) return per * c->num_cus;
    const int cap = std::max(1, 4 * kSmallGridWavesPerSimd / std::max(1, bs / 64)); /* blocks per CU */
    return std::min(per, cap) * c->num_cus;
}

/* the launch configuration a captured graph depends on (kernel arguments) */
struct GraphKey {
    SceneDev s;
    void*    bufs[18];
    uint32_t nitems, npix, nviews, nlights;
    unsigned mask, flags;
    int      bs, refill, gp, gs, gf, stack, reorder, strip, merged;
    float    band[2];
    float    lights[kNL][3];
    uint32_t tmpl;
};

} // namespace

void qbt_ctx::graph_free()
{
    if (graph_exec) (void) hipGraphExecDestroy(graph_exec);
    if (graph) (void) hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph      = nullptr;
    graph_key.clear();
}

/* internal per-batch buffers, sized for the largest batch so far */
int qbt_ctx::ensure_wave_bufs(size_t npix, int nlights)
{
    int          rc;
    const size_t nl = (size_t) (nlights > 0 ? nlights : 1);
    if ((rc = dev_alloc(wb_isp, npix * 16)) || (rc = dev_alloc(wb_ms, npix * 4)) || (rc = dev_alloc(wb_md, npix * 4)) ||
        (rc = dev_alloc(wb_leaf, npix * 4)) || (rc = dev_alloc(wb_flags, npix)) || (rc = dev_alloc(wb_disc, npix * 4)) ||
        (rc = dev_alloc(wb_lstat, npix * nl)) || (rc = dev_alloc(wb_exp, npix * 4)))
        return rc;
    return QBT_OK;
}

template <bool CMP, bool DUAL, bool GRID, bool COUNT, bool GSTK>
static int render_batch_t(qbt_ctx* c, const SceneDev& s, const BatchHost& bh, const BatchDev& b, const LightsDev& L,
                          unsigned mask, unsigned flags)
{
    hipStream_t    st        = c->stream;
    const uint32_t npix      = bh.npix, nitems = bh.nitems;
    const int      nl        = L.n;
    const bool     surrogate = (flags & QBT_RENDER_SURROGATE) && nl > 0;
    const bool     collect   = (flags & QBT_RENDER_COLLECT) && nl > 0 && !surrogate;
    const bool     reorder   = c->tune_reorder && nl > 0;
    const uint32_t nspr      = (uint32_t) b.n * (uint32_t) (nl > 0 ? nl : 0); /* sprite traces */
    int            rc;
    /* every buffer exists before a capture starts */
    if ((rc = c->ensure_wave_bufs(npix, nl)) || (rc = c->dev_alloc(c->wb_cand, (size_t) nitems * 4)) ||
        (rc = c->dev_alloc(c->wb_items, (size_t) npix * 4)) || (rc = c->dev_alloc(c->wb_q, sizeof(unsigned) * Q_N)))
        return rc;
    if (reorder && (rc = c->dev_alloc(c->wb_lists, (size_t) npix * 4 * nl))) return rc;
    const uint32_t nb_max = ((nitems > npix ? nitems : npix) + kSelItems - 1) / kSelItems + 1;
    /* rows of the compactions: one per light, or one per (light, octant) */
    if ((rc = c->dev_alloc(c->wb_bcnt, (size_t) nb_max * 4 * (reorder ? 8 * nl : (nl > 0 ? nl : 1))))) return rc;
    if (surrogate && (rc = c->dev_alloc(c->wb_fb, (size_t) npix * 4 * nl))) return rc;
    if (COUNT && nl > 0 && (rc = c->dev_alloc(c->wb_lexp, (size_t) npix * 4 * nl))) return rc;

    const int      bs     = c->block_threads(s.maxlevel, DUAL);
    const size_t   lds    = GSTK ? 0 : qbt_stack_bytes(s.maxlevel, DUAL, bs);
    /* refill thresholds: the tuned one for both launches, else the measured
     * defaults (camera rays are shorter, so their refills cost more of a ray) */
    const bool     tuned     = c->tune_refill > 0 && c->tune_refill <= 64;
    const unsigned refill    = (unsigned) (tuned ? c->tune_refill : 24);
    const unsigned refill_sh = (unsigned) (tuned ? c->tune_refill : 16);
    auto           kp     = k_primary<CMP, DUAL, GRID, COUNT, GSTK>;
    auto           ks     = k_shadow<CMP, DUAL, GRID, COUNT, GSTK>;
    auto           kf     = k_frame<CMP, DUAL, GRID, COUNT, GSTK>;
    /* exact renders: camera and shadow rays in one launch (k_frame). Tuning
     * merge 2 always; merge 1 where it was measured to pay: at least two
     * lights (shadow work enough to fill the end of the camera pass, a few
     * long rays on mostly idle SIMDs) and a camera pass of 1 to 16 rays per
     * lane of the full grid (longer passes keep the separate launches at the
     * full grid; shorter ones are latency-bound either way) */
    const size_t cam  = (size_t) nitems + nspr;
    const bool   auto_merge = nl >= 2 && cam >= grid_lanes(c, kp, bs, lds) && !fills_grid(c, kp, bs, lds, cam);
    const bool merged = (c->tune_merge == 2 || (c->tune_merge == 1 && auto_merge)) && nl > 0 && !surrogate &&
                        !collect && !reorder;
    if (merged && (rc = c->dev_alloc(c->wb_slots, (size_t) nitems * 16 * nl))) return rc;
    /* resident blocks, but no more than one work item per lane: small
     * batches then start fewer waves (less launch and refill-counter
     * contention); shadow items are at most one per batch pixel and light */
    const size_t   np = (size_t) nitems + nspr, ns = (size_t) npix * (nl > 0 ? nl : 1);
    const int      gp = std::min(persistent_grid(c, kp, bs, lds, np), (int) nblk(np, (unsigned) bs));
    const int      gs = std::min(persistent_grid(c, ks, bs, lds, ns), (int) nblk(ns, (unsigned) bs));
    const int      gf = merged ? std::min(persistent_grid(c, kf, bs, lds, np + ns), (int) nblk(np + ns, (unsigned) bs)) : 0;
    const int gmax = std::max(std::max(gp, gs), gf);
    if (GSTK && (rc = c->dev_alloc(c->wb_gstack, qbt_stack_bytes(s.maxlevel, DUAL, gmax * bs)))) return rc;

    PrimOut po;
    po.isp   = (float4*) c->wb_isp.p;
    po.ms    = (int*) c->wb_ms.p;
    po.md    = (int*) c->wb_md.p;
    po.leaf  = (int*) c->wb_leaf.p;
    po.flags = (uint8_t*) c->wb_flags.p;
    po.cand  = (uint32_t*) c->wb_cand.p;
    po.exp   = (int*) c->wb_exp.p;
    unsigned*           q     = (unsigned*) c->wb_q.p;
    unsigned long long* cnt   = (unsigned long long*) c->counters.p;
    uint32_t*           items = (uint32_t*) c->wb_items.p;
    uint32_t*           gstk  = (uint32_t*) c->wb_gstack.p;
    /* shadow lists: per light (octant order) or one shared list */
    const uint32_t* lists   = reorder ? (const uint32_t*) c->wb_lists.p : items;
    const size_t    lstride = reorder ? npix : 0;

    /* the launches of one batch; traversal kernels bracketed for qbt_profile_*
     * (no-ops unless profiling, which is never on while capturing) */
    auto enqueue = [&]() -> int {
        int err;
        k_zero_u32<<<nblk(Q_N), 256, 0, st>>>(q, Q_N);
        k_fill_u32n<<<nblk(npix), 256, 0, st>>>((unsigned*) c->wb_disc.p, npix, ~0u);
        if (COUNT && nl > 0) /* lights the surrogate decided add no shadow expansions */
            k_fill_u32n<<<nblk((size_t) npix * nl), 256, 0, st>>>((unsigned*) c->wb_lexp.p, (size_t) npix * nl, 0u);
        if (merged) {
            const size_t nw = (size_t) nitems * 4 * nl;
            k_fill_u32n<<<nblk(nw), 256, 0, st>>>((unsigned*) c->wb_slots.p, nw, kSlotFree);
            c->prof_mark_begin(QBT_PROF_FRAME);
            kf<<<gf, bs, lds, st>>>(s, b, po, L, nitems, nspr, c->tune_strip, (int*) c->sprite.p,
                                    (uint32_t*) c->wb_slots.p, (uint8_t*) c->wb_lstat.p, npix,
                                    (unsigned*) c->wb_disc.p, q, refill, refill_sh, gstk, (int*) c->wb_lexp.p, cnt);
            c->prof_mark_end();
            k_shade<DUAL, COUNT><<<nblk(npix), 256, 0, st>>>(s, b, po, (const uint8_t*) c->wb_lstat.p,
                                                             (const int*) c->wb_lexp.p, npix, mask, cnt);
            return QBT_OK;
        }
        c->prof_mark_begin(QBT_PROF_PRIMARY);
        kp<<<gp, bs, lds, st>>>(s, b, po, nitems, nspr, nl > 0 ? nl : 1, c->tune_strip, (int*) c->sprite.p, q,
                                refill, gstk, cnt);
        c->prof_mark_end();
        compact(st, PredCand{po.cand, nitems}, nitems, 1, (uint32_t*) c->wb_bcnt.p, q + Q_NSEL, 1, items, 0);
        if (reorder) {
            const uint32_t nbo = (npix + kSelItems - 1) / kSelItems + 1;
            k_oct_count<<<nbo, 256, 0, st>>>(items, q + Q_NSEL, po.isp, L, (uint32_t*) c->wb_bcnt.p, nbo);
            k_sel_scan<<<8 * nl, 1024, 0, st>>>((uint32_t*) c->wb_bcnt.p, nbo, q + Q_OCT, 1);
            k_oct_scan<<<1, 64, 0, st>>>(q + Q_OCT, q + Q_CUR, nl);
            k_oct_scatter<<<nbo, 256, 0, st>>>(items, q + Q_NSEL, po.isp, L, (const uint32_t*) c->wb_bcnt.p, nbo,
                                                q + Q_CUR, (uint32_t*) c->wb_lists.p, npix);
        }
        if (surrogate) {
            if ((err = qbt_surrogate_decide(c, s, DUAL, L, lists, lstride, q + Q_NSEL, npix))) return err;
            compact(st, PredUnsure{lists, lstride, q + Q_NSEL, (const uint8_t*) c->wb_lstat.p, npix}, npix, nl,
                    (uint32_t*) c->wb_bcnt.p, q + Q_FB, 1, (uint32_t*) c->wb_fb.p, npix);
        }
        if (nl > 0) {
            /* exact mode: every shaded pixel per light; surrogate mode: each
             * light's undecided ones */
            const uint32_t* li  = surrogate ? (const uint32_t*) c->wb_fb.p : lists;
            const size_t    lst = surrogate ? npix : lstride;
            const unsigned* ln  = surrogate ? q + Q_FB : q + Q_NSEL;
            c->prof_mark_begin(QBT_PROF_SHADOW);
            ks<<<gs, bs, lds, st>>>(s, L, li, lst, ln, surrogate ? 1 : 0, po.isp, (uint8_t*) c->wb_lstat.p, npix,
                                    (unsigned*) c->wb_disc.p, q, refill_sh, gstk, (int*) c->wb_lexp.p, cnt);
            c->prof_mark_end();
        }
        if (collect && (err = qbt_collect_batch(c, s, DUAL, L, lists, lstride, q + Q_NSEL, npix))) return err;
        k_shade<DUAL, COUNT><<<nblk(npix), 256, 0, st>>>(s, b, po, (const uint8_t*) c->wb_lstat.p,
                                                         (const int*) c->wb_lexp.p, npix, mask, cnt);
        return QBT_OK;
    };

    const bool use_graph = c->tune_graph && !COUNT && !collect && !c->profiling;
    if (!use_graph) {
        if ((rc = enqueue())) return rc;
        HIPCHK(hipGetLastError());
        c->last.launches += merged ? 5 + (COUNT ? 1 : 0) : 6 + (nl > 0 ? 1 : 0) + (reorder ? 4 : 0) + (surrogate ? 4 : 0);
        return QBT_OK;
    }

    /* graph replay when the launch configuration is unchanged, else capture */
    GraphKey key;
    memset(&key, 0, sizeof(key));
    key.s      = s;
    void* bp[] = {c->wb_isp.p,   c->wb_ms.p,    c->wb_md.p,   c->wb_leaf.p,  c->wb_flags.p, c->wb_disc.p,
                  c->wb_lstat.p, c->wb_exp.p,   c->wb_cand.p, c->wb_items.p, c->wb_q.p,     c->wb_lists.p,
                  c->wb_fb.p,    c->wb_gstack.p, c->bt_tab.p, c->sprite.p, c->wb_bcnt.p, c->wb_slots.p};
    static_assert(sizeof(bp) == sizeof(key.bufs), "graph key buffers");
    memcpy(key.bufs, bp, sizeof(bp));
    key.nitems = nitems, key.npix = npix, key.nviews = (uint32_t) b.n, key.nlights = (uint32_t) nl;
    key.mask = mask, key.flags = flags, key.bs = bs, key.refill = (int) (refill | refill_sh << 8), key.gp = gp, key.gs = gs;
    key.stack = GSTK, key.reorder = reorder, key.strip = c->tune_strip, key.gf = gf, key.merged = merged;
    memcpy(key.lights, L.p, sizeof(key.lights));
    key.tmpl = (CMP ? 1u : 0u) | (DUAL ? 2u : 0u) | (GRID ? 4u : 0u);
    if (surrogate) qbt_surrogate_band(c, &key.band[0], &key.band[1]);
    const uint8_t* kb = reinterpret_cast<const uint8_t*>(&key);
    if (c->graph_exec && c->graph_key.size() == sizeof(key) && memcmp(c->graph_key.data(), kb, sizeof(key)) == 0) {
        c->graph_replays++;
    } else {
        c->graph_free();
        (void) hipGetLastError();
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int        err = enqueue();
        const hipError_t le  = hipGetLastError();
        hipGraph_t       g   = nullptr;
        const hipError_t ce  = hipStreamEndCapture(st, &g);
        if (err || le != hipSuccess || ce != hipSuccess) {
            if (g) (void) hipGraphDestroy(g);
            if (err) return err;
            HIPCHK(le != hipSuccess ? le : ce);
        }
        c->graph = g;
        HIPCHK(hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
        c->graph_key.assign(kb, kb + sizeof(key));
    }
    HIPCHK(hipGraphLaunch(c->graph_exec, st));
    c->last.launches += 1;
    return QBT_OK;
}

int qbt_ctx::render_batch_wavefront(const SceneDev& s, bool cmp, bool dual, bool grid, bool count,
                                    const BatchHost& bh, const BatchDev& b, const LightsDev& L, unsigned mask,
                                    unsigned flags)
{
    if (bh.npix == 0) return QBT_OK;
    const bool gstk = tune_stack == 1;
    return qbt_dispatch4(cmp, dual, grid, count, [&](auto C_, auto D_, auto G_, auto N_) {
        return qbt_dispatch1(gstk, [&](auto S_) {
            return render_batch_t<decltype(C_)::value, decltype(D_)::value, decltype(G_)::value, decltype(N_)::value,
                                  decltype(S_)::value>(this, s, bh, b, L, mask, flags);
        });
    });
}
