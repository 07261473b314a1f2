/*
 * qbt_internal.hpp - context and device-side structures shared by the
 * translation units of libqbt.so (not part of the C ABI).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "qbt.h"
#include "trace.hpp"

#define QBT_OUT_ALL 0x1FFu

enum { CNT_SHADOW = 0, CNT_HITS, CNT_EXP, CNT_ABORT, CNT_PRED, CNT_QUEUE, CNT_EXP_PRIM, CNT_DEFER, CNT_N };
enum { FR_HIT = 0, FR_LEAF, FR_SHADOW, FR_RGBA8, FR_RGBAF, FR_ISP, FR_FLAGS, FR_DYN, FR_EXP, FR_N };

/* per-(light, pixel) shadow status of the wavefront path */
enum : uint8_t {
    LS_OCCLUDED = 0, /* exact trace: occluded */
    LS_LIT      = 1, /* exact trace: visible */
    LS_DISCARD  = 2, /* exact trace rejected at the root: fragment discarded */
    LS_SKIPPED  = 3, /* an earlier light discarded the fragment */
    LS_PRED_OCC = 4, /* surrogate: occluded */
    LS_PRED_LIT = 5, /* surrogate: visible */
    LS_UNSURE   = 6  /* surrogate: inside the band, to be traced exactly */
};

/* Item g of a list of node ranges [ranges[r].x, ranges[r].y) numbered
 * consecutively (roff[r] = items before range r): the node
 * ranges[r].x + g - roff[r] for the last r with roff[r] <= g */
__device__ __forceinline__ size_t range_item(const uint2* __restrict__ ranges, const uint32_t* __restrict__ roff,
                                             int nr, uint32_t g)
{
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (roff[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return (size_t) ranges[lo].x + (g - roff[lo]);
}

struct DevBuf {
    void*  p   = nullptr;
    size_t cap = 0; /* bytes */
};

/* breadth-first compact layout of a tree (layout.hip) */
struct CompactTree {
    DevBuf rec, model, ref; /* cap records each */
    DevBuf r2c;             /* per reference node: its record, -1 none (ref_n valid, r2c_cap allocated) */
    DevBuf scratch;         /* relayout and patch work buffers, kept between calls */
    size_t n       = 0;     /* records in use (in-place patches append) */
    size_t cap     = 0;
    size_t ref_n   = 0;     /* reference nodes the layout covers */
    size_t r2c_cap = 0;
    bool   ok      = false; /* built and the node array is a proper tree */
};

struct TreeBuf {
    DevBuf child, model;
    size_t n         = 0; /* logical node count (size/48 of the last upload) */
    size_t cap_nodes = 0;
    bool   validated = true;
    bool   full_check = false;
    size_t dirty_lo = 0, dirty_hi = 0; /* bounding range of the node ranges written since the last check */
    /* the written ranges themselves (up to kValRanges; vdirty_over: more) */
    std::vector<std::pair<size_t, size_t>> vdirty;
    bool                                   vdirty_over = false;
    CompactTree cmp;
    bool        cmp_stale = true; /* node array replaced since the last layout: relayout */
    /* node ranges written by range uploads since the last layout: patched in
     * place (layout.hip), or a relayout when the patch cannot represent them */
    std::vector<std::pair<size_t, size_t>> patch;
    uint64_t relayouts = 0, patches = 0;
    /* small range uploads, held on the host until the tree is next used:
     * texel index -> its 16 bytes (a later upload of a texel replaces it) */
    std::vector<uint64_t>                   pend_idx;
    std::vector<int4>                       pend_val;
    std::unordered_map<uint64_t, uint32_t>  pend_pos;
};

struct AttrBuf {
    DevBuf ca;
    size_t n = 0, n_col = 0, n_nrm = 0, cap = 0;
};

struct PointsBuf {
    DevBuf xyz;
    size_t n = 0, cap = 0;
};

/* per-view uniforms of the pixel program */
struct ViewDev {
    float cam[3];
    float qz[4], qx[4];
    float W, H;
    int   w, h;
    int   shoot;
    int   nlights;
    float light[QBT_MAX_LIGHTS_RENDER][3];
    float cl[QBT_MAX_LIGHTS_RENDER][3]; /* normalize(light - camfp) */
    float cos_thr;
};

/* the light positions of a render call (shared by all its views) */
struct LightsDev {
    float p[QBT_MAX_LIGHTS_RENDER][3];
    int   n;
};

struct FrameDev {
    int*      hit;
    int*      leaf;
    uint8_t*  shadow;
    uint32_t* rgba8;
    float4*   rgbaf;
    float4*   isp;
    uint8_t*  flags;
    int*      dyn;
    int*      exp;
};

/* A batch of views rendered by one set of launches (device tables).
 * Work items: 64 per 8x8 pixel tile, views in order, item_off[v] the first
 * item of view v; pixels: view v's pixel (px, py) at pix_off[v] + py*w + px. */
struct BatchDev {
    const ViewDev*  views;
    const uint32_t* item_off; /* n + 1 */
    const uint32_t* pix_off;  /* n + 1 */
    const FrameDev* frames;
    const int*      sprite; /* 8 per view: bit 0 discard, bit 1 paint white */
    int             n;
};

/* view of a batch item / pixel: the largest v with off[v] <= x */
__device__ __forceinline__ int batch_view(const uint32_t* __restrict__ off, int n, uint32_t x)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

/* host side of a batch: the views [first, first + n) of a qbt_render call */
struct BatchHost {
    std::vector<ViewDev>  views;
    std::vector<uint32_t> item_off, pix_off;
    std::vector<FrameDev> frames;
    uint32_t              nitems = 0, npix = 0;
    int                   first = 0;
};

struct Frame {
    DevBuf   buf[FR_N];
    int      w = 0, h = 0;
    unsigned mask = 0;
};

struct MlpState; /* surrogate.hip */
struct BuildState; /* build.hip */

struct qbt_ctx {
    int          device    = 0;
    size_t       max_bytes = 0;
    size_t       held      = 0;
    hipStream_t  stream    = nullptr;
    hipEvent_t   ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    DevBuf       stage, counters, sprite, rays, valflag, valrg;
    TreeBuf      tree[2];
    AttrBuf      attr[2];
    PointsBuf    pts[2];
    std::vector<Frame> frames;
    qbt_stats    last{};
    bool         pending_stats   = false;
    bool         stats_has_trace = false;
    uint64_t     aborted         = 0;
    int          iter_cap        = 1 << 22;
    int          tune_block      = 0;
    int          tune_refill     = 0;  /* idle lanes that trigger a wave refill (1..64; 0 = 24 camera, 16 shadow) */
    int          path            = 1;  /* 0 = fused per-pixel kernel, 1 = wavefront stages */
    int          num_cus         = 256;
    /* wavefront render buffers */
    DevBuf       wb_isp, wb_ms, wb_md, wb_leaf, wb_flags, wb_disc, wb_lstat, wb_exp, wb_cand, wb_items, wb_q;
    DevBuf       wb_lists; /* per-light shadow lists in direction-octant order */
    DevBuf       wb_fb;    /* surrogate: per-light lists of the rays traced exactly */
    DevBuf       wb_gstack; /* traversal stacks in global memory (tuning stack = 1) */
    DevBuf       wb_bcnt;   /* per-block counts of the stream compactions */
    DevBuf       wb_lexp;   /* counting renders: shadow expansions per light and pixel */
    DevBuf       wb_slots;  /* k_frame: shadow hand-over slots, 16 B per (light, camera work item) */
    DevBuf       bt_tab; /* batch tables: views, item and pixel offsets, frames */
    static constexpr int kTabSlots = 4;
    void*        tab_host[kTabSlots] = {};      /* pinned staging of the batch tables */
    hipEvent_t   tab_ev[kTabSlots]   = {};      /* the last copy out of each slot */
    int          tab_slot            = 0;
    int          tune_stack   = 0; /* 0 = LDS, 1 = global memory */
    int          tune_reorder = 0; /* shadow rays in direction-octant order (measured: no gain) */
    int          tune_graph   = 1; /* replay batched renders from a captured hipGraph */
    int          tune_strip   = 0; /* tile order: 0 row-major, S > 0 vertical strips of S tiles */
    int          tune_merge   = 1; /* exact renders: camera and shadow rays in one launch (k_frame): 0 never, 1 where it pays, 2 always */
    /* captured render graph and the launch configuration it was captured for */
    hipGraph_t           graph      = nullptr;
    hipGraphExec_t       graph_exec = nullptr;
    std::vector<uint8_t> graph_key;
    uint64_t             graph_replays = 0;
    MlpState*    mlp             = nullptr;
    BuildState*  build           = nullptr;
    /* qbt_profile_begin/end: event pairs around traversal-kernel launches */
    bool                    profiling = false;
    std::vector<hipEvent_t> prof_ev;
    std::vector<uint8_t>    prof_kind;
    uint32_t                prof_n = 0;

    void prof_mark_begin(int kind)
    {
        if (profiling && prof_n < 4096 && ensure_prof_events()) {
            if (prof_kind.size() <= prof_n) prof_kind.resize(prof_n + 1);
            prof_kind[prof_n] = (uint8_t) (kind & 3);
            (void) hipEventRecord(prof_ev[2 * prof_n], stream);
        }
    }
    void prof_mark_end()
    {
        if (profiling && prof_n < 4096 && prof_ev.size() >= 2 * (size_t) (prof_n + 1))
            (void) hipEventRecord(prof_ev[2 * prof_n + 1], stream), prof_n++;
    }
    bool ensure_prof_events()
    {
        while (prof_ev.size() < 2 * (size_t) (prof_n + 1)) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return false;
            prof_ev.push_back(e);
        }
        return true;
    }

    /* qbt_render_noalloc: the device address of this device's static
     * scratch array (qbt_core.hip), and where qbt_get_stats reads the
     * counters of the last call (nullptr: the context's own) */
    void*        scratch = nullptr;
    const void*  cnt_src = nullptr;

    int      dev_alloc(DevBuf& b, size_t bytes);
    void     dev_free(DevBuf& b);
    /* flag: a device word for the check's result (nullptr: the context's own,
     * allocated on first use) */
    int      validate_tree(int t, unsigned* flag = nullptr);
    /* apply the held small uploads of tree t (discard: drop them) */
    int      flush_pending(int t, bool discard = false);
    int      relayout(int t);
    /* apply the range uploads since the last layout to the compact layout of
     * tree t in place; *done = false when they change what the patch cannot
     * represent (then the caller relayouts) */
    int      patch_layout(int t, bool* done);
    /* compact arrays of tree t for nrec records and nref reference nodes */
    int      compact_reserve(int t, size_t nrec, size_t nref);
    int      scene_for(const float basecube[4], int maxlevel, qbt::SceneDev& s, bool& dual, bool& grid, bool& cmp,
                       bool allow_compact, unsigned* valflag_dev = nullptr);
    bool     use_compact = true; /* traverse the compact layout when both trees have one */
    int      ensure_frame(int view, int w, int h, unsigned mask);
    FrameDev frame_dev(int view);
    int      block_threads(int L, bool dual) const;
    int      launch_render(const qbt::SceneDev& s, bool dual, bool grid, bool count, const ViewDev& v, const int* spr,
                           FrameDev f, unsigned mask);
    int      launch_sprite(const qbt::SceneDev& s, bool cmp, bool dual, bool grid, const BatchDev& b,
                           hipStream_t st);
    int      ensure_wave_bufs(size_t npix, int nlights);
    /* flags: QBT_RENDER_SURROGATE, QBT_RENDER_COLLECT */
    int      render_batch_wavefront(const qbt::SceneDev& s, bool cmp, bool dual, bool grid, bool count,
                                    const BatchHost& bh, const BatchDev& b, const LightsDev& L, unsigned mask,
                                    unsigned flags);
    void     graph_free();
    void     mlp_free();
    void     builder_free();
    /* release the builder blocks kept between builds (none while a build
     * runs); true if any were held */
    bool     builder_trim();
};

/* surrogate.hip: the MLP decision pass of a batch. The shaded pixels of
 * light l are lists[l * list_stride + i], i < *nsel (batch pixel indices).
 * For every (pixel, light) query it writes LS_PRED_* where the prediction is
 * outside the band and LS_UNSURE otherwise (to be traced exactly). */
int qbt_surrogate_decide(qbt_ctx* c, const qbt::SceneDev& s, bool dual, const LightsDev& L, const uint32_t* lists,
                         size_t list_stride, const unsigned* nsel, size_t npix);
/* surrogate.hip: append training samples (every stride-th shaded pixel x
 * light of the batch, exact label) to the sample store */
int qbt_collect_batch(qbt_ctx* c, const qbt::SceneDev& s, bool dual, const LightsDev& L, const uint32_t* lists,
                      size_t list_stride, const unsigned* nsel, size_t npix);
bool qbt_surrogate_ready(const qbt_ctx* c);
void qbt_surrogate_band(const qbt_ctx* c, float* lo, float* hi);

/* Stream compaction without same-address atomics (render.hip): the indices
 * i < n for which pred(i, &value) holds, their values written to out in index
 * order, the count to *count. Blocks of kSelItems items: per-block counts, one
 * scan block, scatter. bcnt: (n / kSelItems + 1) * rows words of scratch;
 * rows = gridDim.y of the launches (one independent compaction per row, the
 * predicate gets the row). */
constexpr int kSelItems = 1024;

void   qbt_set_hip_error(hipError_t e, const char* what, const char* file, int line);
bool   qbt_grid_exact(const float b[4], int L);
size_t qbt_stack_bytes(int L, bool dual, int threads);
void   qbt_make_view(const qbt_view& q, const float* lights, int nlights, ViewDev& v);

/* Runtime bools -> template arguments: f receives std::integral_constant<bool>
 * values, one per flag. */
template <class F>
inline int qbt_dispatch1(bool a, F&& f)
{
    return a ? f(std::true_type{}) : f(std::false_type{});
}
template <class F>
inline int qbt_dispatch2(bool a, bool b, F&& f)
{
    return qbt_dispatch1(a, [&](auto A) { return qbt_dispatch1(b, [&](auto B) { return f(A, B); }); });
}
template <class F>
inline int qbt_dispatch3(bool a, bool b, bool c, F&& f)
{
    return qbt_dispatch1(a, [&](auto A) { return qbt_dispatch2(b, c, [&](auto B, auto C) { return f(A, B, C); }); });
}
template <class F>
inline int qbt_dispatch4(bool a, bool b, bool c, bool d, F&& f)
{
    return qbt_dispatch1(a, [&](auto A) {
        return qbt_dispatch3(b, c, d, [&](auto B, auto C, auto D) { return f(A, B, C, D); });
    });
}

#define QBT_HIPCHK(x)                                    \
    do {                                                   \
        hipError_t e_ = (x);                               \
        if (e_ != hipSuccess) {                            \
            qbt_set_hip_error(e_, #x, __FILE__, __LINE__); \
            return QBT_EDEVICE;                            \
        }                                                  \
    } while (0)
