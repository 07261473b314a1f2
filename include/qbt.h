/*
 * qbt.h - C ABI of libqbt.so, the MI355X (gfx950) HIP layer of
 * qubatron-mi355x.
 *
 * It replaces the OpenGL connector of the reference (src/qubatron/octree_glc.c)
 * for the per-pixel sparse-octree ray traversal, shadow test and shading:
 *
 *   reference (octree_glc.c)                 this library
 *   octree_glc_init            :93-247       qbt_create
 *   octree_glc_upload_texbuffer_data :361-498 qbt_upload (same arguments)
 *   octree_glc_update          :249-352      qbt_update (same arguments) /
 *                                            qbt_render (explicit views+lights)
 *   (GL context teardown)                    qbt_destroy
 *   octree_trace_line (octree.c:341-537)     qbt_trace_rays (batched, on GPU)
 *   octree_insert_point loop (modelutil.c:203-217)
 *                                            qbt_build_octree (on GPU)
 *   particle_glc_update (particle_glc.c:118-156)
 *                                            qbt_particle_update
 *
 * Plain pointers and sizes only. All entry points return QBT_OK (0) or a
 * negative QBT_E* code (qbt_host.h); qbt_strerror names it. A context owns one
 * HIP stream and its device buffers; it is not thread-safe, calls on one
 * context must be serialised by the caller (the reference is single-threaded
 * on its GL thread, qubatron.c:678-679).
 */
#ifndef QBT_H
#define QBT_H

#include <stddef.h>
#include <stdint.h>

#include "qbt_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qbt_ctx qbt_ctx;

/* Buffer ids, in the order of octree_glc_buffer_t (octree_glc.c:16-24);
 * the two POINTS buffers are an extension feeding qbt_build_octree. */
enum {
    QBT_BUFFER_STATIC_COLOR   = 0,
    QBT_BUFFER_STATIC_NORMAL  = 1,
    QBT_BUFFER_STATIC_OCTREE  = 2,
    QBT_BUFFER_DYNAMIC_COLOR  = 3,
    QBT_BUFFER_DYNAMIC_NORMAL = 4,
    QBT_BUFFER_DYNAMIC_OCTREE = 5,
    QBT_BUFFER_STATIC_POINTS  = 6,
    QBT_BUFFER_DYNAMIC_POINTS = 7
};

/* Element types: the GL enum values the reference passes (GL_INT, GL_FLOAT). */
#define QBT_TYPE_INT 0x1404
#define QBT_TYPE_FLOAT 0x1406

typedef struct qbt_opts {
    int    device;           /* HIP device ordinal */
    size_t max_device_bytes; /* cap on the library's device allocations, 0 = none */
} qbt_opts;

/* octree_glc_init: picks the device, creates the stream. opts may be NULL. */
int         qbt_create(const qbt_opts* opts, qbt_ctx** out);
void        qbt_destroy(qbt_ctx* ctx);
const char* qbt_strerror(int code);
/* Device memory currently held by the context, bytes. */
size_t      qbt_device_bytes(const qbt_ctx* ctx);

/* octree_glc_upload_texbuffer_data (octree_glc.c:361-498), same arguments.
 * data: host array of `size` bytes; bytes [start, end) are copied into the
 * device buffer `buftype`. itemsize is 16 for an octree (one int4 texel, three
 * per 48-byte node) and 12 for float3 colours/normals/points; start, end and
 * size must be multiples of itemsize, start <= end <= size. When size exceeds
 * the buffer's capacity the buffer is reallocated with slack and the whole
 * array [0, size) is copied, as the reference does on texture regrowth
 * (octree_glc.c:412-428). size also sets the buffer's logical length. Copies
 * are complete when the call returns (data may be freed). */
int qbt_upload(qbt_ctx* ctx, const void* data, int type, size_t size, size_t itemsize, size_t start, size_t end,
               int buftype);

/* Copy bytes [start, end) of a buffer's host-layout image back to host
 * (octree buffers in the 48-byte node layout, float3 buffers as float3). */
int qbt_download(qbt_ctx* ctx, int buftype, void* host, size_t start, size_t end);

/* One camera view: uniforms of octree_fsh.c:9-15. */
typedef struct qbt_view {
    float camfp[3];    /* camera position */
    float angle[2];    /* yaw (angle_in.x), pitch (angle_in.y) */
    float width;       /* render dimensions (uniform `dimensions`) */
    float height;
    int   shoot;       /* gunshot light term */
    float basecube[4]; /* x_min, y_max, z_max, edge */
    int   maxlevel;    /* leaf depth */
} qbt_view;

/* Outputs (device frame of each view). */
#define QBT_OUT_HIT 0x001u    /* int32 static model index of the hit leaf, -1 none */
#define QBT_OUT_LEAF 0x002u   /* int32 static leaf node index, -1 none */
#define QBT_OUT_SHADOW 0x004u /* uint8 bit l = visible from light l */
#define QBT_OUT_RGBA8 0x008u  /* uint32 RGBA8, r in the low byte */
#define QBT_OUT_RGBAF 0x010u  /* float4 colour before clamping */
#define QBT_OUT_ISP 0x020u    /* float4 primary entry point (x, y, z, t) */
#define QBT_OUT_FLAGS 0x040u  /* uint8 QBT_PX_* */
#define QBT_OUT_DYN 0x080u    /* int32 dynamic model index (slot 8), -1 none */
#define QBT_OUT_EXP 0x100u    /* int32 node expansions (primary + shadow) */
#define QBT_OUT_DEFAULT (QBT_OUT_HIT | QBT_OUT_LEAF | QBT_OUT_SHADOW | QBT_OUT_RGBA8)

/* Render flags (high bits of the same word). */
#define QBT_RENDER_COUNT 0x10000u     /* count node expansions (stats + QBT_OUT_EXP) */
#define QBT_RENDER_SURROGATE 0x20000u /* shadow bits from the MLP outside its band */
#define QBT_RENDER_COLLECT 0x40000u   /* sample exact shadow results for training */

/* Pixel flags. */
#define QBT_PX_DISCARD 1u /* fragment discarded: colour (0,0,0,0) */
#define QBT_PX_LEAF 2u    /* primary trace reached a leaf */
#define QBT_PX_SHADED 4u  /* isp.t > 0: shadow traces and shading ran */
#define QBT_PX_SPRITE 8u  /* light sprite painted white */
#define QBT_PX_PREDICTED 16u /* surrogate decided the shadow bits */

#define QBT_MAX_LIGHTS_RENDER 8
#define QBT_MAX_VIEWS 256

/* Render nviews views with the same nlights point lights (lights: nlights x 3
 * floats) into the context's device frames (one per view index). Pixel
 * (px, py), py = 0 the bottom row, is stored at py * (int)width + px.
 * out_mask selects outputs (QBT_OUT_*) and flags (QBT_RENDER_*). Asynchronous
 * on the context stream; qbt_read_frame / qbt_get_stats synchronise. */
int qbt_render(qbt_ctx* ctx, const qbt_view* views, int nviews, const float* lights, int nlights,
               unsigned out_mask);

/* octree_glc_update (octree_glc.c:249-352), same arguments: derives the
 * light (420, 200 - sin(lighta)*20, 680 - sin(lighta)*200), basecube
 * {0, basesize, basesize, basesize} and dimensions width/(6 - quality/2),
 * then renders view 0 with QBT_OUT_RGBA8 (+ hit, leaf, shadow). */
int qbt_update(qbt_ctx* ctx, float width, float height, const float position[3], const float angle[3],
               float lighta, uint8_t quality, int maxlevel, float basesize, int shoot);

/* Host copies of one view's frame; NULL members are skipped. Sizes are
 * w*h elements (w = (int)width, h = (int)height of that view). */
typedef struct qbt_frame {
    int       w, h;
    int32_t*  hit;
    int32_t*  leaf;
    uint8_t*  shadow;
    uint32_t* rgba8;
    float*    rgbaf; /* 4 per pixel */
    float*    isp;   /* 4 per pixel */
    uint8_t*  flags;
    int32_t*  dyn;
    int32_t*  exp;
} qbt_frame;

int qbt_read_frame(qbt_ctx* ctx, int view, qbt_frame* out);
/* Device pointers of one view's frame (valid until the next render with a
 * larger frame); members not rendered are NULL. */
int qbt_frame_device(qbt_ctx* ctx, int view, qbt_frame* out);

/* Size of the static device array behind qbt_render_noalloc, per GPU. */
#define QBT_NOALLOC_SCRATCH_BYTES 8704

/* qbt_render without any allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). It renders like
 * qbt_render with tuning path 0 (the fused per-pixel kernel, direct node
 * layout; same results as every path) into caller-owned DEVICE frames:
 * frames[i] holds device pointers for view i with w = (int)width,
 * h = (int)height, one for every output out_mask selects (rgbaf and isp
 * 16-byte aligned); qbt_read_frame does not see these frames. Everything
 * else the kernels need (counters, sprite flags, the node-array check word)
 * lives in one static device array of QBT_NOALLOC_SCRATCH_BYTES on each GPU,
 * held until the process ends and shared by every call on that GPU: calls
 * that overlap on one GPU race. nviews above QBT_MAX_VIEWS: QBT_ERANGE;
 * QBT_RENDER_SURROGATE / QBT_RENDER_COLLECT: QBT_ENOTSUP. The first call
 * after an octree upload checks the node array (one synchronisation).
 * Asynchronous on the context stream; qbt_get_stats synchronises and reports
 * the call's counters. */
int qbt_render_noalloc(qbt_ctx* ctx, const qbt_view* views, int nviews, const float* lights, int nlights,
                       unsigned out_mask, const qbt_frame* frames);

typedef struct qbt_stats {
    double   kernel_ms;      /* device time of the last qbt_render / qbt_trace_rays */
    double   trace_ms;       /* time of the traversal kernel(s) only */
    uint64_t primary_rays;   /* pixels */
    uint64_t shadow_rays;    /* shadow traces performed */
    uint64_t hits;           /* pixels with isp.t > 0 */
    uint64_t expansions;     /* node expansions (QBT_RENDER_COUNT only) */
    uint64_t predicted;      /* shadow queries answered by the surrogate */
    uint32_t launches;       /* kernel launches of the last call */
    uint32_t grid_exact;     /* the basecube grid was exact in fp32 */
    uint32_t dual;           /* the dynamic tree took part */
    uint32_t compact;        /* the compact node layout was traversed */
    uint64_t expansions_primary; /* of which by primary rays (QBT_RENDER_COUNT only) */
    uint64_t shadow_deferred;    /* of shadow_rays, traced by their pixel's camera lane (tuning merge) */
} qbt_stats;

int qbt_get_stats(qbt_ctx* ctx, qbt_stats* out);

/* The compact node layout of tree (QBT_BUFFER_STATIC_OCTREE or
 * QBT_BUFFER_DYNAMIC_OCTREE): how many times it was rebuilt from the node
 * array (relayouts) and patched in place after range uploads (patches) since
 * the context was created, and its record count (patches append records).
 * The layout is brought up to date at the next render; qbt_build_octree
 * writes it directly (neither count changes). Any pointer may be NULL. */
int qbt_layout_stats(qbt_ctx* ctx, int tree, uint64_t* relayouts, uint64_t* patches, uint64_t* records);

/* Wait for all work queued on the context stream. */
int qbt_synchronize(qbt_ctx* ctx);

/* Kernel timing: between qbt_profile_begin and qbt_profile_end every launch
 * of the traversal kernel (render or ray batch) is bracketed by a HIP event
 * pair on the context stream (up to 4096 launches). qbt_profile_end
 * synchronises and reports launches and their durations in milliseconds. */
#define QBT_PROF_PRIMARY 0   /* fused render kernel, primary kernel, ray batches */
#define QBT_PROF_SHADOW 1    /* shadow kernel of the wavefront path (all lights) */
#define QBT_PROF_SURROGATE 2 /* surrogate decision kernel */
#define QBT_PROF_FRAME 3     /* camera and shadow rays in one launch (tuning merge) */
typedef struct qbt_profile {
    uint32_t launches;
    uint32_t reserved;
    double   total_ms;
    double   avg_ms;
    double   min_ms;
    double   max_ms;
    uint32_t kind_launches[4]; /* per QBT_PROF_* */
    double   kind_ms[4];       /* summed duration per QBT_PROF_* */
} qbt_profile;

int qbt_profile_begin(qbt_ctx* ctx);
int qbt_profile_end(qbt_ctx* ctx, qbt_profile* out);

/* Launch parameters of the render path (what the autotuner chooses).
 * path: 0 = one fused kernel per view (every lane runs its pixel's primary
 * and shadow rays back to back), 1 = wavefront stages with persistent,
 * lane-refilling traversal kernels over a batch of views. block_threads:
 * 64..256 (0 = default 256; halved while the LDS stack exceeds 64 KiB).
 * refill: idle lanes that make a wave claim new rays, 1..64 (0 = 24 for
 * camera rays and 16 for shadow rays): the rays a wave takes per refill. layout: 0 = traverse the compact
 * breadth-first node layout (8 bytes per node, children contiguous) whenever
 * the node array is a proper tree, 1 = always the direct 32-byte child
 * blocks. stack: 0 = per-lane traversal stacks in LDS, 1 = in global memory.
 * reorder: 1 = shadow rays in direction-octant order per light. graph: 1 =
 * capture the batch's launches into a hipGraph and replay it while the
 * launch configuration is unchanged. strip: order of the 8x8 pixel tiles,
 * 0 = row-major, S (1..1024) = vertical strips S tiles wide, each row-major.
 * merge: whether an exact render (no surrogate, no sample collection,
 * reorder 0) traces its camera and shadow rays in one persistent launch, the
 * shadow rays of finished pixels taken by waves whose camera work has run
 * dry, instead of a camera launch, then a shadow launch: 0 = never, 1 =
 * where it pays (default: two or more lights and 1 to 16 camera rays per lane
 * of the full persistent grid), 2 = always. Results never depend on these. */
typedef struct qbt_tuning {
    int path;
    int block_threads;
    int refill;
    int layout;
    int stack;
    int reorder;
    int graph;
    int strip;
    int merge; /* 1: camera and shadow rays of an exact render in one launch (k_frame) */
} qbt_tuning;

int qbt_get_tuning(qbt_ctx* ctx, qbt_tuning* out);
int qbt_set_tuning(qbt_ctx* ctx, const qbt_tuning* in);

/* Autotuner: times qbt_render of the given views and lights (QBT_RENDER_*
 * counting/collection flags are dropped) under candidate tunings of the
 * wavefront path - coordinate descent over block_threads {64, 128, 256},
 * refill {defaults, 8, 16, 24, 32, 48}, stack {LDS, global}, reorder {0, 1}
 * and merge {0, 1, 2} (merge takes effect with reorder 0), two
 * sweeps, a candidate kept only if it is at least 1% faster - and leaves the
 * fastest set on the context. results (up to max_results) receive every
 * timed candidate, *nresults how many were timed. reps: timed renders per
 * candidate (after one untimed render). Outputs of the last render are
 * those of the last candidate timed. */
typedef struct qbt_tune_result {
    qbt_tuning tuning;
    double     ms; /* mean time of one qbt_render call */
} qbt_tune_result;

int qbt_autotune(qbt_ctx* ctx, const qbt_view* views, int nviews, const float* lights, int nlights,
                 unsigned mask, int reps, qbt_tune_result* results, int max_results, int* nresults);

/* Diagnostic: copy `bytes` (a multiple of 16) between two scratch buffers
 * with a plain streaming kernel (k_copy16) `reps` times and report the
 * average launch time and (read + write) GB/s. Used to calibrate hardware
 * byte counters and as the achievable-bandwidth reference. */
int qbt_copy_bandwidth(qbt_ctx* ctx, size_t bytes, int reps, double* avg_ms, double* gbps);

/* Batched octree_trace_line: n rays (origins, dirs: n x 3 floats, host).
 * semantics 0 = shader (octree_fsh.c cube_trace_line, dual tree), 1 = CPU
 * twin (octree.c:341-537: static tree only, its plane-test returns), 2 = the
 * particle shader's copy (particle_vsh.c:121-343: static tree only, (0,0,0,0)
 * plane-test returns, its slot address, see qbt_particle_update). Outputs
 * (host, any may be NULL): status (1 hit, 0 miss, -1 no entry),
 * model (slot 8 of the static leaf, -1 none), leaf node, isp (4 floats),
 * tlf (leaf cube, 4 floats). basecube/maxlevel as in qbt_view. */
int qbt_trace_rays(qbt_ctx* ctx, const float* origins, const float* dirs, size_t n, int semantics,
                   const float basecube[4], int maxlevel, int32_t* status, int32_t* model, int32_t* leaf,
                   float* isp, float* tlf);

/* The reference's particle pass: particle_glc_alloc_in (particle_glc.c:158-167)
 * + particle_glc_update (:118-156) + the read-back into pos_out/spd_out
 * (:147-150). n particles, position and velocity (n x 3 floats each, host)
 * in, the updated ones out (host; may alias the inputs). Each particle runs
 * the particle shader's main (shaders/particle_vsh.c:345-373): a particle
 * with spd.x > -90000 gets spd.y -= 0.4, traces the static octree from pos
 * along spd with that shader's copy of the traversal (:62-343: one tree,
 * (0,0,0,0) plane-test defaults, its slot address, see trace.hpp), stops at
 * the leaf corner if it hit a leaf closer than 10 units (spd.x = -100000),
 * else moves by spd and stops below y = -10. basecube = (0, basesize,
 * basesize, basesize) as particle_glc.c:130. */
int qbt_particle_update(qbt_ctx* ctx, const float* pos, const float* spd, size_t n, int maxlevel, float basesize,
                        float* pos_out, float* spd_out);

/* Build the static or dynamic octree (tree = QBT_BUFFER_STATIC_OCTREE or
 * QBT_BUFFER_DYNAMIC_OCTREE) on the GPU from the first npoints points of the
 * matching POINTS buffer, giving the node array that inserting the points in
 * index order with octree_insert_point (octree.c:95-147) produces, node for
 * node. *nodes_out receives the node count (the reference's len). A
 * dynamic-tree build keeps up to 2 GB of its scratch memory for the next
 * build (a per-frame rebuild then allocates nothing); a static build releases
 * its scratch when it ends. Kept scratch counts against max_device_bytes; it
 * is released when another allocation of the context would exceed that cap
 * or find the device full, and by qbt_destroy. */
int qbt_build_octree(qbt_ctx* ctx, int tree, const float basecube[4], int levels, size_t npoints,
                     size_t* nodes_out);

/* ---------------- learned shadow surrogate (no reference counterpart) ----
 * A 42 -> 64 -> 64 -> 64 -> 1 MLP (ReLU, sigmoid; bf16 weights and
 * activations on the matrix cores, fp32 accumulation and biases) predicting
 * the shadow bit of a (hit point, normal, light) query. Inputs: the hit
 * position scaled to the base cube, u = (p - (x, y - w, z - w)) / w, as
 * sin/cos of 2^k pi u for k = 0..5 (36 values: index 6k + 2a + {0 sin,
 * 1 cos} for axis a), the normal (3), the light position scaled the same
 * way (3). Parameters: flat fp32, row-major W[out][in]: W1 64x42, b1, W2
 * 64x64, b2, W3 64x64, b3, W4 1x64, b4.
 *
 * qbt_render with QBT_RENDER_SURROGATE (wavefront path): primary rays are
 * traced exactly; each (shaded pixel, light) query whose predicted
 * probability is below band_lo or above band_hi takes the predicted bit
 * (QBT_PX_PREDICTED set), the others are traced exactly. stats.predicted
 * counts the predicted queries. With QBT_RENDER_COLLECT an exact render
 * appends every stride-th (shaded pixel, light) query with its exact bit to
 * the training store (capacity from qbt_surrogate_reserve). */
#define QBT_MLP_INPUTS 42
#define QBT_MLP_PARAMS 11137

/* He-uniform weights from seed, zero biases, Adam state reset. */
int qbt_surrogate_init(qbt_ctx* ctx, uint64_t seed);
int qbt_surrogate_set_params(qbt_ctx* ctx, const float* params, size_t n);
int qbt_surrogate_get_params(qbt_ctx* ctx, float* params, size_t n);
/* Uncertainty band, 0 < lo <= hi < 1 (default 0.1, 0.9). */
int qbt_surrogate_set_band(qbt_ctx* ctx, float lo, float hi);
/* (Re)allocate the training store for nsamples queries and empty it;
 * collecting renders keep every stride-th query (stride >= 1). */
int qbt_surrogate_reserve(qbt_ctx* ctx, size_t nsamples, uint32_t stride);
/* Samples held (saturates at the capacity); lit = how many are labelled 1. */
int qbt_surrogate_samples(qbt_ctx* ctx, size_t* n, size_t* lit);
/* Copy samples [start, start + n) of the store: 10 floats each (hit point
 * xyz, normal xyz, light xyz, label). */
int qbt_surrogate_read_samples(qbt_ctx* ctx, size_t start, size_t n, float* out);
/* Probabilities of n feature rows (x: n x 42 host floats). */
int qbt_surrogate_forward(qbt_ctx* ctx, const float* x, size_t n, float* p);
/* Probabilities of n queries from raw inputs (host, n x 3 floats each), the
 * feature computation and fused inference kernel the render path uses. */
int qbt_surrogate_predict(qbt_ctx* ctx, const float* isp, const float* nrm, const float* light, size_t n,
                          const float basecube[4], float* p);
/* Mean binary cross-entropy of n labelled rows and its gradient (flat
 * parameter order; grads may be NULL). */
int qbt_surrogate_grad(qbt_ctx* ctx, const float* x, const float* y, size_t n, float* grads, float* loss);
/* qbt_surrogate_grad followed by one Adam step (beta1 0.9, beta2 0.999,
 * eps 1e-8, bias-corrected) with learning rate lr. */
int qbt_surrogate_step(qbt_ctx* ctx, const float* x, const float* y, size_t n, float lr, float* loss);
/* steps Adam steps on minibatches (batch: multiple of 16) drawn without
 * replacement from a seeded shuffle of the store; loss of the last batch. */
int qbt_surrogate_train(qbt_ctx* ctx, const float basecube[4], int steps, int batch, float lr, uint64_t seed,
                        float* loss);

/* Library version string. */
const char* qbt_version(void);
/* Text of the last HIP runtime error seen by the library (QBT_EDEVICE). */
const char* qbt_last_hip_error(void);

#ifdef __cplusplus
}
#endif
#endif
