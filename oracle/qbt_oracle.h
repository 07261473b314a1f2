/*
 * qbt_oracle.h - CPU restatement of Qubatron's octree ray traversal, shadow
 * test and shading, used ONLY as the parity reference of the HIP path.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load liboracle.so; the product library never links it.
 *
 * What it restates (reference paths relative to the milgra/qubatron tree):
 *   orc_trace      src/qubatron/shaders/octree_fsh.c:62-99 (plane tests),
 *                  :127-136 (slot addressing), :138-379 (cube_trace_line,
 *                  static + dynamic tree in lockstep);
 *                  with ORC_SEM_CPU: src/qubatron/octree.c:293-537
 *                  (octree_trace_line, single tree, CPU plane-test returns);
 *                  with ORC_SEM_PARTICLE: src/qubatron/shaders/
 *                  particle_vsh.c:62-343 (single tree, (0,0,0,0) plane-test
 *                  returns, its slot address :110-119, return for discard).
 *   orc_particle_step  particle_vsh.c:345-373 (gravity step, trace, stop).
 *   orc_render     src/qubatron/shaders/octree_fsh.c:381-464 (camera
 *                  quaternions, primary trace, per-light shadow trace,
 *                  shading, light sprite), pixel centres of
 *                  src/qubatron/shaders/octree_vsh.c:8-12.
 *   orc_mlp_*      the shadow-visibility MLP (no reference counterpart):
 *                  fp32 forward/backward with bf16 rounding points, Adam.
 *
 * Pinned definitions where GLSL is implementation-defined: tan(pi/4) = 1;
 * normalize(v) = v * (1/sqrtf(dot(v,v))); dot = (x*x' + y*y') + z*z';
 * acos(d) < 0.02 is d > cosf(0.02f); quaternion sin/cos are sinf/cosf of the
 * host; pixel centre (px+0.5, py+0.5) with py = 0 the bottom row.
 * Built with -ffp-contract=off: no fused multiply-add anywhere.
 *
 * Parity pins (see tests/): brute-force nearest-leaf search on small trees,
 * axis-aligned closed forms, invariants (hit leaf contains isp, entry t
 * monotone), the 5-point OCTTEST scene, and the single-tree reduction of the
 * dual-tree traversal.
 */
#ifndef QBT_ORACLE_H
#define QBT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_SEM_SHADER 0 /* octree_fsh.c semantics (the GPU hot path) */
#define ORC_SEM_CPU 1    /* octree.c octree_trace_line semantics */
#define ORC_SEM_PARTICLE 2 /* particle_vsh.c semantics */

#define ORC_HIT 1
#define ORC_MISS 0
#define ORC_DISCARD (-1) /* root entry rejected: the shader discards the fragment */
#define ORC_ERR (-2)     /* tree references a node out of range / stack too deep */

#define ORC_MAXLEVEL 23

typedef struct orc_tree {
    const int32_t* octs; /* len nodes x 12 int32; NULL = root-only empty tree */
    size_t         len;
} orc_tree;

typedef struct orc_scene {
    orc_tree     stat, dyn;
    const float *col_s, *nrm_s; /* float3 per point */
    size_t       npts_s;
    const float *col_d, *nrm_d;
    size_t       npts_d;
    float        basecube[4];
    int          maxlevel;
} orc_scene;

typedef struct orc_trace_out {
    int     status;
    float   isp[4];   /* entry point of the hit leaf (xyz, t) */
    float   tlf[4];   /* hit leaf cube */
    int32_t socti;    /* static leaf node (0 if only dynamic) */
    int32_t docti;    /* dynamic leaf node */
    int32_t ms, md;   /* slot 8 of the static / dynamic leaf */
    float   col[4];   /* leaf colour (rgb, 1) */
    float   nrm[4];   /* leaf normal (xyz, 1) */
    int64_t expansions; /* node expansions (child-block reads), static tree */
    int64_t expansions_dyn;
    int64_t iterations;
} orc_trace_out;

int orc_trace(const orc_scene* sc, const float pos[3], const float dir[3], int semantics, orc_trace_out* out);

/* One particle of the particle shader's transform-feedback pass: pos/spd in,
 * pos_out/spd_out out (particle_vsh.c:345-373). Returns 0, or ORC_ERR if the
 * trace hit an out-of-range node. length() is sqrtf(dot) with dot as above. */
int orc_particle_step(const orc_scene* sc, const float pos[3], const float spd[3], float pos_out[3],
                      float spd_out[3]);

typedef struct orc_view {
    float camfp[3];
    float angle[2]; /* angle_in.x (yaw), angle_in.y (pitch) */
    float W, H;     /* render dimensions (uniform `dimensions`) */
    int   shoot;
} orc_view;

/* pixel flags */
#define ORC_PX_DISCARD 1u /* fragment discarded: colour stays (0,0,0,0) */
#define ORC_PX_LEAF 2u    /* primary trace reached a leaf */
#define ORC_PX_SHADED 4u  /* isp.t > 0: shadow traces + shading ran */
#define ORC_PX_SPRITE 8u  /* light sprite painted white */

typedef struct orc_frame {
    int       w, h;       /* pixel grid (int)W x (int)H */
    int32_t*  hit;        /* static model index, -1 when no leaf */
    int32_t*  leaf;       /* static leaf node index, -1 when no leaf */
    int32_t*  dyn;        /* dynamic model index (slot 8), -1 when no leaf */
    uint8_t*  shadow;     /* bit l = light l visible */
    uint8_t*  flags;      /* ORC_PX_* */
    float*    rgba;       /* final colour before clamping, 4 floats */
    uint32_t* rgba8;      /* RGBA8 packed (r in the low byte) */
    float*    isp;        /* primary isp, 4 floats */
    int32_t*  expansions; /* primary + shadow expansions per pixel */
} orc_frame;

/* Render rows [row0, row1) of the frame (py = 0 is the bottom row). Any
 * output pointer may be NULL. Returns 0, or ORC_ERR if a trace hit an
 * invalid tree. nthreads > 1 splits rows over pthreads. */
int orc_render(const orc_scene* sc, const orc_view* v, int nlights, const float* lights, orc_frame* fr, int row0,
               int row1, int nthreads);

/* The per-view camera quaternions (qz, vx, qx) of octree_fsh.c:406-408. */
void orc_camera(const float angle[2], float qz[4], float vx[3], float qx[4]);
/* Ray direction (csv) of pixel (px, py). */
void orc_pixel_dir(const orc_view* v, const float qz[4], const float qx[4], int px, int py, float csv[3]);

/* RGBA8 conversion used for the frame (clamp to [0,1], *255, round half up;
 * NaN -> 0). */
uint32_t orc_pack_rgba8(const float c[4]);

/* ---------------- shadow-visibility MLP (no reference counterpart) -------- */

#define ORC_MLP_IN 42
#define ORC_MLP_H 64

/* Feature vector of one shadow query: 6-band sin/cos of the hit position
 * scaled to [0,1] by the cube (36), unit normal (3), light position scaled
 * to [0,1] (3). */
void  orc_mlp_features(const float isp[3], const float nrm[3], const float light[3], const float basecube[4],
                       float out[ORC_MLP_IN]);
float orc_bf16_round(float x); /* round to nearest even bf16, back to fp32 */

/* Parameters, fp32 master copies, row-major W[out][in]:
 * W1 64x42, b1 64, W2 64x64, b2 64, W3 64x64, b3 64, W4 1x64, b4 1. */
#define ORC_MLP_NPARAM (64 * 42 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1)

/* Forward of n rows (x: n x 42). bf16 = 1: inputs, weights and post-ReLU
 * activations rounded to bf16 (round to nearest even), fp32 accumulation and
 * fp32 biases, as the GPU kernels compute; bf16 = 0: plain fp32 (for
 * finite-difference checks). Writes sigmoid outputs p (n); logits (n) if
 * not NULL. */
void orc_mlp_forward(const float* params, const float* x, int n, float* p, float* logits, int bf16);
/* Mean binary cross-entropy over n rows with labels y in {0,1}: writes the
 * gradient (ORC_MLP_NPARAM) and returns the mean loss. In bf16 mode the
 * back-propagated pre-activation gradients are rounded to bf16 before they
 * meet weights or activations, as in the GPU kernels. */
float orc_mlp_backward(const float* params, const float* x, const float* y, int n, float* grads, int bf16);
/* Offsets of the parameter blocks in the flat array. */
#define ORC_MLP_W1 0
#define ORC_MLP_B1 (64 * 42)
#define ORC_MLP_W2 (ORC_MLP_B1 + 64)
#define ORC_MLP_B2 (ORC_MLP_W2 + 64 * 64)
#define ORC_MLP_W3 (ORC_MLP_B2 + 64)
#define ORC_MLP_B3 (ORC_MLP_W3 + 64 * 64)
#define ORC_MLP_W4 (ORC_MLP_B3 + 64)
#define ORC_MLP_B4 (ORC_MLP_W4 + 64)
/* One Adam step (t = 1-based step count). */
void orc_mlp_adam(float* params, const float* grads, float* m, float* v, int t, float lr, float beta1, float beta2,
                  float eps);

#ifdef __cplusplus
}
#endif
#endif
