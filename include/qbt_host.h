/*
 * qbt_host.h - host-side C library of qubatron-mi355x (libqbt_host.so).
 *
 * Plain C11, no GPU dependency. It holds the data layer that the reference
 * keeps on the host and that the HIP layer (qbt.h) consumes:
 *
 *   - the flat sparse octree (node = 12 x int32: 8 child indices, the model
 *     index of the point that created the node, 3 zero pad words) with the
 *     reference's insertion, path insertion, removal and lookup semantics
 *     (src/qubatron/octree.c:11-252),
 *   - the flat point-model files .pnt/.nrm/.col/.rng (src/qubatron/model.c:53-111)
 *     and the qmc quantise/sort/dedup step that writes them (src/qubatron/qmc.c:41-327),
 *   - a seeded, counter-based synthetic scene generator (points on planes,
 *     boxes, spheres and ellipsoids; camera poses; lights) used by the tests
 *     and the benchmark.
 *
 * Every function returns 0 (or a count) on success and a negative QBT_E*
 * code on failure.
 */
#ifndef QBT_HOST_H
#define QBT_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QBT_OK 0
#define QBT_EINVAL (-22)   /* bad argument (NULL pointer, range, size) */
#define QBT_ENOMEM (-12)   /* allocation failed or allocation cap reached */
#define QBT_ERANGE (-34)   /* point outside the octree's addressable range */
#define QBT_EIO (-5)       /* file missing, short or unreadable */
#define QBT_EDEVICE (-19)  /* no usable GPU / HIP runtime error */
#define QBT_ETREE (-74)    /* node array references a node or point out of range */
#define QBT_EABORT (-125)  /* a ray exceeded the traversal iteration cap */
#define QBT_ENOTSUP (-95)  /* feature/size not supported by this build */

#define QBT_NODE_INTS 12 /* octets_t: src/qubatron/octree.c:11-14 */

/* ------------------------------------------------------------------------ */
/* Octree (src/qubatron/octree.c)                                           */
/* ------------------------------------------------------------------------ */

typedef struct qbt_octree {
    float    basecube[4]; /* x_min, y_max, z_max, edge (octree.c:18) */
    int32_t* octs;        /* size nodes x 12 int32, first len in use */
    size_t   size;        /* capacity in nodes */
    size_t   len;         /* nodes in use; node 0 is the root */
    int      levels;      /* leaf depth */
} qbt_octree;

/* octree_create (octree.c:55-67): root node all zero, len 1. */
int  qbt_octree_init(qbt_octree* t, const float basecube[4], int levels);
void qbt_octree_free(qbt_octree* t);
/* octree_reset (octree.c:89-93): len back to 1, root zeroed. */
void qbt_octree_reset(qbt_octree* t);
/* Reserve capacity for at least n nodes (no change to content). */
int  qbt_octree_reserve(qbt_octree* t, size_t n);

/* Octet of point p at a level whose child edge is `size`
 * (octree.c:104-109): (int)(x/size)%2, +2 if (int)(y/size)%2==0,
 * +4 if (int)(z/size)%2==0. Returns QBT_ERANGE for a point the reference
 * would index out of the node (coordinate <= -size, non-finite or too big). */
int qbt_octet_of(const float p[3], float size);

/* octree_insert_point (octree.c:95-147). Walks from node `index`; when a
 * child is created at level l, octindarr[l] = parent, octindarr[l+1] = new
 * node (octindarr needs levels+1 entries; may be NULL). path_out (levels
 * bytes, may be NULL) receives the octet of every level. */
int qbt_octree_insert_point(qbt_octree* t, size_t index, int32_t modind, const float p[3],
                            int32_t* octindarr, uint8_t* path_out);

/* octree_insert_path (octree.c:149-180): the octets come from three int
 * arrays (levels 0-3, 4-7, 8-11), as the GPU skinning pass writes them. */
int qbt_octree_insert_path(qbt_octree* t, size_t index, int32_t modind, const int32_t* octs14,
                           const int32_t* octs54, const int32_t* octs94);

/* octree_remove_point (octree.c:182-218): zeroes the leaf's slot in its
 * parent; *modind = leaf model index, *octind = parent node. Leaves both
 * untouched when the path does not exist. */
int qbt_octree_remove_point(qbt_octree* t, const float p[3], int32_t* modind, int32_t* octind);

/* octree_model_index (octree.c:220-252): same walk, no change. */
int qbt_octree_model_index(const qbt_octree* t, const float p[3], int32_t* modind, int32_t* octind);

/* Insert points [0,n) in order with model index first_modind + i, exactly as
 * the reference's load loop (modelutil.c:203-217). Returns QBT_OK or error. */
int qbt_octree_insert_points(qbt_octree* t, const float* xyz, size_t n, int32_t first_modind);

/* Octet paths of n points: out[i*levels + l] = octet at level l, computed
 * with the same fp32 formula as insertion (QBT_ERANGE on a bad point). */
int qbt_octree_paths(const float basecube[4], int levels, const float* xyz, size_t n, uint8_t* out,
                     int nthreads);

/* FNV-1a 64 hash of the first len nodes (48 bytes each). */
uint64_t qbt_octree_hash(const int32_t* octs, size_t len);

/* ------------------------------------------------------------------------ */
/* Flat model files (src/qubatron/model.c:53-111, src/qubatron/qmc.c)       */
/* ------------------------------------------------------------------------ */

/* Number of float3 points in a .pnt file (QBT_EIO if missing). */
int64_t qbt_flat_count(const char* pnt_path);
/* Read .pnt/.col/.nrm (raw float32 x3, little endian) into caller arrays of
 * capacity `cap` points; .rng (int32 triples) is optional (rng may be NULL).
 * Returns the point count or QBT_EIO on a missing/short file. */
int64_t qbt_flat_read(const char* pnt_path, const char* col_path, const char* nrm_path, float* pnt,
                      float* col, float* nrm, size_t cap);
int     qbt_flat_write(const char* pnt_path, const char* col_path, const char* nrm_path,
                       const float* pnt, const float* col, const float* nrm, size_t n);
/* qmc quantisation (qmc.c:41-127, 259, 291-327): drop points whose cell
 * floor(p/precision) (precision = size / 2^(levels+1)) lies outside
 * [0, 2^(levels+1)), order the rest by cell (x, then y, then z; ties keep
 * input order), keep the first point of every cell, and emit a range record
 * (first point index, x cell, y cell) whenever the (x, y) column changes.
 * Arrays are compacted in place; returns the kept count (rng_count set). */
int64_t qbt_qmc_quantize(float* pnt, float* col, float* nrm, size_t n, float size, int levels,
                         int32_t* rng_out, size_t* rng_count);

/* ------------------------------------------------------------------------ */
/* Seeded synthetic scenes                                                  */
/* ------------------------------------------------------------------------ */

#define QBT_MAX_BOXES 256
#define QBT_MAX_SPHERES 64
#define QBT_MAX_ELLIPSOIDS 16
#define QBT_MAX_LIGHTS 8

typedef struct qbt_scene_spec {
    int      config;       /* 1..5 preset id, 0 custom */
    uint64_t seed;
    uint64_t npoints;      /* total points */
    int      depth;        /* octree levels */
    float    size;         /* cube edge; basecube = {0, size, size, size} */
    float    inset;        /* cube faces inset (half a leaf) */
    /* point budget: [0, n_planes) cube faces, [n_planes, n_planes+n_objects)
     * boxes + spheres (area-proportional), rest ellipsoid cluster */
    uint64_t n_planes;
    uint64_t n_objects;
    uint64_t n_cluster;
    int      nboxes, nspheres, nellipsoids;
    float    box_lo[QBT_MAX_BOXES][3], box_hi[QBT_MAX_BOXES][3];
    float    sph_c[QBT_MAX_SPHERES][3], sph_r[QBT_MAX_SPHERES];
    float    ell_c[QBT_MAX_ELLIPSOIDS][3], ell_a[QBT_MAX_ELLIPSOIDS][3];
    float    ell_jitter;   /* radial jitter amplitude of cluster points */
    int      nlights;
    float    lights[QBT_MAX_LIGHTS][3];
} qbt_scene_spec;

/* Fill a preset: config 1..5 as described in the project documentation
 * (1: 1M pts depth 10 faces+20 boxes; 2: 10M depth 12 faces+40 boxes+10
 * spheres; 3/4: 90M depth 12 planes 60% + 200 boxes + 50 spheres + 10M
 * ellipsoid cluster; 5: the config-2 scene). npoints 0 keeps the preset's
 * count; a smaller count scales every budget proportionally. */
int qbt_scene_preset(qbt_scene_spec* s, int config, uint64_t seed, uint64_t npoints);

/* Generate points [first, first+count) (model index = generation index):
 * positions, normals (unit face/radial/gradient normals) and colours in
 * [0,1). Deterministic for a given spec, independent of nthreads. */
int qbt_scene_points(const qbt_scene_spec* s, uint64_t first, uint64_t count, float* pnt,
                     float* nrm, float* col, int nthreads);

/* Is p inside (or within margin of) a solid of the scene or the cube walls? */
int qbt_scene_blocked(const qbt_scene_spec* s, const float p[3], float margin);

/* Seeded camera poses: position uniform inside the cube, rejected when
 * within `margin` of a solid or wall; yaw U[0,2pi), pitch U[-pi/4,pi/4].
 * pos_out: n x 3, ang_out: n x 2 (yaw, pitch). */
int qbt_scene_poses(const qbt_scene_spec* s, uint64_t seed, int n, float margin, float* pos_out,
                    float* ang_out);

/* Counter-based RNG used by the generator (splitmix64 finaliser). */
uint64_t qbt_rng_u64(uint64_t key, uint64_t counter);
float    qbt_rng_f32(uint64_t key, uint64_t counter); /* [0,1), 24-bit */

#ifdef __cplusplus
}
#endif
#endif
