/*
 * Flat sparse octree on the host: construction and edits with the node
 * numbering of src/qubatron/octree.c, so that node indices produced here are
 * the indices the reference would produce for the same insertion sequence.
 *
 * Node layout (octree.c:11-14): 12 int32 per node, [0..7] child node index
 * (0 = no child; node 0 is the root and never a child), [8] model index of the
 * point whose insertion created the node, [9..11] zero.
 *
 * Octets (octree.c:41-53): bit0 = high-x half, bit1 = low-y half, bit2 = low-z
 * half of the parent cube; the cube spans x in [x, x+w], y in [y-w, y],
 * z in [z-w, z] for basecube (x, y, z, w).
 */
#include "qbt_host.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define NODE_BYTES (QBT_NODE_INTS * sizeof(int32_t))

int qbt_octree_init(qbt_octree* t, const float basecube[4], int levels)
{
    if (!t || !basecube || levels < 0 || levels > 30) return QBT_EINVAL;
    memcpy(t->basecube, basecube, sizeof(t->basecube));
    t->levels = levels;
    t->size   = 1024;
    t->octs   = (int32_t*) calloc(t->size, NODE_BYTES);
    if (!t->octs) return QBT_ENOMEM;
    t->len = 1; /* root, all zero (octree.c:63-64) */
    return QBT_OK;
}

void qbt_octree_free(qbt_octree* t)
{
    if (!t) return;
    free(t->octs);
    t->octs = NULL;
    t->size = t->len = 0;
}

void qbt_octree_reset(qbt_octree* t)
{
    if (!t || !t->octs) return;
    memset(t->octs, 0, NODE_BYTES);
    t->len = 1;
}

int qbt_octree_reserve(qbt_octree* t, size_t n)
{
    if (!t) return QBT_EINVAL;
    if (n <= t->size) return QBT_OK;
    int32_t* p = (int32_t*) realloc(t->octs, n * NODE_BYTES);
    if (!p) return QBT_ENOMEM;
    memset(p + t->size * QBT_NODE_INTS, 0, (n - t->size) * NODE_BYTES);
    t->octs = p;
    t->size = n;
    return QBT_OK;
}

/* (int)(v/size)%2 with the C truncation of the reference. A quotient outside
 * the int range (or NaN) is undefined behaviour in the reference and is
 * flagged; negative quotients are kept as the reference computes them. */
static inline int half_bit(float v, float size, int* bad)
{
    float q = v / size;
    if (!(q > -2147483520.0f && q < 2147483520.0f)) {
        *bad = 1;
        return 0;
    }
    return (int) q % 2;
}

/* A final octet < 0 (x bit -1 with y and z bits non-zero) would index before
 * the node in the reference; it is reported as QBT_ERANGE. */
int qbt_octet_of(const float p[3], float size)
{
    int bad = 0;
    int o   = half_bit(p[0], size, &bad);
    int yi  = half_bit(p[1], size, &bad);
    int zi  = half_bit(p[2], size, &bad);
    if (yi == 0) o += 2;
    if (zi == 0) o += 4;
    if (bad || o < 0) return QBT_ERANGE;
    return o;
}

/* Append a node {0,..,0, modind}; grows geometrically (the reference grows by
 * 100000 nodes, octree.c:128-134; growth policy does not change indices). */
static int append_node(qbt_octree* t, int32_t modind, size_t* idx)
{
    if (t->len == t->size) {
        size_t ns = t->size < 1024 ? 1024 : t->size + t->size / 2;
        int    rc = qbt_octree_reserve(t, ns);
        if (rc) return rc;
    }
    if (t->len >= (size_t) INT32_MAX) return QBT_ENOMEM;
    int32_t* n = t->octs + t->len * QBT_NODE_INTS;
    memset(n, 0, NODE_BYTES);
    n[8] = modind;
    *idx = t->len++;
    return QBT_OK;
}

int qbt_octree_insert_point(qbt_octree* t, size_t index, int32_t modind, const float p[3],
                            int32_t* octindarr, uint8_t* path_out)
{
    if (!t || !t->octs || !p || index >= t->len) return QBT_EINVAL;
    float size = t->basecube[3];
    for (int level = 0; level < t->levels; level++) {
        size /= 2.0f;
        int octet = qbt_octet_of(p, size);
        if (octet < 0) return octet;
        int32_t* node = t->octs + index * QBT_NODE_INTS;
        if (node[octet] == 0) {
            size_t nidx;
            if (octindarr) {
                octindarr[level]     = (int32_t) index;
                octindarr[level + 1] = (int32_t) t->len;
            }
            int rc = append_node(t, modind, &nidx);
            if (rc) return rc;
            t->octs[index * QBT_NODE_INTS + octet] = (int32_t) nidx; /* re-read: append may realloc */
        }
        index = (size_t) t->octs[index * QBT_NODE_INTS + octet];
        if (path_out) path_out[level] = (uint8_t) octet;
    }
    return QBT_OK;
}

int qbt_octree_insert_path(qbt_octree* t, size_t index, int32_t modind, const int32_t* octs14,
                           const int32_t* octs54, const int32_t* octs94)
{
    if (!t || !t->octs || !octs14 || index >= t->len) return QBT_EINVAL;
    if (t->levels > 4 && !octs54) return QBT_EINVAL;
    if (t->levels > 8 && !octs94) return QBT_EINVAL;
    if (t->levels > 12) return QBT_ENOTSUP; /* three int4 buffers hold 12 levels */
    for (int level = 0; level < t->levels; level++) {
        int octet = level > 7 ? octs94[level - 8] : level > 3 ? octs54[level - 4] : octs14[level];
        if (octet < 0 || octet > 7) return QBT_ERANGE;
        if (t->octs[index * QBT_NODE_INTS + octet] == 0) {
            size_t nidx;
            int    rc = append_node(t, modind, &nidx);
            if (rc) return rc;
            t->octs[index * QBT_NODE_INTS + octet] = (int32_t) nidx;
        }
        index = (size_t) t->octs[index * QBT_NODE_INTS + octet];
    }
    return QBT_OK;
}

/* Shared walk of octree_remove_point / octree_model_index: levels+1 steps,
 * the last one lands on the leaf. */
static int walk_leaf(qbt_octree* t, const float p[3], int32_t* modind, int32_t* octind, int zero)
{
    if (!t || !t->octs || !p || !modind || !octind) return QBT_EINVAL;
    float  size   = t->basecube[3];
    size_t index  = 0;
    size_t lindex = 0;
    int    loctet = 0;
    for (int level = 0; level < t->levels + 1; level++) {
        size /= 2.0f;
        int octet = qbt_octet_of(p, size);
        if (octet < 0) return octet;
        if (level == t->levels) {
            *modind = t->octs[index * QBT_NODE_INTS + 8];
            *octind = (int32_t) lindex;
            if (zero) t->octs[lindex * QBT_NODE_INTS + loctet] = 0;
        }
        lindex = index;
        loctet = octet;
        index  = (size_t) t->octs[index * QBT_NODE_INTS + octet];
        if (index == 0) return QBT_OK;
        if (index >= t->len) return QBT_ETREE;
    }
    return QBT_OK;
}

int qbt_octree_remove_point(qbt_octree* t, const float p[3], int32_t* modind, int32_t* octind)
{
    return walk_leaf(t, p, modind, octind, 1);
}

int qbt_octree_model_index(const qbt_octree* t, const float p[3], int32_t* modind, int32_t* octind)
{
    return walk_leaf((qbt_octree*) t, p, modind, octind, 0);
}

int qbt_octree_insert_points(qbt_octree* t, const float* xyz, size_t n, int32_t first_modind)
{
    if (!t || (!xyz && n)) return QBT_EINVAL;
    for (size_t i = 0; i < n; i++) {
        int rc = qbt_octree_insert_point(t, 0, first_modind + (int32_t) i, xyz + 3 * i, NULL, NULL);
        if (rc) return rc;
    }
    return QBT_OK;
}

typedef struct paths_job {
    const float* xyz;
    uint8_t*     out;
    size_t       lo, hi;
    float        edge;
    int          levels;
    int          rc;
} paths_job;

static void* paths_worker(void* arg)
{
    paths_job* j = (paths_job*) arg;
    for (size_t i = j->lo; i < j->hi; i++) {
        float size = j->edge;
        for (int l = 0; l < j->levels; l++) {
            size /= 2.0f;
            int o = qbt_octet_of(j->xyz + 3 * i, size);
            if (o < 0) {
                j->rc = o;
                return NULL;
            }
            j->out[i * (size_t) j->levels + l] = (uint8_t) o;
        }
    }
    return NULL;
}

int qbt_octree_paths(const float basecube[4], int levels, const float* xyz, size_t n, uint8_t* out,
                     int nthreads)
{
    if (!basecube || (!xyz && n) || (!out && n) || levels < 0) return QBT_EINVAL;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    pthread_t th[64];
    paths_job jobs[64];
    size_t    per = (n + nthreads - 1) / nthreads;
    for (int k = 0; k < nthreads; k++) {
        jobs[k] = (paths_job){xyz, out, k * per < n ? k * per : n, (k + 1) * per < n ? (k + 1) * per : n,
                              basecube[3], levels, 0};
        pthread_create(&th[k], NULL, paths_worker, &jobs[k]);
    }
    int rc = QBT_OK;
    for (int k = 0; k < nthreads; k++) {
        pthread_join(th[k], NULL);
        if (jobs[k].rc) rc = jobs[k].rc;
    }
    return rc;
}

uint64_t qbt_octree_hash(const int32_t* octs, size_t len)
{
    uint64_t       h = 1469598103934665603ull;
    const uint8_t* b = (const uint8_t*) octs;
    size_t         nb = len * NODE_BYTES;
    for (size_t i = 0; i < nb; i++) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}
