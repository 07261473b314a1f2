/*
 * Seeded synthetic point scenes for the traversal benchmarks and tests.
 *
 * Every point i draws from a counter-based RNG keyed by (seed, i), so a scene
 * is identical for any thread count and any chunking of the index range, and
 * generation order is the model index (the order the reference inserts points
 * in, src/qubatron/modelutil.c:203-217).
 *
 * Surfaces: the six inner faces of the cube [inset, size-inset]^3 (inset =
 * half a leaf, so no point sits on x|y|z == size, which the reference's
 * truncating octet formula would wrap to the low half, octree.c:104-109),
 * axis-aligned boxes, spheres and a cluster of jittered ellipsoids.
 */
#include "qbt_host.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define GOLDEN 0x9E3779B97F4A7C15ull
#define TWO_PI 6.28318530717958647692f

static inline uint64_t mix64(uint64_t z)
{
    z += GOLDEN;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t qbt_rng_u64(uint64_t key, uint64_t counter) { return mix64(key ^ mix64(counter)); }

float qbt_rng_f32(uint64_t key, uint64_t counter)
{
    return (float) (qbt_rng_u64(key, counter) >> 40) * (1.0f / 16777216.0f);
}

static inline double rng_f64(uint64_t key, uint64_t counter)
{
    return (double) (qbt_rng_u64(key, counter) >> 11) * (1.0 / 9007199254740992.0);
}

/* stream keys */
#define K_LAYOUT 0x4C41594F55540001ull
#define K_POINT 0x504F494E54530002ull
#define K_POSE 0x504F534553530003ull

/* ------------------------------------------------------------------------ */
/* presets                                                                  */
/* ------------------------------------------------------------------------ */

static int near_keepout(const float* kp, int nk, const float lo[3], const float hi[3], float margin)
{
    for (int k = 0; k < nk; k++) {
        const float* p = kp + 3 * k;
        if (p[0] > lo[0] - margin && p[0] < hi[0] + margin && p[1] > lo[1] - margin &&
            p[1] < hi[1] + margin && p[2] > lo[2] - margin && p[2] < hi[2] + margin)
            return 1;
    }
    return 0;
}

static double box_area(const float lo[3], const float hi[3])
{
    double a = hi[0] - lo[0], b = hi[1] - lo[1], c = hi[2] - lo[2];
    return 2.0 * (a * b + b * c + c * a);
}

int qbt_scene_preset(qbt_scene_spec* s, int config, uint64_t seed, uint64_t npoints)
{
    if (!s || config < 1 || config > 5) return QBT_EINVAL;
    memset(s, 0, sizeof(*s));
    s->config = config;
    s->seed   = seed;
    s->size   = 1800.0f;
    uint64_t n_default;
    int      nb, ns, ne;
    switch (config) {
    case 1: n_default = 1000000ull, s->depth = 10, nb = 20, ns = 0, ne = 0; break;
    case 2:
    case 5: n_default = 10000000ull, s->depth = 12, nb = 40, ns = 10, ne = 0; break;
    default: n_default = 90000000ull, s->depth = 12, nb = 200, ns = 50, ne = 15; break;
    }
    s->npoints = npoints ? npoints : n_default;
    s->inset   = s->size / (float) (1u << s->depth) * 0.5f;

    /* lights. Config 1 uses the light of the reference's demo scene,
     * (900, 1700, 900), which lies on two mid-planes of the root cube: its
     * shadow rays take the reference's start-on-a-plane path, which enters a
     * single quarter of the cube. The larger scenes keep every light off the
     * mid-planes so that their shadow rays are generic traversals. */
    static const float L1[3]   = {900.0f, 1700.0f, 900.0f};
    static const float L[4][3] = {
        {880.25f, 1700.0f, 915.75f}, {300.0f, 1500.0f, 1400.0f}, {1500.0f, 1600.0f, 300.0f}, {1450.0f, 950.0f, 1500.0f}};
    s->nlights = config == 3 ? 2 : config == 4 ? 4 : 1;
    if (config == 1) memcpy(s->lights, L1, sizeof(float) * 3);
    else memcpy(s->lights, L, sizeof(float) * 3 * s->nlights);

    /* keep-out points: lights, and the fixed config-1 camera */
    float kp[8][3];
    int   nk = 0;
    for (int l = 0; l < s->nlights; l++) memcpy(kp[nk++], s->lights[l], sizeof(float) * 3);
    kp[nk][0] = kp[nk][1] = kp[nk][2] = 900.0f;
    nk++;

    const uint64_t key = mix64(seed) ^ K_LAYOUT;
    uint64_t       ctr = 0;
    const float    in = s->inset, sz = s->size;

    /* ellipsoid cluster (a standing figure): centre x,z in [400,1400] */
    if (ne) {
        float cx = 400.0f + 1000.0f * qbt_rng_f32(key, ctr++);
        float cz = 400.0f + 1000.0f * qbt_rng_f32(key, ctr++);
        s->nellipsoids = ne;
        s->ell_jitter  = 1.0f;
        for (int k = 0; k < ne; k++) {
            for (int tries = 0;; tries++) {
                float a[3], c[3];
                for (int d = 0; d < 3; d++) a[d] = 20.0f + 100.0f * qbt_rng_f32(key, ctr++);
                c[0]       = cx - 150.0f + 300.0f * qbt_rng_f32(key, ctr++);
                c[1]       = 130.0f + 600.0f * qbt_rng_f32(key, ctr++);
                c[2]       = cz - 150.0f + 300.0f * qbt_rng_f32(key, ctr++);
                float lo[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
                float hi[3] = {c[0] + a[0], c[1] + a[1], c[2] + a[2]};
                if (!near_keepout(&kp[0][0], nk, lo, hi, 40.0f) || tries > 1000) {
                    memcpy(s->ell_c[k], c, sizeof(c));
                    memcpy(s->ell_a[k], a, sizeof(a));
                    break;
                }
            }
        }
        s->n_cluster = s->npoints / 9; /* 10M of 90M */
    }

    /* boxes: edges U[50,400] per axis, inside the cube */
    s->nboxes = nb;
    for (int b = 0; b < nb; b++) {
        for (int tries = 0;; tries++) {
            float lo[3], hi[3];
            for (int d = 0; d < 3; d++) {
                float e = 50.0f + 350.0f * qbt_rng_f32(key, ctr++);
                lo[d]   = in + (sz - 2.0f * in - e) * qbt_rng_f32(key, ctr++);
                hi[d]   = lo[d] + e;
            }
            if (!near_keepout(&kp[0][0], nk, lo, hi, 40.0f) || tries > 1000) {
                memcpy(s->box_lo[b], lo, sizeof(lo));
                memcpy(s->box_hi[b], hi, sizeof(hi));
                break;
            }
        }
    }
    /* spheres: radius U[30,300], inside the cube */
    s->nspheres = ns;
    for (int k = 0; k < ns; k++) {
        for (int tries = 0;; tries++) {
            float r = 30.0f + 270.0f * qbt_rng_f32(key, ctr++);
            float c[3];
            for (int d = 0; d < 3; d++) c[d] = in + r + (sz - 2.0f * in - 2.0f * r) * qbt_rng_f32(key, ctr++);
            float lo[3] = {c[0] - r, c[1] - r, c[2] - r}, hi[3] = {c[0] + r, c[1] + r, c[2] + r};
            if (!near_keepout(&kp[0][0], nk, lo, hi, 40.0f) || tries > 1000) {
                memcpy(s->sph_c[k], c, sizeof(c));
                s->sph_r[k] = r;
                break;
            }
        }
    }

    /* point budget */
    uint64_t rest = s->npoints - s->n_cluster;
    if (config >= 3 && config <= 4) {
        s->n_planes  = (uint64_t) ((double) s->npoints * 0.6);
        s->n_objects = rest - s->n_planes;
    } else {
        double a_faces = 6.0 * (double) (sz - 2 * in) * (double) (sz - 2 * in);
        double a_obj   = 0.0;
        for (int b = 0; b < nb; b++) a_obj += box_area(s->box_lo[b], s->box_hi[b]);
        for (int k = 0; k < ns; k++) a_obj += 4.0 * 3.14159265358979 * s->sph_r[k] * s->sph_r[k];
        s->n_planes  = (uint64_t) llround((double) rest * a_faces / (a_faces + a_obj));
        s->n_objects = rest - s->n_planes;
    }
    return QBT_OK;
}

/* ------------------------------------------------------------------------ */
/* points                                                                   */
/* ------------------------------------------------------------------------ */

typedef struct surf_table {
    int     nitems;
    double* cdf;   /* cumulative area, nitems */
    int*    kind;  /* 0 box face, 1 sphere */
    int*    which; /* box*6+face or sphere id */
} surf_table;

static int build_table(const qbt_scene_spec* s, surf_table* t)
{
    t->nitems = s->nboxes * 6 + s->nspheres;
    t->cdf    = (double*) malloc(sizeof(double) * (t->nitems + 1));
    t->kind   = (int*) malloc(sizeof(int) * (t->nitems + 1));
    t->which  = (int*) malloc(sizeof(int) * (t->nitems + 1));
    if (!t->cdf || !t->kind || !t->which) return QBT_ENOMEM;
    double acc = 0.0;
    int    k   = 0;
    for (int b = 0; b < s->nboxes; b++)
        for (int f = 0; f < 6; f++) {
            int    ax = f >> 1, a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
            double e1 = s->box_hi[b][a1] - s->box_lo[b][a1], e2 = s->box_hi[b][a2] - s->box_lo[b][a2];
            acc += e1 * e2;
            t->cdf[k] = acc, t->kind[k] = 0, t->which[k] = b * 6 + f;
            k++;
        }
    for (int q = 0; q < s->nspheres; q++) {
        acc += 4.0 * 3.14159265358979 * s->sph_r[q] * s->sph_r[q];
        t->cdf[k] = acc, t->kind[k] = 1, t->which[k] = q;
        k++;
    }
    for (int i = 0; i < t->nitems; i++) t->cdf[i] /= acc > 0 ? acc : 1.0;
    return QBT_OK;
}

static void free_table(surf_table* t)
{
    free(t->cdf);
    free(t->kind);
    free(t->which);
}

static inline void unit_dir(float u1, float u2, float d[3])
{
    float z   = 1.0f - 2.0f * u1;
    float rxy = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float ph  = TWO_PI * u2;
    d[0]      = rxy * cosf(ph);
    d[1]      = rxy * sinf(ph);
    d[2]      = z;
}

typedef struct gen_job {
    const qbt_scene_spec* s;
    const surf_table*     t;
    uint64_t              lo, hi, first;
    float *               pnt, *nrm, *col;
} gen_job;

static void gen_point(const qbt_scene_spec* s, const surf_table* t, uint64_t i, float* p, float* n, float* c)
{
    const uint64_t key = mix64(s->seed) ^ K_POINT;
    const uint64_t c0  = i * 8ull;
    float          u[7];
    for (int j = 0; j < 7; j++) u[j] = qbt_rng_f32(key, c0 + (uint64_t) j);
    const float in = s->inset, sz = s->size;
    n[0] = n[1] = n[2] = 0.0f;
    if (i < s->n_planes) {
        int f = (int) ((qbt_rng_u64(key, c0 + 7) >> 32) % 6u);
        int ax = f >> 1, a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
        p[ax] = (f & 1) ? sz - in : in;
        n[ax] = (f & 1) ? -1.0f : 1.0f;
        p[a1] = in + (sz - 2.0f * in) * u[0];
        p[a2] = in + (sz - 2.0f * in) * u[1];
    } else if (i < s->n_planes + s->n_objects) {
        double ud = rng_f64(key, c0 + 7);
        int    lo = 0, hi = t->nitems - 1;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (t->cdf[mid] > ud) hi = mid;
            else lo = mid + 1;
        }
        if (t->kind[lo] == 0) {
            int b = t->which[lo] / 6, f = t->which[lo] % 6;
            int ax = f >> 1, a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
            p[ax] = (f & 1) ? s->box_hi[b][ax] : s->box_lo[b][ax];
            n[ax] = (f & 1) ? 1.0f : -1.0f;
            p[a1] = s->box_lo[b][a1] + (s->box_hi[b][a1] - s->box_lo[b][a1]) * u[0];
            p[a2] = s->box_lo[b][a2] + (s->box_hi[b][a2] - s->box_lo[b][a2]) * u[1];
        } else {
            int   q = t->which[lo];
            float d[3];
            unit_dir(u[0], u[1], d);
            for (int k = 0; k < 3; k++) {
                p[k] = s->sph_c[q][k] + s->sph_r[q] * d[k];
                n[k] = d[k];
            }
        }
    } else {
        int   e = (int) ((i - s->n_planes - s->n_objects) % (uint64_t) (s->nellipsoids ? s->nellipsoids : 1));
        float d[3], g[3];
        unit_dir(u[0], u[1], d);
        float jit = (2.0f * u[2] - 1.0f) * s->ell_jitter;
        for (int k = 0; k < 3; k++) {
            p[k] = s->ell_c[e][k] + s->ell_a[e][k] * d[k] + jit * d[k];
            g[k] = d[k] / s->ell_a[e][k];
        }
        float inv = 1.0f / sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        for (int k = 0; k < 3; k++) n[k] = g[k] * inv;
    }
    for (int k = 0; k < 3; k++) {
        if (p[k] < in) p[k] = in;
        if (p[k] > sz - in) p[k] = sz - in;
    }
    c[0] = u[4], c[1] = u[5], c[2] = u[6];
}

static void* gen_worker(void* arg)
{
    gen_job* j = (gen_job*) arg;
    for (uint64_t i = j->lo; i < j->hi; i++) {
        uint64_t o = (i - j->first) * 3;
        gen_point(j->s, j->t, i, j->pnt + o, j->nrm + o, j->col + o);
    }
    return NULL;
}

int qbt_scene_points(const qbt_scene_spec* s, uint64_t first, uint64_t count, float* pnt, float* nrm,
                     float* col, int nthreads)
{
    if (!s || (count && (!pnt || !nrm || !col))) return QBT_EINVAL;
    if (first + count > s->npoints) return QBT_EINVAL;
    surf_table t;
    int        rc = build_table(s, &t);
    if (rc) {
        free_table(&t);
        return rc;
    }
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    pthread_t th[64];
    gen_job   jobs[64];
    uint64_t  per = (count + nthreads - 1) / nthreads;
    for (int k = 0; k < nthreads; k++) {
        uint64_t lo = first + (k * per < count ? k * per : count);
        uint64_t hi = first + ((k + 1) * per < count ? (k + 1) * per : count);
        jobs[k]     = (gen_job){s, &t, lo, hi, first, pnt, nrm, col};
        pthread_create(&th[k], NULL, gen_worker, &jobs[k]);
    }
    for (int k = 0; k < nthreads; k++) pthread_join(th[k], NULL);
    free_table(&t);
    return QBT_OK;
}

int qbt_scene_blocked(const qbt_scene_spec* s, const float p[3], float margin)
{
    for (int d = 0; d < 3; d++)
        if (p[d] < margin || p[d] > s->size - margin) return 1;
    for (int b = 0; b < s->nboxes; b++)
        if (p[0] > s->box_lo[b][0] - margin && p[0] < s->box_hi[b][0] + margin &&
            p[1] > s->box_lo[b][1] - margin && p[1] < s->box_hi[b][1] + margin &&
            p[2] > s->box_lo[b][2] - margin && p[2] < s->box_hi[b][2] + margin)
            return 1;
    for (int q = 0; q < s->nspheres; q++) {
        float dx = p[0] - s->sph_c[q][0], dy = p[1] - s->sph_c[q][1], dz = p[2] - s->sph_c[q][2];
        float r  = s->sph_r[q] + margin;
        if (dx * dx + dy * dy + dz * dz < r * r) return 1;
    }
    for (int e = 0; e < s->nellipsoids; e++) {
        float am = fmaxf(s->ell_a[e][0], fmaxf(s->ell_a[e][1], s->ell_a[e][2])) + s->ell_jitter + margin;
        float dx = p[0] - s->ell_c[e][0], dy = p[1] - s->ell_c[e][1], dz = p[2] - s->ell_c[e][2];
        if (dx * dx + dy * dy + dz * dz < am * am) return 1;
    }
    return 0;
}

int qbt_scene_poses(const qbt_scene_spec* s, uint64_t seed, int n, float margin, float* pos_out, float* ang_out)
{
    if (!s || n < 0 || (n && (!pos_out || !ang_out))) return QBT_EINVAL;
    const uint64_t key = mix64(seed) ^ K_POSE;
    uint64_t       ctr = 0;
    for (int k = 0; k < n; k++) {
        int found = 0;
        for (int tries = 0; tries < 100000 && !found; tries++) {
            float p[3];
            for (int d = 0; d < 3; d++) p[d] = margin + (s->size - 2.0f * margin) * qbt_rng_f32(key, ctr++);
            if (!qbt_scene_blocked(s, p, margin)) {
                memcpy(pos_out + 3 * k, p, sizeof(p));
                found = 1;
            }
        }
        if (!found) return QBT_ERANGE;
        ang_out[2 * k]     = TWO_PI * qbt_rng_f32(key, ctr++);
        ang_out[2 * k + 1] = (qbt_rng_f32(key, ctr++) - 0.5f) * 1.57079632679f;
    }
    return QBT_OK;
}
