/*
 * qbt_oracle.c - CPU restatement of the reference octree traversal and pixel
 * program. TEST INFRASTRUCTURE ONLY (the parity checker of the HIP path);
 * never linked into the product library. Build: -O2 -ffp-contract=off.
 *
 * The traversal keeps the reference's own data structure on purpose: an
 * explicit per-level record with the cube, up to four pending entry points
 * and octets, the packed `ispsi` word and the static/dynamic node indices
 * (octree_fsh.c:50-58), so that every comparison, tie and flip happens in
 * the reference's order.
 */
#include "qbt_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct v4 {
    float x, y, z, w;
} v4;

/* octree_fsh.c:25-31 / octree.c:47-53 */
static const float XSFT[8] = {0, 1, 0, 1, 0, 1, 0, 1};
static const float YSFT[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const float ZSFT[8] = {0, 0, 0, 0, 1, 1, 1, 1};
static const int   HORPAIRS[8] = {1, 0, 3, 2, 5, 4, 7, 6};
static const int   VERPAIRS[8] = {2, 3, 0, 1, 6, 7, 4, 5};
static const int   DEPPAIRS[8] = {4, 5, 6, 7, 0, 1, 2, 3};

/* is_cube_xplane/yplane/zplane: octree_fsh.c:62-99 (shader: all FLT_MAX when
 * the direction component is 0), octree.c:302-339 (CPU: (0,0,0,FLT_MAX)) and
 * particle_vsh.c:62-99 (particle: (0,0,0,0)). */
static v4 plane_default(int sem)
{
    if (sem == ORC_SEM_CPU) return (v4){0.0f, 0.0f, 0.0f, FLT_MAX};
    if (sem == ORC_SEM_PARTICLE) return (v4){0.0f, 0.0f, 0.0f, 0.0f};
    return (v4){FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX};
}

static v4 plane_x(float c, const float* p, const float* d, int cpu)
{
    v4 r = plane_default(cpu);
    if (d[0] != 0.0f) {
        r.w = (c - p[0]) / d[0];
        r.y = p[1] + d[1] * r.w;
        r.z = p[2] + d[2] * r.w;
        r.x = c;
    }
    return r;
}

static v4 plane_y(float c, const float* p, const float* d, int cpu)
{
    v4 r = plane_default(cpu);
    if (d[1] != 0.0f) {
        r.w = (c - p[1]) / d[1];
        r.x = p[0] + d[0] * r.w;
        r.z = p[2] + d[2] * r.w;
        r.y = c;
    }
    return r;
}

static v4 plane_z(float c, const float* p, const float* d, int cpu)
{
    v4 r = plane_default(cpu);
    if (d[2] != 0.0f) {
        r.w = (c - p[2]) / d[2];
        r.x = p[0] + d[0] * r.w;
        r.y = p[1] + d[1] * r.w;
        r.z = c;
    }
    return r;
}

/* oct_from_octets_for_index (octree_fsh.c:127-136): slot k of node i; an
 * index 0 below the root means "no node" (absent branch, mostly in the
 * dynamic tree). A NULL tree is the empty root-only tree. */
static int32_t slot(const orc_tree* t, int k, int32_t i, int level, int* err)
{
    if (i == 0 && level > 0) return 0;
    if (!t->octs) return 0;
    if (i < 0 || (size_t) i >= t->len) {
        *err = 1;
        return 0;
    }
    return t->octs[(size_t) i * 12 + (size_t) k];
}

/* The particle shader's oct_from_octets_for_index (particle_vsh.c:110-119):
 * texel column 3 i mod 8192 + k / 4 on row 3 i / 8192, without carrying into
 * the next row. A column past 8191 is outside the 8192-wide texture; GLSL ES
 * leaves such a fetch undefined, and it is defined here as 0 (robust access). */
static int32_t slot_particle(const orc_tree* t, int k, int32_t i, int level, int* err)
{
    if (i == 0 && level > 0) return 0;
    const int64_t e  = (int64_t) i * 3;
    const int64_t cy = e / 8192;
    const int64_t cx = e - cy * 8192 + k / 4;
    if (cx >= 8192) return 0;
    return slot(t, k, i, level, err);
}

static void fetch_attr(const float* arr, size_t n, int32_t m, float out[4])
{
    if (!arr || m < 0 || (size_t) m >= n) { /* outside the texture: zeros */
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        return;
    }
    out[0] = arr[3 * (size_t) m];
    out[1] = arr[3 * (size_t) m + 1];
    out[2] = arr[3 * (size_t) m + 2];
    out[3] = 1.0f; /* RGB32F texelFetch returns alpha 1 */
}

typedef struct level_rec {
    v4      cube;
    v4      isps[4];
    int     octs[4];
    int     ispsi;
    int32_t socti;
    int32_t docti;
} level_rec;

int orc_trace(const orc_scene* sc, const float pos[3], const float dir[3], int semantics, orc_trace_out* out)
{
    const int cpu = semantics; /* plane-test defaults, see plane_default */
    const int one = semantics != ORC_SEM_SHADER; /* a single tree */
    const int prt = semantics == ORC_SEM_PARTICLE;
#define SLOT(t, k, i, lv, e) (prt ? slot_particle(t, k, i, lv, e) : slot(t, k, i, lv, e))
    memset(out, 0, sizeof(*out));
    out->status = ORC_MISS;
    out->socti = out->docti = -1;
    out->ms = out->md = -1;
    if (sc->maxlevel < 0 || sc->maxlevel > ORC_MAXLEVEL) {
        out->status = ORC_ERR;
        return ORC_ERR;
    }
    const orc_tree* dyn   = one ? &(orc_tree){NULL, 0} : &sc->dyn; /* CPU twin, particles: one tree */
    int             err   = 0;
    int             level = 0;
    level_rec       st[ORC_MAXLEVEL + 1];

    v4 base;
    memcpy(&base, sc->basecube, sizeof(base));
    st[0].cube  = base;
    st[0].socti = 0;
    st[0].docti = 0;
    st[0].ispsi = 0;

    v4  tlf = base;
    v4  brb = {tlf.x + tlf.w, tlf.y - tlf.w, tlf.z - tlf.w, 0.0f};
    v4  act;
    v4  hitp[6]; /* up to 6 face hits through edges/corners (reference has 4) */
    int hitc = 0;

    /* root faces in the reference order (octree_fsh.c:164-192); the CPU twin
     * also requires act.w < FLT_MAX (octree.c:359-386) */
#define ROOT_OK(a) (semantics != ORC_SEM_CPU || (a).w < FLT_MAX)
    act = plane_z(tlf.z, pos, dir, cpu); /* front */
    if (ROOT_OK(act) && tlf.x < act.x && act.x <= brb.x && tlf.y > act.y && act.y >= brb.y) hitp[hitc++] = act;
    act = plane_z(brb.z, pos, dir, cpu); /* back */
    if (ROOT_OK(act) && tlf.x < act.x && act.x <= brb.x && tlf.y > act.y && act.y >= brb.y) hitp[hitc++] = act;
    act = plane_x(tlf.x, pos, dir, cpu); /* left */
    if (ROOT_OK(act) && tlf.y > act.y && act.y >= brb.y && tlf.z > act.z && act.z >= brb.z) hitp[hitc++] = act;
    act = plane_x(brb.x, pos, dir, cpu); /* right */
    if (ROOT_OK(act) && tlf.y > act.y && act.y >= brb.y && tlf.z > act.z && act.z >= brb.z) hitp[hitc++] = act;
    act = plane_y(tlf.y, pos, dir, cpu); /* top */
    if (ROOT_OK(act) && tlf.x < act.x && act.x <= brb.x && tlf.z > act.z && act.z >= brb.z) hitp[hitc++] = act;
    act = plane_y(brb.y, pos, dir, cpu); /* bottom */
    if (ROOT_OK(act) && tlf.x < act.x && act.x <= brb.x && tlf.z > act.z && act.z >= brb.z) hitp[hitc++] = act;
#undef ROOT_OK

    /* octree_fsh.c:194-211 */
    if (hitc < 2) {
        out->status = ORC_DISCARD;
        return ORC_DISCARD;
    }
    if (hitp[0].w < 0.0f && hitp[1].w < 0.0f) {
        out->status = ORC_DISCARD;
        return ORC_DISCARD;
    }
    if (hitp[1].w < hitp[0].w) hitp[0] = hitp[1];
    if (hitp[0].w < 0.0f) hitp[0] = (v4){pos[0], pos[1], pos[2], 0.0f};
    st[level].isps[0] = hitp[0];

    for (;;) {
        out->iterations++;
        tlf = st[level].cube;

        /* leaf (octree_fsh.c:217-248) */
        if (level == sc->maxlevel) {
            v4 isp = st[level].isps[0];
            memcpy(out->isp, &isp, sizeof(isp));
            memcpy(out->tlf, &tlf, sizeof(tlf));
            out->socti = st[level].socti;
            out->docti = st[level].docti;
            out->ms    = SLOT(&sc->stat, 8, st[level].socti, level, &err);
            out->md    = slot(dyn, 8, st[level].docti, level, &err);
            fetch_attr(sc->col_s, sc->npts_s, out->ms, out->col);
            fetch_attr(sc->nrm_s, sc->npts_s, out->ms, out->nrm);
            if (out->md > 0) {
                fetch_attr(sc->col_d, sc->npts_d, out->md, out->col);
                fetch_attr(sc->nrm_d, sc->npts_d, out->md, out->nrm);
            }
            out->status = err ? ORC_ERR : ORC_HIT;
            return out->status;
        }

        /* expansion (octree_fsh.c:251-330) */
        if (st[level].ispsi == 0) {
            st[level].ispsi = 128;
            if (st[level].socti != 0 || level == 0) out->expansions++;
            if (dyn->octs && dyn->len > 1 && (st[level].docti != 0 || level == 0)) out->expansions_dyn++;

            v4 b = {tlf.x + tlf.w, tlf.y - tlf.w, tlf.z - tlf.w, 0.0f};
            v4 hlf;
            hlf.x = b.x + (tlf.x - b.x) * 0.5f;
            hlf.y = b.y + (tlf.y - b.y) * 0.5f;
            hlf.z = b.z + (tlf.z - b.z) * 0.5f;
            hlf.w = b.w + (tlf.w - b.w) * 0.5f;

            hitc    = 1;
            hitp[0] = st[level].isps[0];

            act = plane_z(hlf.z, pos, dir, cpu);
            if (act.w > 0.0f && tlf.x < act.x && act.x <= b.x && tlf.y > act.y && act.y >= b.y) hitp[hitc++] = act;
            act = plane_x(hlf.x, pos, dir, cpu);
            if (act.w > 0.0f && tlf.y > act.y && act.y >= b.y && tlf.z > act.z && act.z >= b.z) hitp[hitc++] = act;
            act = plane_y(hlf.y, pos, dir, cpu);
            if (act.w > 0.0f && tlf.x < act.x && act.x <= b.x && tlf.z > act.z && act.z >= b.z) hitp[hitc++] = act;

            int oct = 0;
            int pre = -1;
            for (int i = 0; i < hitc; ++i) {
                if (i < hitc - 1) {
                    for (int j = i + 1; j < hitc; ++j) {
                        if (hitp[j].w < hitp[i].w) {
                            act     = hitp[i];
                            hitp[i] = hitp[j];
                            hitp[j] = act;
                        }
                    }
                }
                act = hitp[i];
                oct = 0;
                if (act.x > hlf.x) oct = 1;
                if (act.y < hlf.y) oct += 2;
                if (act.z < hlf.z) oct += 4;
                if (oct == pre) {
                    if (act.x == hlf.x) oct = HORPAIRS[oct];
                    else if (act.y == hlf.y) oct = VERPAIRS[oct];
                    else if (act.z == hlf.z) oct = DEPPAIRS[oct];
                }
                pre = oct;

                int32_t s = SLOT(&sc->stat, oct, st[level].socti, level, &err);
                int32_t d = slot(dyn, oct, st[level].docti, level, &err);
                if (s > 0 || d > 0) {
                    int ind               = st[level].ispsi;
                    int len               = ind & 0x0F;
                    st[level].octs[len]   = oct;
                    st[level].isps[len]   = act;
                    st[level].ispsi       = (ind & 0xF0) | (len + 1);
                }
            }
            if (err) {
                out->status = ORC_ERR;
                return ORC_ERR;
            }
        }

        /* descend or backtrack (octree_fsh.c:334-375) */
        int cur_len = st[level].ispsi & 0x0F;
        if (cur_len > 0) {
            int   nxt_ind = (st[level].ispsi >> 4) & 7;
            v4    nxt_isp = st[level].isps[nxt_ind];
            int   nxt_oct = st[level].octs[nxt_ind];
            float halfs   = tlf.w / 2.0f;
            tlf.x += XSFT[nxt_oct] * halfs;
            tlf.y -= YSFT[nxt_oct] * halfs;
            tlf.z -= ZSFT[nxt_oct] * halfs;
            tlf.w = halfs;
            nxt_ind++;
            cur_len--;
            st[level].ispsi = 128 | (nxt_ind << 4) | cur_len;
            int32_t s = SLOT(&sc->stat, nxt_oct, st[level].socti, level, &err);
            int32_t d = slot(dyn, nxt_oct, st[level].docti, level, &err);
            if (err) {
                out->status = ORC_ERR;
                return ORC_ERR;
            }
            level += 1;
            st[level].cube    = tlf;
            st[level].ispsi   = 0;
            st[level].socti   = s;
            st[level].docti   = d;
            st[level].isps[0] = nxt_isp;
        } else {
            st[level--].ispsi = 0;
            if (level < 0) {
                out->status = ORC_MISS;
                return ORC_MISS;
            }
        }
    }
#undef SLOT
}

/* particle_vsh.c:345-373 */
int orc_particle_step(const orc_scene* sc, const float pos[3], const float spd[3], float pos_out[3],
                      float spd_out[3])
{
    for (int a = 0; a < 3; a++) pos_out[a] = pos[a], spd_out[a] = spd[a];
    if (!(spd_out[0] > -90000.0f)) return 0;
    spd_out[1] -= 0.4f;
    orc_trace_out r;
    if (orc_trace(sc, pos, spd_out, ORC_SEM_PARTICLE, &r) == ORC_ERR) return ORC_ERR;
    if (r.status == ORC_HIT && r.isp[3] > 0.0f) { /* res.isp.w > 0 (a miss leaves it 0) */
        const float dx = r.tlf[0] - pos[0], dy = r.tlf[1] - pos[1], dz = r.tlf[2] - pos[2];
        if (sqrtf(dx * dx + dy * dy + dz * dz) < 10.0f) {
            pos_out[0] = r.tlf[0], pos_out[1] = r.tlf[1], pos_out[2] = r.tlf[2];
            spd_out[0] = -100000.0f;
            return 0;
        }
    }
    for (int a = 0; a < 3; a++) pos_out[a] = pos[a] + spd_out[a];
    if (pos_out[1] < -10.0f) spd_out[0] = -100000.0f;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* pixel program (octree_fsh.c:381-464)                                     */
/* ------------------------------------------------------------------------ */

static void cross3(const float a[3], const float b[3], float r[3])
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

/* quat_rotate: v + 2 * cross(q.xyz, cross(q.xyz, v) + q.w * v) (octree_fsh.c:392-395) */
static void quat_rotate(const float q[4], const float v[3], float r[3])
{
    float c1[3], s[3], c2[3];
    cross3(q, v, c1);
    s[0] = c1[0] + q[3] * v[0];
    s[1] = c1[1] + q[3] * v[1];
    s[2] = c1[2] + q[3] * v[2];
    cross3(q, s, c2);
    r[0] = v[0] + 2.0f * c2[0];
    r[1] = v[1] + 2.0f * c2[1];
    r[2] = v[2] + 2.0f * c2[2];
}

/* quat_from_axis_angle (octree_fsh.c:381-390) */
static void quat_axis(const float axis[3], float angle, float q[4])
{
    float h = angle * 0.5f;
    q[0]    = axis[0] * sinf(h);
    q[1]    = axis[1] * sinf(h);
    q[2]    = axis[2] * sinf(h);
    q[3]    = cosf(h);
}

void orc_camera(const float angle[2], float qz[4], float vx[3], float qx[4])
{
    static const float ay[3] = {0.0f, 1.0f, 0.0f};
    static const float mx[3] = {-1.0f, 0.0f, 0.0f};
    quat_axis(ay, -angle[0], qz);
    quat_rotate(qz, mx, vx);
    quat_axis(vx, -angle[1], qx);
}

void orc_pixel_dir(const orc_view* v, const float qz[4], const float qx[4], int px, int py, float csv[3])
{
    float ctp[3] = {(float) px + 0.5f, (float) py + 0.5f, 0.0f};
    float cfp[3] = {v->W / 2.0f, v->H / 2.0f, (v->W / 2.0f) / 1.0f};
    float c0[3]  = {ctp[0] - cfp[0], ctp[1] - cfp[1], ctp[2] - cfp[2]};
    float c1[3];
    quat_rotate(qz, c0, c1);
    quat_rotate(qx, c1, csv);
}

static inline float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static inline void normalize3(const float v[3], float r[3])
{
    float s = 1.0f / sqrtf(dot3(v, v));
    r[0]    = v[0] * s;
    r[1]    = v[1] * s;
    r[2]    = v[2] * s;
}

/* GLSL max(x, y): y if x < y, else x */
static inline float gmax(float x, float y) { return x < y ? y : x; }

uint32_t orc_pack_rgba8(const float c[4])
{
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) {
        float x = c[k];
        if (!(x > 0.0f)) x = 0.0f; /* also NaN */
        if (x > 1.0f) x = 1.0f;
        uint32_t b = (uint32_t) (x * 255.0f + 0.5f);
        r |= b << (8 * k);
    }
    return r;
}

typedef struct sprite_info {
    int   discard; /* the sprite trace discarded */
    int   white;   /* resv.x / lghtv.x > 1 */
    float camlight_n[3];
} sprite_info;

typedef struct render_job {
    const orc_scene* sc;
    const orc_view*  v;
    int              nlights;
    const float*     lights;
    orc_frame*       fr;
    int              row0, row1;
    float            qz[4], qx[4];
    float            cos_thr;
    sprite_info*     spr;
    int              err;
} render_job;

static void render_pixel(render_job* j, int px, int py)
{
    const orc_scene* sc = j->sc;
    const orc_view*  v  = j->v;
    orc_frame*       fr = j->fr;
    size_t           o  = (size_t) py * (size_t) fr->w + (size_t) px;

    float csv[3];
    orc_pixel_dir(v, j->qz, j->qx, px, py, csv);

    orc_trace_out res;
    int           st   = orc_trace(sc, v->camfp, csv, ORC_SEM_SHADER, &res);
    int32_t       hit  = -1, leaf = -1, dyn = -1;
    uint8_t       flg  = 0, sbits = 0;
    float         col[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int64_t       exps   = 0;
    if (st == ORC_ERR) j->err = 1;

    if (st == ORC_DISCARD) {
        flg |= ORC_PX_DISCARD;
    } else {
        exps += res.expansions + res.expansions_dyn;
        if (st == ORC_HIT) {
            flg |= ORC_PX_LEAF;
            hit  = res.ms;
            leaf = res.socti;
            dyn  = res.md;
            memcpy(col, res.col, sizeof(col));
        }
        /* shading (octree_fsh.c:424-450); a miss has isp = 0 */
        if (st == ORC_HIT && res.isp[3] > 0.0f) {
            flg |= ORC_PX_SHADED;
            float nn[3], mc[3], nmc[3];
            normalize3(res.nrm, nn);
            mc[0] = -csv[0], mc[1] = -csv[1], mc[2] = -csv[2];
            normalize3(mc, nmc);
            float camv = gmax(dot3(nmc, nn), 0.0f);
            float f    = 0.1f + 0.2f * camv;
            for (int l = 0; l < j->nlights && !(flg & ORC_PX_DISCARD); l++) {
                const float*  L = j->lights + 3 * l;
                float         lghtv[3] = {res.isp[0] - L[0], res.isp[1] - L[1], res.isp[2] - L[2]};
                orc_trace_out lc;
                int           ls = orc_trace(sc, L, lghtv, ORC_SEM_SHADER, &lc);
                if (ls == ORC_ERR) j->err = 1;
                if (ls == ORC_DISCARD) {
                    flg |= ORC_PX_DISCARD;
                    break;
                }
                exps += lc.expansions + lc.expansions_dyn;
                float ispv[3] = {lc.isp[0] - res.isp[0], lc.isp[1] - res.isp[1], lc.isp[2] - res.isp[2]};
                float sqr     = ispv[0] * ispv[0] + ispv[1] * ispv[1] + ispv[2] * ispv[2];
                float bit     = (15.0f < sqr) ? 0.0f : 1.0f; /* step(sqr, 15.0) */
                if (bit != 0.0f) sbits |= (uint8_t) (1u << l);
                float ml[3] = {-lghtv[0], -lghtv[1], -lghtv[2]}, nml[3];
                normalize3(ml, nml);
                float lam = gmax(dot3(nml, nn), 0.0f);
                f         = f + lam * bit * 0.7f;
            }
            if (!(flg & ORC_PX_DISCARD)) {
                col[0] = col[0] * f;
                col[1] = col[1] * f;
                col[2] = col[2] * f;
                col[2] *= 0.7f;
                float g = (float) v->shoot * camv * 0.1f;
                col[0] += g;
                col[1] += g;
                col[2] += g;
            }
        }
        /* light sprite (octree_fsh.c:415-418, 452-459) */
        if (!(flg & ORC_PX_DISCARD)) {
            float ncsv[3];
            normalize3(csv, ncsv);
            for (int l = 0; l < j->nlights; l++) {
                if (dot3(j->spr[l].camlight_n, ncsv) > j->cos_thr) {
                    if (j->spr[l].discard) {
                        flg |= ORC_PX_DISCARD;
                        break;
                    }
                    if (j->spr[l].white) {
                        col[0] = col[1] = col[2] = col[3] = 1.0f;
                        flg |= ORC_PX_SPRITE;
                    }
                }
            }
        }
    }
    if (flg & ORC_PX_DISCARD) {
        col[0] = col[1] = col[2] = col[3] = 0.0f;
        sbits                             = 0;
    }
    if (fr->hit) fr->hit[o] = hit;
    if (fr->leaf) fr->leaf[o] = leaf;
    if (fr->dyn) fr->dyn[o] = dyn;
    if (fr->shadow) fr->shadow[o] = sbits;
    if (fr->flags) fr->flags[o] = flg;
    if (fr->rgba) memcpy(fr->rgba + 4 * o, col, sizeof(col));
    if (fr->rgba8) fr->rgba8[o] = orc_pack_rgba8(col);
    if (fr->isp) {
        float z[4] = {0, 0, 0, 0};
        memcpy(fr->isp + 4 * o, (st == ORC_HIT) ? res.isp : z, sizeof(z));
    }
    if (fr->expansions) fr->expansions[o] = (int32_t) exps;
}

static void* render_worker(void* arg)
{
    render_job* j = (render_job*) arg;
    for (int py = j->row0; py < j->row1; py++)
        for (int px = 0; px < j->fr->w; px++) render_pixel(j, px, py);
    return NULL;
}

int orc_render(const orc_scene* sc, const orc_view* v, int nlights, const float* lights, orc_frame* fr, int row0,
               int row1, int nthreads)
{
    if (nlights < 0 || nlights > 8 || row0 < 0 || row1 > fr->h || row0 > row1) return ORC_ERR;
    render_job base;
    memset(&base, 0, sizeof(base));
    base.sc      = sc;
    base.v       = v;
    base.nlights = nlights;
    base.lights  = lights;
    base.fr      = fr;
    float vx[3];
    orc_camera(v->angle, base.qz, vx, base.qx);
    base.cos_thr = cosf(0.02f);

    /* the light-sprite trace is the same for every pixel of the view */
    sprite_info spr[8];
    for (int l = 0; l < nlights; l++) {
        const float*  L      = lights + 3 * l;
        float         lv[3]  = {L[0] - v->camfp[0], L[1] - v->camfp[1], L[2] - v->camfp[2]};
        orc_trace_out lc;
        int           ls     = orc_trace(sc, v->camfp, lv, ORC_SEM_SHADER, &lc);
        spr[l].discard       = ls == ORC_DISCARD;
        float resvx          = lc.isp[0] - v->camfp[0];
        spr[l].white         = (resvx / lv[0]) > 1.0f;
        normalize3(lv, spr[l].camlight_n);
    }
    base.spr = spr;

    if (nthreads < 1) nthreads = 1;
    if (nthreads > 128) nthreads = 128;
    int rows = row1 - row0;
    if (nthreads > rows) nthreads = rows > 0 ? rows : 1;
    render_job* jobs = (render_job*) calloc((size_t) nthreads, sizeof(render_job));
    pthread_t*  th   = (pthread_t*) calloc((size_t) nthreads, sizeof(pthread_t));
    if (!jobs || !th) {
        free(jobs);
        free(th);
        return ORC_ERR;
    }
    /* interleaved row blocks so threads see similar work */
    int err = 0;
    if (nthreads == 1) {
        jobs[0]      = base;
        jobs[0].row0 = row0;
        jobs[0].row1 = row1;
        render_worker(&jobs[0]);
        err = jobs[0].err;
    } else {
        int per = (rows + nthreads - 1) / nthreads;
        for (int k = 0; k < nthreads; k++) {
            jobs[k]      = base;
            jobs[k].row0 = row0 + k * per < row1 ? row0 + k * per : row1;
            jobs[k].row1 = row0 + (k + 1) * per < row1 ? row0 + (k + 1) * per : row1;
            pthread_create(&th[k], NULL, render_worker, &jobs[k]);
        }
        for (int k = 0; k < nthreads; k++) {
            pthread_join(th[k], NULL);
            err |= jobs[k].err;
        }
    }
    free(jobs);
    free(th);
    return err ? ORC_ERR : 0;
}

/* ------------------------------------------------------------------------ */
/* shadow-visibility MLP: 42 -> 64 -> 64 -> 64 -> 1 (no reference counterpart) */
/* ------------------------------------------------------------------------ */

float orc_bf16_round(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) { /* inf / nan: keep, quiet nan */
        if (u & 0x007FFFFFu) u |= 0x00400000u;
        u &= 0xFFFF0000u;
    } else {
        uint32_t lsb = (u >> 16) & 1u;
        u += 0x7FFFu + lsb;
        u &= 0xFFFF0000u;
    }
    float r;
    memcpy(&r, &u, 4);
    return r;
}

void orc_mlp_features(const float isp[3], const float nrm[3], const float light[3], const float basecube[4],
                      float out[ORC_MLP_IN])
{
    const float w     = basecube[3];
    const float lo[3] = {basecube[0], basecube[1] - w, basecube[2] - w};
    float       u[3];
    for (int a = 0; a < 3; a++) u[a] = (isp[a] - lo[a]) / w;
    for (int k = 0; k < 6; k++) {
        float f = (float) (1 << k) * 3.14159265358979f;
        for (int a = 0; a < 3; a++) {
            float arg          = f * u[a];
            out[6 * k + 2 * a]     = sinf(arg);
            out[6 * k + 2 * a + 1] = cosf(arg);
        }
    }
    for (int a = 0; a < 3; a++) {
        out[36 + a] = nrm[a];
        out[39 + a] = (light[a] - lo[a]) / w;
    }
}

static inline float rb(float x, int bf16) { return bf16 ? orc_bf16_round(x) : x; }

/* one dense layer: z[o] = sum_i W[o][i] * h[i] + b[o], sequential fp32 sum */
static void dense(const float* W, const float* b, const float* h, int nin, int nout, float* z, int bf16)
{
    for (int o = 0; o < nout; o++) {
        float acc = 0.0f;
        for (int i = 0; i < nin; i++) acc += rb(W[o * nin + i], bf16) * h[i];
        z[o] = acc + b[o];
    }
}

typedef struct mlp_acts {
    float x[ORC_MLP_IN];
    float z1[64], h1[64], z2[64], h2[64], z3[64], h3[64];
    float logit, p;
} mlp_acts;

static void mlp_forward_one(const float* P, const float* xin, mlp_acts* a, int bf16)
{
    for (int i = 0; i < ORC_MLP_IN; i++) a->x[i] = rb(xin[i], bf16);
    dense(P + ORC_MLP_W1, P + ORC_MLP_B1, a->x, ORC_MLP_IN, 64, a->z1, bf16);
    for (int o = 0; o < 64; o++) a->h1[o] = rb(a->z1[o] > 0.0f ? a->z1[o] : 0.0f, bf16);
    dense(P + ORC_MLP_W2, P + ORC_MLP_B2, a->h1, 64, 64, a->z2, bf16);
    for (int o = 0; o < 64; o++) a->h2[o] = rb(a->z2[o] > 0.0f ? a->z2[o] : 0.0f, bf16);
    dense(P + ORC_MLP_W3, P + ORC_MLP_B3, a->h2, 64, 64, a->z3, bf16);
    for (int o = 0; o < 64; o++) a->h3[o] = rb(a->z3[o] > 0.0f ? a->z3[o] : 0.0f, bf16);
    float z4;
    dense(P + ORC_MLP_W4, P + ORC_MLP_B4, a->h3, 64, 1, &z4, bf16);
    a->logit = z4;
    a->p     = 1.0f / (1.0f + expf(-z4));
}

void orc_mlp_forward(const float* params, const float* x, int n, float* p, float* logits, int bf16)
{
    mlp_acts a;
    for (int s = 0; s < n; s++) {
        mlp_forward_one(params, x + (size_t) s * ORC_MLP_IN, &a, bf16);
        p[s] = a.p;
        if (logits) logits[s] = a.logit;
    }
}

float orc_mlp_backward(const float* params, const float* x, const float* y, int n, float* g, int bf16)
{
    const float* P = params;
    memset(g, 0, sizeof(float) * ORC_MLP_NPARAM);
    double   loss = 0.0;
    mlp_acts a;
    float    dz4, dz3[64], dz2[64], dz1[64], dh[64];
    for (int s = 0; s < n; s++) {
        mlp_forward_one(P, x + (size_t) s * ORC_MLP_IN, &a, bf16);
        float pc = a.p < 1e-7f ? 1e-7f : a.p > 1.0f - 1e-7f ? 1.0f - 1e-7f : a.p;
        loss -= y[s] > 0.5f ? log(pc) : log(1.0 - pc);
        dz4 = rb((a.p - y[s]) / (float) n, bf16);
        /* layer 4 */
        for (int i = 0; i < 64; i++) g[ORC_MLP_W4 + i] += dz4 * a.h3[i];
        g[ORC_MLP_B4] += dz4;
        for (int i = 0; i < 64; i++) dh[i] = rb(P[ORC_MLP_W4 + i], bf16) * dz4;
        /* ReLU derivative from the stored activation (h > 0), as the GPU sees it */
        for (int o = 0; o < 64; o++) dz3[o] = rb(a.h3[o] > 0.0f ? dh[o] : 0.0f, bf16);
        /* layer 3 */
        for (int o = 0; o < 64; o++) {
            for (int i = 0; i < 64; i++) g[ORC_MLP_W3 + o * 64 + i] += dz3[o] * a.h2[i];
            g[ORC_MLP_B3 + o] += dz3[o];
        }
        for (int i = 0; i < 64; i++) {
            float acc = 0.0f;
            for (int o = 0; o < 64; o++) acc += rb(P[ORC_MLP_W3 + o * 64 + i], bf16) * dz3[o];
            dh[i] = acc;
        }
        for (int o = 0; o < 64; o++) dz2[o] = rb(a.h2[o] > 0.0f ? dh[o] : 0.0f, bf16);
        /* layer 2 */
        for (int o = 0; o < 64; o++) {
            for (int i = 0; i < 64; i++) g[ORC_MLP_W2 + o * 64 + i] += dz2[o] * a.h1[i];
            g[ORC_MLP_B2 + o] += dz2[o];
        }
        for (int i = 0; i < 64; i++) {
            float acc = 0.0f;
            for (int o = 0; o < 64; o++) acc += rb(P[ORC_MLP_W2 + o * 64 + i], bf16) * dz2[o];
            dh[i] = acc;
        }
        for (int o = 0; o < 64; o++) dz1[o] = rb(a.h1[o] > 0.0f ? dh[o] : 0.0f, bf16);
        /* layer 1 */
        for (int o = 0; o < 64; o++) {
            for (int i = 0; i < ORC_MLP_IN; i++) g[ORC_MLP_W1 + o * ORC_MLP_IN + i] += dz1[o] * a.x[i];
            g[ORC_MLP_B1 + o] += dz1[o];
        }
    }
    return (float) (loss / (n > 0 ? n : 1));
}

void orc_mlp_adam(float* params, const float* grads, float* m, float* v, int t, float lr, float beta1, float beta2,
                  float eps)
{
    const float c1 = 1.0f - powf(beta1, (float) t);
    const float c2 = 1.0f - powf(beta2, (float) t);
    for (int i = 0; i < ORC_MLP_NPARAM; i++) {
        float gi  = grads[i];
        m[i]      = beta1 * m[i] + (1.0f - beta1) * gi;
        v[i]      = beta2 * v[i] + (1.0f - beta2) * gi * gi;
        float mh  = m[i] / c1;
        float vh  = v[i] / c2;
        params[i] = params[i] - lr * mh / (sqrtf(vh) + eps);
    }
}
