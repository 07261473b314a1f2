/*
 * glc_demo.c - a C caller of the octree_glc_* shim (octree_glc_qbt.c) that
 * replays the reference's calls for a static scene, for
 * tests/test_gpu_cshim.py:
 *
 *   model_load_flat of .pnt/.col/.nrm files (model.c:53-111)
 *   the insertion loop over the points (modelutil.c:203-217)
 *   the colour, normal and octree uploads (modelutil.c:227-255), and an
 *   empty dynamic tree
 *   octree_glc_update                                   -> frame 1
 *   the punch-hole edit (modelutil.c:405-500): each point within `radius`
 *   of (cx, cy, cz) is removed (its parent node uploaded, 48 bytes), then
 *   re-inserted 5 units out along its normal (every node it touched
 *   uploaded)
 *   octree_glc_update                                   -> frame 2
 *
 * Output file: int32 w, h, nnodes, nremoved; frame 1 and frame 2 as hit
 * (int32), leaf (int32), shadow (uint8), rgba8 (uint32) planes; the final
 * node array (nnodes x 12 int32).
 *
 * usage: glc_demo PNT COL NRM OUT W H px py pz ax ay lighta quality maxlevel
 *        basesize shoot cx cy cz radius
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "glc_types.h"
#include "qbt.h"

#define GL_INT_ 0x1404
#define GL_FLOAT_ 0x1406

static int write_frame(octree_glc_t* rc, FILE* f, int* w, int* h)
{
    qbt_frame d = {0};
    if (qbt_frame_device((qbt_ctx*) rc->qbt, 0, &d) != QBT_OK) return -1;
    const size_t n   = (size_t) d.w * (size_t) d.h;
    qbt_frame    fr  = {0};
    fr.hit           = malloc(4 * n);
    fr.leaf          = malloc(4 * n);
    fr.shadow        = malloc(n);
    fr.rgba8         = malloc(4 * n);
    int rc_          = -1;
    if (fr.hit && fr.leaf && fr.shadow && fr.rgba8 && qbt_read_frame((qbt_ctx*) rc->qbt, 0, &fr) == QBT_OK &&
        fwrite(fr.hit, 4, n, f) == n && fwrite(fr.leaf, 4, n, f) == n && fwrite(fr.shadow, 1, n, f) == n &&
        fwrite(fr.rgba8, 4, n, f) == n)
        rc_ = 0;
    *w = d.w, *h = d.h;
    free(fr.hit), free(fr.leaf), free(fr.shadow), free(fr.rgba8);
    return rc_;
}

int main(int argc, char** argv)
{
    if (argc != 21) {
        fprintf(stderr, "usage: %s PNT COL NRM OUT W H px py pz ax ay lighta quality maxlevel basesize shoot cx cy cz "
                        "radius\n", argv[0]);
        return 2;
    }
    const float width = (float) atof(argv[5]), height = (float) atof(argv[6]);
    v3_t        pos   = {(float) atof(argv[7]), (float) atof(argv[8]), (float) atof(argv[9])};
    v3_t        ang   = {(float) atof(argv[10]), (float) atof(argv[11]), 0.0f};
    const float lighta   = (float) atof(argv[12]);
    const int   quality  = atoi(argv[13]), maxlevel = atoi(argv[14]);
    const float basesize = (float) atof(argv[15]);
    const int   shoot    = atoi(argv[16]);
    const float cx = (float) atof(argv[17]), cy = (float) atof(argv[18]), cz = (float) atof(argv[19]);
    const float radius = (float) atof(argv[20]);

    /* model_load_flat */
    const int64_t n = qbt_flat_count(argv[1]);
    if (n <= 0) return fprintf(stderr, "flat count: %lld\n", (long long) n), 1;
    float* pnt = malloc(12 * (size_t) n);
    float* col = malloc(12 * (size_t) n);
    float* nrm = malloc(12 * (size_t) n);
    if (!pnt || !col || !nrm || qbt_flat_read(argv[1], argv[2], argv[3], pnt, col, nrm, (size_t) n) != n)
        return fprintf(stderr, "flat read failed\n"), 1;

    /* the insertion loop */
    qbt_octree  t;
    const float bc[4] = {0.0f, basesize, basesize, basesize};
    if (qbt_octree_init(&t, bc, maxlevel) != QBT_OK) return 1;
    for (int64_t i = 0; i < n; i++)
        if (qbt_octree_insert_point(&t, 0, (int32_t) i, pnt + 3 * i, NULL, NULL) != QBT_OK) return 1;

    octree_glc_t rc = octree_glc_init("");
    if (!rc.ok) return 1;
    const size_t fbytes = 12 * (size_t) n;
    octree_glc_upload_texbuffer_data(&rc, col, GL_FLOAT_, fbytes, 12, 0, fbytes, OCTREE_GLC_BUFFER_STATIC_COLOR);
    octree_glc_upload_texbuffer_data(&rc, nrm, GL_FLOAT_, fbytes, 12, 0, fbytes, OCTREE_GLC_BUFFER_STATIC_NORMAL);
    octree_glc_upload_texbuffer_data(&rc, t.octs, GL_INT_, t.len * 48, 16, 0, t.len * 48,
                                     OCTREE_GLC_BUFFER_STATIC_OCTREE);
    int32_t dyn_root[12] = {0};
    float   one[3]       = {0.0f, 0.0f, 0.0f};
    octree_glc_upload_texbuffer_data(&rc, dyn_root, GL_INT_, 48, 16, 0, 48, OCTREE_GLC_BUFFER_DYNAMIC_OCTREE);
    octree_glc_upload_texbuffer_data(&rc, one, GL_FLOAT_, 12, 12, 0, 12, OCTREE_GLC_BUFFER_DYNAMIC_COLOR);
    octree_glc_upload_texbuffer_data(&rc, one, GL_FLOAT_, 12, 12, 0, 12, OCTREE_GLC_BUFFER_DYNAMIC_NORMAL);

    FILE* out = fopen(argv[4], "wb");
    if (!out) return 1;
    int32_t hdr[4] = {0, 0, 0, 0};
    fwrite(hdr, 4, 4, out); /* rewritten at the end */
    int w = 0, h = 0;
    octree_glc_update(&rc, width, height, pos, ang, lighta, (uint8_t) quality, maxlevel, basesize, shoot);
    if (!rc.ok || write_frame(&rc, out, &w, &h)) return fprintf(stderr, "frame 1 failed\n"), 1;

    /* the punch-hole edit: remove, then re-insert moved points */
    int32_t* mods = malloc(4 * (size_t) n);
    float*   mv   = malloc(12 * (size_t) n);
    int      nrem = 0;
    for (int64_t i = 0; i < n; i++) {
        const float dx = pnt[3 * i] - cx, dy = pnt[3 * i + 1] - cy, dz = pnt[3 * i + 2] - cz;
        if (dx * dx + dy * dy + dz * dz >= radius * radius) continue;
        int32_t modind = -1, octind = -1;
        qbt_octree_remove_point(&t, pnt + 3 * i, &modind, &octind);
        if (octind < 0) continue;
        octree_glc_upload_texbuffer_data(&rc, t.octs, GL_INT_, t.len * 48, 16, (size_t) octind * 48,
                                         (size_t) (octind + 1) * 48, OCTREE_GLC_BUFFER_STATIC_OCTREE);
        for (int k = 0; k < 3; k++) mv[3 * nrem + k] = pnt[3 * i + k] + nrm[3 * modind + k] * 5.0f;
        mods[nrem++] = modind;
    }
    for (int i = 0; i < nrem; i++) {
        int32_t octindarr[24] = {0};
        if (qbt_octree_insert_point(&t, 0, mods[i], mv + 3 * i, octindarr, NULL) != QBT_OK) continue;
        for (int j = 0; j <= maxlevel; j++)
            if (octindarr[j] > 0)
                octree_glc_upload_texbuffer_data(&rc, t.octs, GL_INT_, t.len * 48, 16, (size_t) octindarr[j] * 48,
                                                 (size_t) (octindarr[j] + 1) * 48, OCTREE_GLC_BUFFER_STATIC_OCTREE);
    }
    octree_glc_update(&rc, width, height, pos, ang, lighta, (uint8_t) quality, maxlevel, basesize, shoot);
    if (!rc.ok || write_frame(&rc, out, &w, &h)) return fprintf(stderr, "frame 2 failed\n"), 1;
    if (fwrite(t.octs, 48, t.len, out) != t.len) return 1;
    hdr[0] = w, hdr[1] = h, hdr[2] = (int32_t) t.len, hdr[3] = nrem;
    fseek(out, 0, SEEK_SET);
    fwrite(hdr, 4, 4, out);
    fclose(out);
    printf("glc_demo: %lld points, %d removed and re-inserted, %dx%d frames, %zu nodes\n", (long long) n, nrem, w,
           h, t.len);
    qbt_destroy((qbt_ctx*) rc.qbt);
    qbt_octree_free(&t);
    free(pnt), free(col), free(nrm), free(mods), free(mv);
    return 0;
}
