/*
 * octree_glc_qbt.c - the reference's renderer connector (octree_glc_init,
 * octree_glc_upload_texbuffer_data, octree_glc_update of
 * src/qubatron/octree_glc.c) on top of libqbt: the binding a maintainer adds
 * to the reference's C build in place of octree_glc.c. Every argument passes
 * through unchanged; the GL buffer types GL_INT / GL_FLOAT are the
 * QBT_TYPE_* values.
 */
#include <stdio.h>

#include "glc_types.h"
#include "qbt.h"

octree_glc_t octree_glc_init(char* base_path) /* octree_glc.c:93-247 */
{
    octree_glc_t rc = {0};
    qbt_opts     o  = {0, 0};
    qbt_ctx*     c  = NULL;
    (void) base_path; /* no shaders to compile */
    int e = qbt_create(&o, &c);
    if (e != QBT_OK) fprintf(stderr, "qbt_create: %s\n", qbt_strerror(e));
    rc.qbt = c;
    rc.ok  = e == QBT_OK;
    return rc;
}

void octree_glc_upload_texbuffer_data(octree_glc_t* rc, void* data, int type, size_t size, size_t itemsize,
                                      size_t start, size_t end, octree_glc_buffer_t buftype) /* :361-498 */
{
    int e = qbt_upload((qbt_ctx*) rc->qbt, data, type, size, itemsize, start, end, (int) buftype);
    if (e != QBT_OK) {
        fprintf(stderr, "qbt_upload: %s\n", qbt_strerror(e));
        rc->ok = 0;
    }
}

void octree_glc_update(octree_glc_t* rc, float width, float height, v3_t position, v3_t angle, float lighta,
                       uint8_t quality, int maxlevel, float basesize, int shoot) /* :249-352 */
{
    const float p[3] = {position.x, position.y, position.z};
    const float a[3] = {angle.x, angle.y, angle.z};
    int         e    = qbt_update((qbt_ctx*) rc->qbt, width, height, p, a, lighta, quality, maxlevel, basesize, shoot);
    if (e != QBT_OK) {
        fprintf(stderr, "qbt_update: %s\n", qbt_strerror(e));
        rc->ok = 0;
    }
    /* the frame (qbt_read_frame: RGBA8, r in the low byte, bottom row first)
     * replaces the FBO blit of :338-351 */
}
