/*
 * glc_types.h - the three types of the reference's renderer interface that
 * the octree_glc_* shim needs, declared with the reference's layouts:
 * v3_t (src/mt_math/mt_vector_3d.c:6-10, three floats), the buffer ids of
 * octree_glc_buffer_t (src/qubatron/octree_glc.c:16-24, same order) and an
 * octree_glc_t handle (the reference's struct holds GL objects, which the
 * shim replaces by one libqbt context).
 */
#ifndef GLC_TYPES_H
#define GLC_TYPES_H

#include <stddef.h>
#include <stdint.h>

typedef struct _v3_t {
    float x, y, z;
} v3_t;

typedef enum _octree_glc_buffer_t {
    OCTREE_GLC_BUFFER_STATIC_COLOR,
    OCTREE_GLC_BUFFER_STATIC_NORMAL,
    OCTREE_GLC_BUFFER_STATIC_OCTREE,
    OCTREE_GLC_BUFFER_DYNAMIC_COLOR,
    OCTREE_GLC_BUFFER_DYNAMIC_NORMAL,
    OCTREE_GLC_BUFFER_DYNAMIC_OCTREE
} octree_glc_buffer_t;

typedef struct octree_glc_t {
    void* qbt; /* the libqbt context (qbt_ctx*) */
    int   ok;
} octree_glc_t;

/* the reference's prototypes (octree_glc.c:64-85) */
octree_glc_t octree_glc_init(char* path);
void octree_glc_update(octree_glc_t* rc, float width, float height, v3_t position, v3_t angle, float lighta,
                       uint8_t quality, int maxlevel, float basesize, int shoot);
void octree_glc_upload_texbuffer_data(octree_glc_t* rc, void* data, int type, size_t size, size_t itemsize,
                                      size_t start, size_t end, octree_glc_buffer_t buftype);

#endif
