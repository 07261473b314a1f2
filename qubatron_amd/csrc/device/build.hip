/*
 * build.hip - GPU octree construction (qbt_build_octree).
 */
#include <hip/hip_runtime.h>

#include "qbt.h"
#include "qbt_internal.hpp"

struct BuildState {
    int unused;
};

void qbt_ctx::builder_free()
{
    delete build;
    build = nullptr;
}

extern "C" int qbt_build_octree(qbt_ctx* c, int tree, const float basecube[4], int levels, size_t npoints,
                                size_t* nodes_out)
{
    (void) c, (void) tree, (void) basecube, (void) levels, (void) npoints, (void) nodes_out;
    return QBT_ENOTSUP;
}
