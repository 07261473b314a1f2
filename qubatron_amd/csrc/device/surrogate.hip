/*
 * surrogate.hip - shadow-visibility MLP (forward, backward, Adam) and the
 * surrogate render path.
 */
#include <hip/hip_runtime.h>

#include "qbt.h"
#include "qbt_internal.hpp"

struct MlpState {
    int unused;
};

void qbt_ctx::mlp_free()
{
    delete mlp;
    mlp = nullptr;
}

int qbt_ctx::render_surrogate(const qbt_view* views, int nviews, const float* lights, int nlights, unsigned out_mask)
{
    (void) views, (void) nviews, (void) lights, (void) nlights, (void) out_mask;
    return QBT_ENOTSUP;
}
