// sg_batch.cpp -- batched shots (placeholder until the batched sweeps land).
#include <hip/hip_runtime.h>
#include "../../include/sgfd.h"
#include "sg_error.hpp"

extern "C" {
int sg_batch_create(sg_solver *const *, int, sg_batch **out)
{
    if (out) *out = nullptr;
    return sg::fail(SG_ERR_UNSUPPORTED, "sg_batch_create: not available in this build");
}
int sg_batch_step(sg_batch *, int) { return sg::fail(SG_ERR_UNSUPPORTED, "sg_batch_step: not available"); }
int sg_batch_synchronize(sg_batch *) { return sg::fail(SG_ERR_UNSUPPORTED, "sg_batch_synchronize: not available"); }
int sg_batch_destroy(sg_batch *) { return SG_OK; }
}
