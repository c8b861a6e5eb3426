// wavefront.h — batched wavefront renderer (generate / extend / shade / shadow /
// accumulate over shared device queues, many jobs per launch).
#pragma once

#include "internal.h"

namespace ry {

void wave_destroy(rayon_ctx* ctx);
// Drops cached pass graphs that were captured for `job` (called before the job is freed).
void wave_forget_job(rayon_ctx* ctx, rayon_job* job);

}  // namespace ry
