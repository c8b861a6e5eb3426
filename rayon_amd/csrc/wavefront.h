// wavefront.h — batched wavefront renderer (generate / extend / shade / shadow /
// accumulate over shared device queues, many jobs per launch).
#pragma once

#include "internal.h"

namespace ry {

void wave_destroy(rayon_ctx* ctx);

}  // namespace ry
