// wavefront.hip — batched wavefront renderer. (Stage kernels follow.)
#include <hip/hip_runtime.h>

#include "internal.h"
#include "wavefront.h"

namespace ry {

void wave_destroy(rayon_ctx*) {}

}  // namespace ry

extern "C" rayon_status rayon_render_batched(rayon_ctx* ctx, rayon_job* const* jobs, int32_t n, int32_t samples) {
   (void)ctx;
   (void)jobs;
   (void)n;
   (void)samples;
   ry::set_error("rayon_render_batched: not built yet");
   return RAYON_ERR_UNSUPPORTED;
}

extern "C" rayon_status rayon_load_yaml(const char* path, rayon_generated** out) {
   (void)path;
   (void)out;
   ry::set_error("rayon_load_yaml: not built yet");
   return RAYON_ERR_UNSUPPORTED;
}
