/*
 * rayon_legacy.h — the reference's GPU-render entry points, exported under their own
 * names by librayon_hip.so so that the reference's host renderers link against this
 * library unchanged.
 *
 * Each declaration replaces the same-named function of
 * src/rayon/gpu_renderers/renderer_cuda_device.cuh (line cited per entry); the scene
 * handle replaces the device CudaScene::Scene* that Scene::CudaSceneBuilder::buildGPUScene
 * returned (scenes/scene_builder.hpp:234-239, scene_builder_cuda.cu:148-303).
 *
 * Semantics kept from the reference: the sequential per-pixel random stream seeded
 * 1234 + idx * 747796405 (cuda_utils.cu:26), the reference material quirks (metal
 * roughness 0 -> 0.3, global metal fuzziness and glass IOR), lazily created state behind
 * caller-owned void* handles, a host float3 accumulation buffer that is uploaded on the
 * first call and downloaded after every call when non-NULL, the sign-encoded adaptive
 * sample counts, and the primary-sample return value. Errors are reported on stderr and
 * the call returns, as in the reference; rayon_last_error() holds the message.
 * Calls run on one library-owned context on device 0 and are synchronous.
 */
#ifndef RAYON_LEGACY_H
#define RAYON_LEGACY_H

#include "rayon_hip.h"

#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Scene upload / release: the role of CudaSceneBuilder::buildGPUScene / freeGPUScene
 * (scene_builder.hpp:234-239). Returns an opaque scene handle, NULL on failure. */
void* rayon_legacy_scene_create(const rayon_scene_desc* desc);
void rayon_legacy_scene_destroy(void* scene);

/* renderer_cuda_device.cuh:12-13 */
void initCudaStreams(void);
void cleanupCudaStreams(void);

/* renderer_cuda_device.cuh:18-24 (min_adaptive_samples / adaptive_threshold have the
 * defaults 32 / 0.01 in the reference's C++ declaration). */
unsigned long long renderPixelsCUDAAccumulative(
    unsigned char* image, float* accum_buffer, void* scene, int width, int height, double cam_center_x,
    double cam_center_y, double cam_center_z, double pixel00_x, double pixel00_y, double pixel00_z, double delta_u_x,
    double delta_u_y, double delta_u_z, double delta_v_x, double delta_v_y, double delta_v_z, int samples_to_add,
    int total_samples_so_far, int max_depth, void** d_rand_states_ptr, void** d_accum_buffer_ptr, double cam_u_x,
    double cam_u_y, double cam_u_z, double cam_v_x, double cam_v_y, double cam_v_z, void* d_pixel_sample_counts,
    int min_adaptive_samples, float adaptive_threshold);

/* renderer_cuda_device.cuh:27, :30, :33 */
void freeDeviceRandomStates(void* d_rand_states);
void freeDeviceAccumBuffer(void* d_accum_buffer);
void resetDeviceAccumBuffer(void* d_accum_buffer, int num_pixels);

/* renderer_cuda_device.cuh:36-50 */
void setLightIntensity(float intensity);
void setBackgroundIntensity(float intensity);
void setMetalFuzziness(float fuzziness);
void setGlassRefractionIndex(float index);
void setDOFEnabled(bool enabled);
void setDOFAperture(float aperture);
void setDOFFocusDistance(float distance);

/* renderer_cuda_device.cuh:54-55 */
void convertAccumToDisplayCUDA(void* d_accum_buffer, unsigned char* display_image, int width, int height,
                               int channels, int num_samples, float gamma, void* d_pixel_sample_counts);

/* renderer_cuda_device.cuh:58-63 */
void allocateAdaptiveBuffer(void** d_pixel_sample_counts, int num_pixels);
void resetAdaptiveBuffer(void* d_pixel_sample_counts, int num_pixels);
void freeAdaptiveBuffer(void* d_pixel_sample_counts);
int countConvergedPixels(void* d_pixel_sample_counts, int num_pixels);

/* renderer_cuda_device.cuh:66-67 */
void renderSampleHeatmapCUDA(void* d_pixel_sample_counts, unsigned char* display_image, int width, int height,
                             int channels, int max_samples_for_scale);

#ifdef __cplusplus
}
#endif

#endif /* RAYON_LEGACY_H */
