// ref_cuda_host_drive.cpp — TEST INFRASTRUCTURE ONLY. Runs the reference's own offline GPU
// renderer, RendererCUDA (gpu_renderers/renderer_cuda_host.hpp:23-74, included unchanged from
// the reference sources by oracle/build_ref.sh), linked against librayon_hip.so through the
// scene shim rayon_amd/integration/scene_builder_hip.cc. It loads a scene file with the
// reference's loader, sets up the camera as main.cc:367-373 does, renders, and writes the
// 8-bit RGB image (raw bytes) and the ray count the renderer reports.
//
//   usage: ref_cuda_host_drive scene.yaml width height spp out.rgb
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gpu_renderers/renderer_cuda_host.hpp"
#include "yaml_scene_loader.hpp"

int main(int argc, char** argv) {
   if (argc != 6) {
      std::fprintf(stderr, "usage: %s scene.yaml width height spp out.rgb\n", argv[0]);
      return 2;
   }
   Scene::SceneDescription scene;
   if (!Scene::loadSceneFromYAML(argv[1], scene)) {
      std::fprintf(stderr, "scene %s did not load\n", argv[1]);
      return 1;
   }
   const int W = std::atoi(argv[2]), H = std::atoi(argv[3]), spp = std::atoi(argv[4]);
   CameraBase camera(Vec3(0, 0, 0), W, H, 3, spp);
   camera.look_from = scene.camera_position;
   camera.look_at = scene.camera_look_at;
   camera.vup = scene.camera_up;
   camera.vfov = scene.camera_fov;
   camera.updateFrame();
   std::vector<unsigned char> image((size_t)W * H * 3);
   RenderTargetView target{&image, W, H, 3};
   std::atomic<long long> rays{0};
   RenderContext context{rays, 2.0f};
   RenderRequest request{camera, scene, target};
   RendererCUDA renderer;
   renderer.render(request, context);
   FILE* f = std::fopen(argv[5], "wb");
   if (!f || std::fwrite(image.data(), 1, image.size(), f) != image.size()) {
      std::fprintf(stderr, "cannot write %s\n", argv[5]);
      return 1;
   }
   std::fclose(f);
   std::printf("rays %lld\n", (long long)rays.load());
   return 0;
}
