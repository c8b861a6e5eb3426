// scene_builder_hip.cc — the scene upload of RayON's GPU render modes, for a RayON build that
// links librayon_hip.so in place of its CUDA objects. A RayON maintainer adds this file to
// src/rayon/scenes/ and drops scene_builder_cuda.cu from the build; renderer_cuda_host.hpp
// and the other callers of renderer_cuda_device.cuh stay as they are.
//
// It defines the two statics RayON declares in scenes/scene_builder.hpp:234-239
// (Scene::CudaSceneBuilder::buildGPUScene / freeGPUScene, implemented for CUDA in
// scene_builder_cuda.cu:148-303 and :308) by converting the SceneDescription to the plain
// rayon_scene_desc arrays of include/rayon_hip.h and creating the library's scene (which
// builds its own BVH). The returned handle is opaque; RayON's renderers only pass it on to
// renderPixelsCUDAAccumulative, whose C declaration in renderer_cuda_device.cuh matches the
// definition in librayon_hip.so (include/rayon_legacy.h).
#include <vector>

#include "rayon_legacy.h"
#include "scene_description.hpp"

namespace CudaScene {
struct Scene;  // opaque here: the handle is the library's scene
}

namespace Scene {

// The class of scene_builder.hpp:234-239, declared here as scene_builder_cuda.cu:18-22 does:
// scene_builder.hpp pulls in the CPU shapes and interval.hpp, whose static members are
// defined in the header, so a second translation unit including it would not link.
class CudaSceneBuilder {
 public:
   static CudaScene::Scene* buildGPUScene(const SceneDescription& desc);
   static void freeGPUScene(CudaScene::Scene* scene);
};

namespace {

void f3(float* dst, const Vec3& v) {
   dst[0] = (float)v.x();
   dst[1] = (float)v.y();
   dst[2] = (float)v.z();
}

}  // namespace

CudaScene::Scene* CudaSceneBuilder::buildGPUScene(const SceneDescription& sd) {
   std::vector<rayon_material> mats(sd.materials.size());
   for (size_t i = 0; i < mats.size(); ++i) {
      const MaterialDesc& m = sd.materials[i];
      rayon_material& r = mats[i];
      r = rayon_material{};
      r.type = (int32_t)m.type;  // RAYON_MAT_* follow MaterialType (scene_description.hpp:29-43)
      f3(r.albedo, m.albedo);
      f3(r.emission, m.emission);
      r.roughness = m.roughness;
      r.metallic = m.metallic;
      r.refractive_index = m.refractive_index;
      r.transmission = m.transmission;
      r.anisotropy = m.anisotropy;
      f3(r.eta, m.eta);
      f3(r.k, m.k);
      r.film_thickness = m.film_thickness;
      r.film_ior = m.film_ior;
      r.pattern = (int32_t)m.pattern;
      f3(r.pattern_color, m.pattern_color);
      r.pattern_param1 = m.pattern_param1;
      r.pattern_param2 = m.pattern_param2;
   }
   std::vector<rayon_geometry> geoms;
   geoms.reserve(sd.geometries.size());
   for (const GeometryDesc& g : sd.geometries) {
      rayon_geometry r{};
      r.material_id = g.material_id;
      r.visible = g.visible ? 1 : 0;
      switch (g.type) {
      case GeometryType::SPHERE:
         r.type = RAYON_GEOM_SPHERE;
         f3(r.p, g.data.sphere.center);
         r.p[3] = (float)g.data.sphere.radius;
         break;
      case GeometryType::DISPLACED_SPHERE:
         r.type = RAYON_GEOM_DISPLACED_SPHERE;
         f3(r.p, g.data.displaced_sphere.center);
         r.p[3] = (float)g.data.displaced_sphere.radius;
         break;
      case GeometryType::RECTANGLE:
         r.type = RAYON_GEOM_RECTANGLE;
         f3(r.p, g.data.rectangle.corner);
         f3(r.p + 3, g.data.rectangle.u);
         f3(r.p + 6, g.data.rectangle.v);
         break;
      case GeometryType::TRIANGLE:
         r.type = RAYON_GEOM_TRIANGLE;
         f3(r.p, g.data.triangle.v0);
         f3(r.p + 3, g.data.triangle.v1);
         f3(r.p + 6, g.data.triangle.v2);
         r.has_normals = g.data.triangle.has_normals ? 1 : 0;
         f3(r.n, g.data.triangle.n0);
         f3(r.n + 3, g.data.triangle.n1);
         f3(r.n + 6, g.data.triangle.n2);
         break;
      default:
         // CUBE, TRIANGLE_MESH and SDF_PRIMITIVE entries: the CUDA path does not intersect
         // them either (cuda_raytracer.cuh:353-375 returns no hit)
         continue;
      }
      geoms.push_back(r);
   }
   rayon_scene_desc d{};
   d.materials = mats.data();
   d.num_materials = (int32_t)mats.size();
   d.geometries = geoms.data();
   d.num_geometries = (int64_t)geoms.size();
   return static_cast<CudaScene::Scene*>(rayon_legacy_scene_create(&d));
}

void CudaSceneBuilder::freeGPUScene(CudaScene::Scene* scene) { rayon_legacy_scene_destroy(scene); }

}  // namespace Scene
