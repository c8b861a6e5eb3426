// ref_yaml_dump.cpp — TEST INFRASTRUCTURE ONLY. Loads a scene file with the reference's own
// loader (Scene::loadSceneFromYAML, compiled from the reference sources where they lie by
// oracle/build_ref.sh into oracle/_ref/) and writes what it produced as a flat binary of
// doubles, so the tests can compare rayon_load_yaml with it field by field:
//
//   header   : n_materials, n_geometries, status (1 = loaded)
//   camera   : position[3], look_at[3], up[3], fov, background_color[3],
//              background_intensity, ambient_light, use_bvh, adaptive_sampling   (17)
//   material : type, albedo[3], emission[3], roughness, metallic, refractive_index,
//              transmission, anisotropy, eta[3], k[3], film_thickness, film_ior,
//              pattern, pattern_color[3], pattern_param1, pattern_param2          (26 each)
//   geometry : type, material_id, visible, has_normals, p[9], n[9]                 (22 each)
//
// p holds sphere/displaced sphere centre and radius, rectangle corner/u/v, or triangle
// v0/v1/v2; n holds triangle vertex normals.
#include <cstdio>
#include <vector>

#include "yaml_scene_loader.hpp"

using Scene::GeometryType;

static void put3(std::vector<double>& o, const Vec3& v) {
   o.push_back(v.x());
   o.push_back(v.y());
   o.push_back(v.z());
}

int main(int argc, char** argv) {
   if (argc != 3) {
      std::fprintf(stderr, "usage: %s scene.yaml out.bin\n", argv[0]);
      return 2;
   }
   Scene::SceneDescription s;
   const bool ok = Scene::loadSceneFromYAML(argv[1], s);
   std::vector<double> o = {(double)s.materials.size(), (double)s.geometries.size(), ok ? 1.0 : 0.0};
   put3(o, s.camera_position);
   put3(o, s.camera_look_at);
   put3(o, s.camera_up);
   o.push_back(s.camera_fov);
   put3(o, s.background_color);
   o.push_back(s.background_intensity);
   o.push_back(s.ambient_light);
   o.push_back(s.use_bvh ? 1 : 0);
   o.push_back(s.adaptive_sampling ? 1 : 0);
   for (const auto& m : s.materials) {
      o.push_back((double)(int)m.type);
      put3(o, m.albedo);
      put3(o, m.emission);
      o.push_back(m.roughness);
      o.push_back(m.metallic);
      o.push_back(m.refractive_index);
      o.push_back(m.transmission);
      o.push_back(m.anisotropy);
      put3(o, m.eta);
      put3(o, m.k);
      o.push_back(m.film_thickness);
      o.push_back(m.film_ior);
      o.push_back((double)(int)m.pattern);
      put3(o, m.pattern_color);
      o.push_back(m.pattern_param1);
      o.push_back(m.pattern_param2);
   }
   for (const auto& g : s.geometries) {
      double p[9] = {0}, n[9] = {0};
      double has_n = 0;
      switch (g.type) {
      case GeometryType::SPHERE:
         p[0] = g.data.sphere.center.x(), p[1] = g.data.sphere.center.y(), p[2] = g.data.sphere.center.z();
         p[3] = g.data.sphere.radius;
         break;
      case GeometryType::DISPLACED_SPHERE:
         p[0] = g.data.displaced_sphere.center.x(), p[1] = g.data.displaced_sphere.center.y();
         p[2] = g.data.displaced_sphere.center.z(), p[3] = g.data.displaced_sphere.radius;
         break;
      case GeometryType::RECTANGLE: {
         const Vec3* v[3] = {&g.data.rectangle.corner, &g.data.rectangle.u, &g.data.rectangle.v};
         for (int k = 0; k < 3; ++k) p[3 * k] = v[k]->x(), p[3 * k + 1] = v[k]->y(), p[3 * k + 2] = v[k]->z();
         break;
      }
      case GeometryType::TRIANGLE: {
         const Vec3* v[3] = {&g.data.triangle.v0, &g.data.triangle.v1, &g.data.triangle.v2};
         const Vec3* w[3] = {&g.data.triangle.n0, &g.data.triangle.n1, &g.data.triangle.n2};
         for (int k = 0; k < 3; ++k) p[3 * k] = v[k]->x(), p[3 * k + 1] = v[k]->y(), p[3 * k + 2] = v[k]->z();
         has_n = g.data.triangle.has_normals ? 1 : 0;
         if (has_n)
            for (int k = 0; k < 3; ++k) n[3 * k] = w[k]->x(), n[3 * k + 1] = w[k]->y(), n[3 * k + 2] = w[k]->z();
         break;
      }
      default:
         break;
      }
      o.push_back((double)(int)g.type);
      o.push_back(g.material_id);
      o.push_back(g.visible ? 1 : 0);
      o.push_back(has_n);
      o.insert(o.end(), p, p + 9);
      o.insert(o.end(), n, n + 9);
   }
   FILE* f = std::fopen(argv[2], "wb");
   if (!f) return 1;
   std::fwrite(o.data(), sizeof(double), o.size(), f);
   std::fclose(f);
   return 0;
}
