// generated.h — host container behind rayon_generated: a scene description with the
// arrays it points into, and the camera and settings that came with it (from a generator
// or a YAML scene file).
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/rayon_hip.h"

struct rayon_generated {
   std::vector<rayon_material> mats;
   std::vector<rayon_geometry> geoms;
   std::vector<std::vector<float>> positions;
   std::vector<std::vector<float>> normals;
   std::vector<std::vector<int32_t>> indices;
   std::vector<rayon_mesh> meshes;
   std::vector<rayon_sdf_prim> sdf;
   rayon_scene_desc desc{};
   double look_from[3] = {0, 0, 0}, look_at[3] = {0, 0, -1}, vup[3] = {0, 1, 0};
   double vfov = 90.0;
   float background_intensity = 1.0f;
   float background_color[3] = {0.5f, 0.7f, 1.0f};
   float ambient_light = 0.1f;
   int32_t use_bvh = 0, adaptive_sampling = 1;
   float sdf_k = 0.1f;
   int sdf_mat = 0, sdf_steps = 256;
   float sdf_eps = 1e-4f;

   void finalize() {
      meshes.clear();
      for (size_t i = 0; i < positions.size(); ++i) {
         rayon_mesh m{};
         m.positions = positions[i].data();
         m.normals = normals[i].empty() ? nullptr : normals[i].data();
         m.indices = indices[i].data();
         m.num_vertices = (int64_t)positions[i].size() / 3;
         m.num_triangles = (int64_t)indices[i].size() / 3;
         m.material_id = mesh_mat[i];
         m.visible = 1;
         meshes.push_back(m);
      }
      desc = {};
      desc.materials = mats.data();
      desc.num_materials = (int32_t)mats.size();
      desc.geometries = geoms.data();
      desc.num_geometries = (int64_t)geoms.size();
      desc.meshes = meshes.data();
      desc.num_meshes = (int32_t)meshes.size();
      desc.sdf_prims = sdf.data();
      desc.num_sdf_prims = (int32_t)sdf.size();
      desc.sdf_smooth_k = sdf_k;
      desc.sdf_material_id = sdf_mat;
      desc.sdf_max_steps = sdf_steps;
      desc.sdf_eps = sdf_eps;
   }
   std::vector<int> mesh_mat;
};

