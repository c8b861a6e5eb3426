// generators.cpp — deterministic synthetic scenes (host) for the benchmark workloads.
//
//   RAYON_GEN_BOOK_SPHERES     ground r=1000, three r=1 spheres, jittered r=0.2 grid
//                              (the sphere field of the reference's book scene style).
//   RAYON_GEN_CORNELL_ICOSPHERE Cornell box (10 triangles) + quad light (radiance 15)
//                              around a value-noise-displaced icosphere.
//   RAYON_GEN_SDF_TERRAIN      smooth-union SDF primitives over a heightfield terrain.
//
// All randomness comes from a seeded splitmix64 stream, so a (kind, seed, detail) triple
// always yields the same scene.
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "generated.h"
#include "internal.h"

namespace ry {
void set_error(const char* fmt, ...);
}

namespace {

struct SplitMix {
   uint64_t s;
   uint64_t next() {
      uint64_t z = (s += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return z ^ (z >> 31);
   }
   double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
   double uni(double a, double b) { return a + (b - a) * uni(); }
};

rayon_material mat_base(int type, float r, float g, float b) {
   rayon_material m{};
   m.type = type;
   m.albedo[0] = r;
   m.albedo[1] = g;
   m.albedo[2] = b;
   m.refractive_index = 1.0f;
   m.film_thickness = 400.0f;
   m.film_ior = 1.33f;
   return m;
}

int add_mat(rayon_generated* g, const rayon_material& m) {
   g->mats.push_back(m);
   return (int)g->mats.size() - 1;
}

void add_sphere(rayon_generated* g, double x, double y, double z, double r, int mat) {
   rayon_geometry s{};
   s.type = RAYON_GEOM_SPHERE;
   s.material_id = mat;
   s.visible = 1;
   s.p[0] = (float)x;
   s.p[1] = (float)y;
   s.p[2] = (float)z;
   s.p[3] = (float)r;
   g->geoms.push_back(s);
}

void add_tri(rayon_generated* g, const double* a, const double* b, const double* c, int mat) {
   rayon_geometry t{};
   t.type = RAYON_GEOM_TRIANGLE;
   t.material_id = mat;
   t.visible = 1;
   for (int k = 0; k < 3; ++k) {
      t.p[k] = (float)a[k];
      t.p[3 + k] = (float)b[k];
      t.p[6 + k] = (float)c[k];
   }
   g->geoms.push_back(t);
}

// Book-style sphere field, seed 1984 by default.
void gen_book(rayon_generated* g, uint64_t seed) {
   SplitMix R{seed};
   int ground = add_mat(g, mat_base(RAYON_MAT_LAMBERTIAN, 0.5f, 0.5f, 0.5f));
   add_sphere(g, 0, -1000, 0, 1000, ground);
   for (int a = -11; a < 11; ++a) {
      for (int b = -11; b < 11; ++b) {
         double choose = R.uni();
         double cx = a + 0.9 * R.uni(), cz = b + 0.9 * R.uni();
         double dx = cx - 4, dz = cz - 0;
         if (std::sqrt(dx * dx + dz * dz) <= 0.9) continue;  // |c - (4,0.2,0)|, same y
         rayon_material m;
         if (choose < 0.8) {
            float c[3];
            for (int k = 0; k < 3; ++k) c[k] = (float)(R.uni() * R.uni());
            m = mat_base(RAYON_MAT_LAMBERTIAN, c[0], c[1], c[2]);
         } else if (choose < 0.95) {
            float c[3];
            for (int k = 0; k < 3; ++k) c[k] = (float)R.uni(0.5, 1.0);
            m = mat_base(RAYON_MAT_METAL, c[0], c[1], c[2]);
            m.roughness = (float)R.uni(0.0, 0.5);
         } else {
            m = mat_base(RAYON_MAT_DIELECTRIC, 1, 1, 1);
            m.refractive_index = 1.5f;
         }
         add_sphere(g, cx, 0.2, cz, 0.2, add_mat(g, m));
      }
   }
   rayon_material glass = mat_base(RAYON_MAT_DIELECTRIC, 1, 1, 1);
   glass.refractive_index = 1.5f;
   add_sphere(g, 0, 1, 0, 1.0, add_mat(g, glass));
   add_sphere(g, -4, 1, 0, 1.0, add_mat(g, mat_base(RAYON_MAT_LAMBERTIAN, 0.4f, 0.2f, 0.1f)));
   rayon_material metal = mat_base(RAYON_MAT_METAL, 0.7f, 0.6f, 0.5f);
   metal.roughness = 0.0f;
   add_sphere(g, 4, 1, 0, 1.0, add_mat(g, metal));
   const double lf[3] = {13, 2, 3}, la[3] = {0, 0, 0};
   std::memcpy(g->look_from, lf, sizeof lf);
   std::memcpy(g->look_at, la, sizeof la);
   g->vfov = 20.0;
}

// Value noise on the integer lattice: hashed corner values in [-1,1], smoothstep-weighted
// trilinear interpolation.
double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed) {
   SplitMix h{seed ^ ((uint64_t)x * 0x8DA6B343ull) ^ ((uint64_t)y * 0xD8163841ull) ^ ((uint64_t)z * 0xCB1AB31Full)};
   h.next();
   return 2.0 * h.uni() - 1.0;
}

double vnoise(double x, double y, double z, uint64_t seed) {
   double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
   int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
   double tx = x - fx, ty = y - fy, tz = z - fz;
   tx = tx * tx * (3 - 2 * tx);
   ty = ty * ty * (3 - 2 * ty);
   tz = tz * tz * (3 - 2 * tz);
   double c[2][2][2];
   for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
         for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
   auto lerp = [](double p, double q, double t) { return p + (q - p) * t; };
   double x00 = lerp(c[0][0][0], c[1][0][0], tx), x10 = lerp(c[0][1][0], c[1][1][0], tx);
   double x01 = lerp(c[0][0][1], c[1][0][1], tx), x11 = lerp(c[0][1][1], c[1][1][1], tx);
   return lerp(lerp(x00, x10, ty), lerp(x01, x11, ty), tz);
}

void icosphere(int levels, std::vector<double>& V, std::vector<int32_t>& F) {
   const double t = (1.0 + std::sqrt(5.0)) / 2.0;
   const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                               {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
   const int faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                             {11, 10, 2}, {10, 7, 6}, {7, 1, 8},   {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                             {3, 8, 9},  {4, 9, 5},  {2, 4, 11},  {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
   V.clear();
   F.clear();
   for (auto& v : base) {
      double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      V.push_back(v[0] / l);
      V.push_back(v[1] / l);
      V.push_back(v[2] / l);
   }
   for (auto& f : faces) F.insert(F.end(), f, f + 3);
   for (int lv = 0; lv < levels; ++lv) {
      std::unordered_map<uint64_t, int32_t> mid;
      mid.reserve(F.size());
      std::vector<int32_t> NF;
      NF.reserve(F.size() * 4);
      auto midpoint = [&](int32_t a, int32_t b) {
         uint64_t key = a < b ? ((uint64_t)a << 32 | (uint32_t)b) : ((uint64_t)b << 32 | (uint32_t)a);
         auto it = mid.find(key);
         if (it != mid.end()) return it->second;
         double m[3];
         for (int k = 0; k < 3; ++k) m[k] = 0.5 * (V[3 * (size_t)a + k] + V[3 * (size_t)b + k]);
         double l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
         int32_t id = (int32_t)(V.size() / 3);
         for (int k = 0; k < 3; ++k) V.push_back(m[k] / l);
         mid.emplace(key, id);
         return id;
      };
      for (size_t f = 0; f < F.size(); f += 3) {
         int32_t a = F[f], b = F[f + 1], c = F[f + 2];
         int32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
         const int32_t nf[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
         NF.insert(NF.end(), nf, nf + 12);
      }
      F.swap(NF);
   }
}

// Cornell box [0,555]^3 (5 walls as 10 triangles), a 130x105 quad light of radiance 15
// at y=554, an icosphere (center (278,200,278), r=150, `levels` subdivisions) displaced
// radially by r*0.05*vnoise(4*dir; seed 7), and a sphere (center (440,70,110), r=70) on the
// floor in front of it. One of the two objects is a GGX conductor (gold), the other a GGX
// dielectric (IOR 1.5), each with its own roughness U(0.05,0.8); which one is the
// icosphere and both roughnesses are drawn from seed+1.
void gen_cornell(rayon_generated* g, uint64_t seed, int levels) {
   g->background_intensity = 0.0f;  // the quad light is the only source
   int white = add_mat(g, mat_base(RAYON_MAT_LAMBERTIAN, 0.73f, 0.73f, 0.73f));
   int red = add_mat(g, mat_base(RAYON_MAT_LAMBERTIAN, 0.65f, 0.05f, 0.05f));
   int green = add_mat(g, mat_base(RAYON_MAT_LAMBERTIAN, 0.12f, 0.45f, 0.15f));
   rayon_material lm = mat_base(RAYON_MAT_LIGHT, 0, 0, 0);
   lm.emission[0] = lm.emission[1] = lm.emission[2] = 15.0f;
   int light = add_mat(g, lm);
   const double S = 555;
   auto quad = [&](const double* p0, const double* p1, const double* p2, const double* p3, int m) {
      add_tri(g, p0, p1, p2, m);
      add_tri(g, p0, p2, p3, m);
   };
   const double f0[3] = {0, 0, 0}, f1[3] = {S, 0, 0}, f2[3] = {S, 0, S}, f3[3] = {0, 0, S};
   quad(f0, f1, f2, f3, white);  // floor
   const double c0[3] = {0, S, 0}, c1[3] = {0, S, S}, c2[3] = {S, S, S}, c3[3] = {S, S, 0};
   quad(c0, c1, c2, c3, white);  // ceiling
   const double b0[3] = {0, 0, S}, b1[3] = {S, 0, S}, b2[3] = {S, S, S}, b3[3] = {0, S, S};
   quad(b0, b1, b2, b3, white);  // back
   const double g0[3] = {S, 0, 0}, g1[3] = {S, S, 0}, g2[3] = {S, S, S}, g3[3] = {S, 0, S};
   quad(g0, g1, g2, g3, green);  // x = 555
   const double r0[3] = {0, 0, 0}, r1[3] = {0, 0, S}, r2[3] = {0, S, S}, r3[3] = {0, S, 0};
   quad(r0, r1, r2, r3, red);  // x = 0
   rayon_geometry q{};
   q.type = RAYON_GEOM_RECTANGLE;
   q.material_id = light;
   q.visible = 1;
   const float qp[9] = {213, 554, 227, 130, 0, 0, 0, 0, 105};
   std::memcpy(q.p, qp, sizeof qp);
   g->geoms.push_back(q);

   SplitMix R{seed + 1};
   const bool conductor = R.uni() < 0.5;  // the icosphere's type; the sphere takes the other
   const float rough = (float)R.uni(0.05, 0.8);
   const float rough2 = (float)R.uni(0.05, 0.8);
   auto ggx = [&](bool cond, float r) {
      rayon_material m;
      if (cond) {
         m = mat_base(RAYON_MAT_GGX_CONDUCTOR, 1, 1, 1);
         const float eta[3] = {0.143f, 0.374f, 1.442f}, kk[3] = {3.983f, 2.385f, 1.603f};  // gold
         std::memcpy(m.eta, eta, sizeof eta);
         std::memcpy(m.k, kk, sizeof kk);
      } else {
         m = mat_base(RAYON_MAT_GGX_DIELECTRIC, 1, 1, 1);
         m.refractive_index = 1.5f;
      }
      m.roughness = r;
      return m;
   };
   int smat = add_mat(g, ggx(conductor, rough));
   rayon_geometry ball{};
   ball.type = RAYON_GEOM_SPHERE;
   ball.material_id = add_mat(g, ggx(!conductor, rough2));
   ball.visible = 1;
   const float bp[4] = {440, 70, 110, 70};
   std::memcpy(ball.p, bp, sizeof bp);
   g->geoms.push_back(ball);
   std::vector<double> V;
   std::vector<int32_t> F;
   icosphere(levels, V, F);
   const double C[3] = {278, 200, 278}, rad = 150;
   std::vector<float> P(V.size());
   for (size_t v = 0; v < V.size() / 3; ++v) {
      const double* d = &V[3 * v];
      double disp = rad * (1.0 + 0.05 * vnoise(4 * d[0], 4 * d[1], 4 * d[2], seed));
      for (int k = 0; k < 3; ++k) P[3 * v + k] = (float)(C[k] + disp * d[k]);
   }
   g->positions.push_back(std::move(P));
   g->normals.emplace_back();
   g->indices.push_back(std::move(F));
   g->mesh_mat.push_back(smat);
   const double lf[3] = {278, 278, -800}, la[3] = {278, 278, 0};
   std::memcpy(g->look_from, lf, sizeof lf);
   std::memcpy(g->look_at, la, sizeof la);
   g->vfov = 40.0;
}

// `count` SDF primitives (sphere r U(.2,.8) / rounded box half U(.2,.6) rounding .05 /
// torus R U(.3,.7) r U(.05,.2)) at U([0,10]^3), smooth union k = 0.1, 256 steps,
// eps 1e-4; plus an n x n heightfield (2(n-1)^2 triangles) of 5-octave value-noise fBm
// (seed+1) under it. Lambertian materials, sky light.
void gen_sdf_terrain(rayon_generated* g, uint64_t seed, int grid, int count) {
   SplitMix R{seed};
   int smat = add_mat(g, mat_base(RAYON_MAT_LAMBERTIAN, 0.8f, 0.55f, 0.35f));
   int tmat = add_mat(g, mat_base(RAYON_MAT_LAMBERTIAN, 0.35f, 0.45f, 0.3f));
   for (int i = 0; i < count; ++i) {
      rayon_sdf_prim p{};
      int type = (int)(R.uni() * 3.0);
      p.type = type > 2 ? 2 : type;
      for (int k = 0; k < 3; ++k) p.center[k] = (float)R.uni(0.0, 10.0);
      if (p.type == RAYON_SDF_SPHERE) {
         p.params[0] = (float)R.uni(0.2, 0.8);
      } else if (p.type == RAYON_SDF_ROUNDED_BOX) {
         for (int k = 0; k < 3; ++k) p.params[k] = (float)R.uni(0.2, 0.6);
         p.params[3] = 0.05f;
      } else {
         p.params[0] = (float)R.uni(0.3, 0.7);
         p.params[1] = (float)R.uni(0.05, 0.2);
      }
      g->sdf.push_back(p);
   }
   g->sdf_mat = smat;
   g->sdf_k = 0.1f;
   g->sdf_steps = 256;
   g->sdf_eps = 1e-4f;
   const int n = grid;
   std::vector<float> P(3 * (size_t)n * n);
   const double x0 = -5, x1 = 15;
   for (int j = 0; j < n; ++j) {
      for (int i = 0; i < n; ++i) {
         double x = x0 + (x1 - x0) * i / (n - 1), z = x0 + (x1 - x0) * j / (n - 1);
         double h = 0, amp = 0.5, fr = 0.25;
         for (int o = 0; o < 5; ++o) {
            h += amp * vnoise(x * fr, 0.0, z * fr, seed + 1);
            amp *= 0.5;
            fr *= 2;
         }
         float* p = &P[3 * ((size_t)j * n + i)];
         p[0] = (float)x;
         p[1] = (float)(-1.0 + 2.0 * h);
         p[2] = (float)z;
      }
   }
   std::vector<int32_t> F;
   F.reserve(6 * (size_t)(n - 1) * (n - 1));
   for (int j = 0; j + 1 < n; ++j) {
      for (int i = 0; i + 1 < n; ++i) {
         int32_t a = j * n + i, b = a + 1, c = a + n, d = c + 1;
         const int32_t t[6] = {a, c, b, b, c, d};
         F.insert(F.end(), t, t + 6);
      }
   }
   g->positions.push_back(std::move(P));
   g->normals.emplace_back();
   g->indices.push_back(std::move(F));
   g->mesh_mat.push_back(tmat);
   const double lf[3] = {5, 7, -11}, la[3] = {5, 3, 5};
   std::memcpy(g->look_from, lf, sizeof lf);
   std::memcpy(g->look_at, la, sizeof la);
   g->vfov = 50.0;
}

}  // namespace

extern "C" {

rayon_status rayon_generate(const rayon_gen_params* p, rayon_generated** out) {
   if (!p || !out) {
      ry::set_error("rayon_generate: NULL argument");
      return RAYON_ERR_INVALID;
   }
   auto* g = new rayon_generated();
   switch (p->kind) {
   case RAYON_GEN_BOOK_SPHERES:
      gen_book(g, p->seed);
      break;
   case RAYON_GEN_CORNELL_ICOSPHERE: {
      int lv = p->detail > 0 ? p->detail : 8;
      if (lv > 10) {
         delete g;
         ry::set_error("rayon_generate: icosphere subdivision %d > 10", lv);
         return RAYON_ERR_INVALID;
      }
      gen_cornell(g, p->seed, lv);
      break;
   }
   case RAYON_GEN_SDF_TERRAIN: {
      int grid = p->detail > 1 ? p->detail : 317;
      int count = p->count > 0 ? p->count : 64;
      if (grid > 8192 || count > 4096) {
         delete g;
         ry::set_error("rayon_generate: terrain grid %d or SDF count %d too large", grid, count);
         return RAYON_ERR_INVALID;
      }
      gen_sdf_terrain(g, p->seed, grid, count);
      break;
   }
   default:
      delete g;
      ry::set_error("rayon_generate: unknown kind %d", p->kind);
      return RAYON_ERR_INVALID;
   }
   g->finalize();
   *out = g;
   return RAYON_OK;
}

const rayon_scene_desc* rayon_generated_desc(const rayon_generated* g) { return g ? &g->desc : nullptr; }

rayon_status rayon_generated_camera(const rayon_generated* g, double lf[3], double la[3], double up[3],
                                    double* vfov) {
   if (!g) {
      ry::set_error("rayon_generated_camera: NULL handle");
      return RAYON_ERR_INVALID;
   }
   for (int k = 0; k < 3; ++k) {
      if (lf) lf[k] = g->look_from[k];
      if (la) la[k] = g->look_at[k];
      if (up) up[k] = g->vup[k];
   }
   if (vfov) *vfov = g->vfov;
   return RAYON_OK;
}

float rayon_generated_background_intensity(const rayon_generated* g) { return g ? g->background_intensity : 1.0f; }

void rayon_generated_free(rayon_generated* g) { delete g; }

// Defined with the YAML loader; declared here so this file links without it.
}
