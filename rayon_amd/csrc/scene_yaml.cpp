// scene_yaml.cpp — the scene file loader (and a lossless writer) of the reference's YAML
// subset, with its observable behaviour (yaml_scene_loader.cc:122-576, obj_loader.hpp:34-171,
// resources/scenes/SCENE_FORMAT.md):
//
//   * Lines are read one at a time: everything after the first '#' is dropped, blank lines
//     are skipped. "materials:", "geometry:"/"geometries:", "camera:" and "settings:" open
//     sections. Keys are flat: "key: value" is stored under "<section>.key", where the
//     section of a list item ("- ...") is "material<i>" or "geom<i>" in the two list
//     sections. Indentation is not interpreted, so a nested block such as "pattern:" does
//     not open a scope: its "type:" replaces the material's own type (which then falls back
//     to lambertian), and a block pattern is never applied. Flow maps "- { k: v, ... }" are
//     split at top-level commas. Empty values are not stored; a later key replaces an
//     earlier one.
//   * Vectors are "[x, y, z, ...]" (the outer characters are dropped, the rest split at
//     commas; fewer than three numbers give (0,0,0), a value without '[' too). Numbers
//     parse their longest valid prefix as the C library does; text that starts no number
//     is a parse error (the reference would abort).
//   * At most 100 materials and 1000 geometry entries are read (RAYON_YAML_LIFT_CAPS lifts
//     both). A geometry naming an unknown material is skipped; an unknown geometry type adds
//     nothing. The entry's "visible" flag is then applied to the last geometry of the scene
//     so far (for an OBJ file, its last triangle).
//   * OBJ: "v", "vn", "vt" and "f" records; faces are fanned into triangles; indices are
//     1-based, negative ones count back from the records read so far, and a missing normal
//     index reads as index 0. Triangles with an index out of range are dropped; a triangle
//     takes vertex normals when the file has normals and all three indices are in range.
//     Positions are scaled per axis, then translated; normals are used as read. Paths are
//     relative to the scene file's directory.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "generated.h"
#include "internal.h"

namespace {

struct ParseError {
   std::string msg;
};

std::string trim(const std::string& s) {
   const char* ws = " \t\r\n";
   const size_t a = s.find_first_not_of(ws);
   if (a == std::string::npos) return "";
   const size_t b = s.find_last_not_of(ws);
   return s.substr(a, b - a + 1);
}

std::string unquote(const std::string& v) {
   std::string s = trim(v);
   if (s.size() >= 2 && s.front() == '"' && s.back() == '"') return s.substr(1, s.size() - 2);
   return s;
}

// Longest numeric prefix (leading white space allowed), as strtod/strtof/strtol take it.
double to_double(const std::string& v, const std::string& key) {
   const char* b = v.c_str();
   char* e = nullptr;
   const double x = std::strtod(b, &e);
   if (e == b) throw ParseError{"'" + key + "': '" + v + "' is not a number"};
   return x;
}

float to_float(const std::string& v, const std::string& key) {
   const char* b = v.c_str();
   char* e = nullptr;
   const float x = std::strtof(b, &e);
   if (e == b) throw ParseError{"'" + key + "': '" + v + "' is not a number"};
   return x;
}

int to_int(const std::string& v, const std::string& key) {
   const char* b = v.c_str();
   char* e = nullptr;
   const long x = std::strtol(b, &e, 10);
   if (e == b) throw ParseError{"'" + key + "': '" + v + "' is not an integer"};
   return (int)x;
}

struct Vec {
   double x = 0, y = 0, z = 0;
};

Vec to_vec(const std::string& v, const std::string& key) {
   const std::string c = trim(v);
   if (c.empty() || c[0] != '[') return {};
   const std::string inner = c.substr(1, c.size() >= 2 ? c.size() - 2 : 0);
   std::vector<double> xs;
   size_t pos = 0;
   while (pos < inner.size()) {
      size_t comma = inner.find(',', pos);
      const std::string tok = trim(inner.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos));
      xs.push_back(to_double(tok, key));
      if (comma == std::string::npos) break;
      pos = comma + 1;
   }
   if (xs.size() < 3) return {};
   return {xs[0], xs[1], xs[2]};
}

// The flat key store of a scene file.
class Keys {
 public:
   void read(std::istream& in) {
      std::string line;
      std::string section;  // "" | "camera" | "settings" | "material<i>" | "geom<i>"
      int list = 0;         // 0 none, 1 materials, 2 geometry
      int nmat = 0, ngeo = 0;
      while (std::getline(in, line)) {
         const size_t hash = line.find('#');
         if (hash != std::string::npos) line.resize(hash);
         std::string t = trim(line);
         if (t.empty()) continue;
         if (t == "materials:") {
            list = 1;
            section.clear();
            nmat = 0;
            continue;
         }
         if (t == "geometry:" || t == "geometries:") {
            list = 2;
            section.clear();
            ngeo = 0;
            continue;
         }
         if (t == "camera:" || t == "settings:") {
            list = 0;
            section = t.substr(0, t.size() - 1);
            continue;
         }
         if (t.size() >= 2 && t[0] == '-' && t[1] == ' ') {
            t = trim(t.substr(2));
            if (list == 1) section = "material" + std::to_string(nmat++);
            if (list == 2) section = "geom" + std::to_string(ngeo++);
            if (!t.empty() && t.front() == '{' && t.back() == '}') {
               const std::string inner = t.substr(1, t.size() - 2);
               int depth = 0;
               std::string cur;
               std::vector<std::string> parts;
               for (char ch : inner) {
                  if (ch == '[') ++depth;
                  if (ch == ']') --depth;
                  if (ch == ',' && depth == 0) {
                     parts.push_back(cur);
                     cur.clear();
                  } else {
                     cur += ch;
                  }
               }
               if (!cur.empty()) parts.push_back(cur);
               for (const std::string& p : parts) {
                  const size_t c = p.find(':');
                  if (c != std::string::npos) put(section, trim(p.substr(0, c)), trim(p.substr(c + 1)));
               }
               continue;
            }
         }
         const size_t c = t.find(':');
         if (c != std::string::npos) put(section, trim(t.substr(0, c)), trim(t.substr(c + 1)));
      }
   }

   bool has(const std::string& k) const { return kv_.count(k) != 0; }
   std::string str(const std::string& k, const std::string& dflt = "") const {
      auto it = kv_.find(k);
      return it == kv_.end() ? dflt : it->second;
   }
   float f(const std::string& k, float dflt) const { return has(k) ? to_float(str(k), k) : dflt; }
   double d(const std::string& k, double dflt) const { return has(k) ? to_double(str(k), k) : dflt; }
   int i(const std::string& k, int dflt) const { return has(k) ? to_int(str(k), k) : dflt; }
   Vec v(const std::string& k, Vec dflt = {}) const { return has(k) ? to_vec(str(k), k) : dflt; }

 private:
   void put(const std::string& section, const std::string& key, const std::string& value) {
      if (value.empty()) return;
      kv_[section.empty() ? key : section + "." + key] = value;
   }
   std::map<std::string, std::string> kv_;
};

int material_type(const std::string& s) {
   static const std::map<std::string, int> names = {
      {"lambertian", RAYON_MAT_LAMBERTIAN},   {"metal", RAYON_MAT_METAL},
      {"mirror", RAYON_MAT_MIRROR},           {"rough_mirror", RAYON_MAT_ROUGH_MIRROR},
      {"glass", RAYON_MAT_GLASS},             {"dielectric", RAYON_MAT_DIELECTRIC},
      {"light", RAYON_MAT_LIGHT},             {"constant", RAYON_MAT_CONSTANT},
      {"show_normals", RAYON_MAT_SHOW_NORMALS}, {"anisotropic_metal", RAYON_MAT_ANISOTROPIC_METAL},
      {"thin_film", RAYON_MAT_THIN_FILM},     {"clear_coat", RAYON_MAT_CLEAR_COAT},
   };
   auto it = names.find(s);
   return it == names.end() ? RAYON_MAT_LAMBERTIAN : it->second;
}

// Pattern numbering of the reference's ProceduralPattern enum.
int pattern_type(const std::string& s) {
   if (s == "fibonacci_dots") return 1;
   if (s == "checkerboard") return 2;
   if (s == "stripes") return 3;
   return 0;
}

void set3(float* dst, Vec v) {
   dst[0] = (float)v.x;
   dst[1] = (float)v.y;
   dst[2] = (float)v.z;
}

// Measured complex indices of refraction (RGB) of the conductor presets of
// anisotropic_metal materials (yaml_scene_loader.cc:365-378).
struct Preset {
   const char* name;
   Vec eta, k;
};
const Preset kPresets[] = {
   {"gold", {0.18, 0.42, 1.37}, {3.42, 2.35, 1.77}},
   {"silver", {0.05, 0.06, 0.05}, {4.18, 3.35, 2.58}},
   {"copper", {0.27, 0.68, 1.22}, {3.60, 2.63, 2.29}},
   {"aluminum", {1.35, 0.97, 0.53}, {7.47, 6.40, 5.28}},
};

rayon_material read_material(const Keys& K, const std::string& p) {
   rayon_material m{};
   m.type = material_type(unquote(K.str(p + ".type", "lambertian")));
   set3(m.albedo, K.v(p + ".albedo", {0.7, 0.7, 0.7}));
   if (K.has(p + ".color") && K.has(p + ".emission_intensity")) {
      const Vec c = K.v(p + ".color", {1, 1, 1});
      const double s = K.f(p + ".emission_intensity", 1.0f);
      set3(m.emission, {c.x * s, c.y * s, c.z * s});
   } else {
      set3(m.emission, K.v(p + ".emission"));
   }
   m.roughness = K.f(p + ".roughness", 0.0f);
   m.metallic = K.f(p + ".metallic", 0.0f);
   m.refractive_index = K.f(p + ".refractive_index", 1.0f);
   m.transmission = K.f(p + ".transmission", 0.0f);
   m.anisotropy = K.f(p + ".anisotropy", 0.0f);
   set3(m.eta, K.v(p + ".eta"));
   set3(m.k, K.v(p + ".k"));
   m.film_thickness = K.f(p + ".film_thickness", 400.0f);
   m.film_ior = K.f(p + ".film_ior", 1.33f);
   if (m.type == RAYON_MAT_ANISOTROPIC_METAL && K.has(p + ".preset")) {
      const std::string name = unquote(K.str(p + ".preset"));
      for (const Preset& pr : kPresets)
         if (name == pr.name) {
            set3(m.eta, pr.eta);
            set3(m.k, pr.k);
         }
   }
   if (K.has(p + ".pattern.type")) {
      m.pattern = pattern_type(unquote(K.str(p + ".pattern.type")));
      set3(m.pattern_color, K.v(p + ".pattern.color"));
      m.pattern_param1 = K.f(p + ".pattern.dot_count", 0.0f);
      m.pattern_param2 = K.f(p + ".pattern.dot_radius", 0.0f);
   }
   return m;
}

void add_triangle(rayon_generated* g, const Vec* v, const Vec* n, int mat) {
   rayon_geometry t{};
   t.type = RAYON_GEOM_TRIANGLE;
   t.material_id = mat;
   t.visible = 1;
   for (int k = 0; k < 3; ++k) set3(t.p + 3 * k, v[k]);
   if (n) {
      t.has_normals = 1;
      for (int k = 0; k < 3; ++k) set3(t.n + 3 * k, n[k]);
   }
   g->geoms.push_back(t);
}

// One face-vertex token "v", "v/vt", "v/vt/vn" or "v//vn"; absent parts read as 0.
void face_vertex(const std::string& tok, int& vi, int& vni) {
   vi = vni = 0;
   size_t a = tok.find('/');
   const std::string p0 = tok.substr(0, a);
   if (!p0.empty()) vi = to_int(p0, "f");
   if (a == std::string::npos) return;
   size_t b = tok.find('/', a + 1);
   const std::string p1 = tok.substr(a + 1, b == std::string::npos ? std::string::npos : b - a - 1);
   if (!p1.empty()) (void)to_int(p1, "f");
   if (b == std::string::npos) return;
   size_t c = tok.find('/', b + 1);
   const std::string p2 = tok.substr(b + 1, c == std::string::npos ? std::string::npos : c - b - 1);
   if (!p2.empty()) vni = to_int(p2, "f");
}

// Returns the number of triangles added, -1 if the file cannot be opened.
long load_obj(rayon_generated* g, const std::string& path, int mat, Vec pos, Vec scale) {
   std::ifstream f(path);
   if (!f.is_open()) return -1;
   std::vector<Vec> vs, ns;
   long count = 0;
   std::string line;
   while (std::getline(f, line)) {
      if (line.empty() || line[0] == '#') continue;
      std::istringstream in(line);
      std::string rec;
      in >> rec;
      if (rec == "v") {
         double x, y, z;
         if (in >> x >> y >> z) vs.push_back({x * scale.x + pos.x, y * scale.y + pos.y, z * scale.z + pos.z});
      } else if (rec == "vn") {
         double x, y, z;
         if (in >> x >> y >> z) ns.push_back({x, y, z});
      } else if (rec == "f") {
         std::vector<int> vi, ni;
         std::string tok;
         while (in >> tok) {
            int a, c;
            face_vertex(tok, a, c);
            if (a > 0) a -= 1;
            else if (a < 0) a += (int)vs.size();
            if (c > 0) c -= 1;
            else if (c < 0) c += (int)ns.size();
            vi.push_back(a);
            ni.push_back(c);
         }
         const int nv = (int)vs.size(), nn = (int)ns.size();
         for (size_t k = 1; k + 1 < vi.size(); ++k) {
            const int i0 = vi[0], i1 = vi[k], i2 = vi[k + 1];
            if (i0 < 0 || i0 >= nv || i1 < 0 || i1 >= nv || i2 < 0 || i2 >= nv) continue;
            const Vec v[3] = {vs[i0], vs[i1], vs[i2]};
            const bool normals = nn > 0 && ni[0] >= 0 && ni[0] < nn && ni[k] >= 0 && ni[k] < nn && ni[k + 1] >= 0 &&
                                 ni[k + 1] < nn;
            if (normals) {
               const Vec n[3] = {ns[ni[0]], ns[ni[k]], ns[ni[k + 1]]};
               add_triangle(g, v, n, mat);
            } else {
               add_triangle(g, v, nullptr, mat);
            }
            ++count;
         }
      }
   }
   return count;
}

void load(rayon_generated* g, const char* path, uint32_t flags) {
   std::ifstream in(path);
   if (!in.is_open()) throw ParseError{std::string("cannot open ") + path};
   Keys K;
   K.read(in);
   const bool lift = (flags & RAYON_YAML_LIFT_CAPS) != 0;
   if (K.has("camera.position")) {
      const Vec v = K.v("camera.position");
      g->look_from[0] = v.x, g->look_from[1] = v.y, g->look_from[2] = v.z;
   }
   if (K.has("camera.look_at")) {
      const Vec v = K.v("camera.look_at");
      g->look_at[0] = v.x, g->look_at[1] = v.y, g->look_at[2] = v.z;
   }
   if (K.has("camera.up")) {
      const Vec v = K.v("camera.up");
      g->vup[0] = v.x, g->vup[1] = v.y, g->vup[2] = v.z;
   }
   if (K.has("camera.fov")) g->vfov = K.f("camera.fov", 90.0f);
   if (K.has("settings.background_color")) set3(g->background_color, K.v("settings.background_color"));
   if (K.has("settings.ambient_light")) g->ambient_light = K.f("settings.ambient_light", 0.1f);
   if (K.has("settings.background_intensity")) g->background_intensity = K.f("settings.background_intensity", 1.0f);
   if (K.has("settings.use_bvh")) {
      const std::string v = unquote(K.str("settings.use_bvh"));
      g->use_bvh = (v == "true" || v == "1");
   }
   if (K.has("settings.adaptive_sampling")) {
      const std::string v = unquote(K.str("settings.adaptive_sampling"));
      g->adaptive_sampling = (v == "true" || v == "1");
   }
   std::string dir = ".";
   {
      const std::string f(path);
      const size_t s = f.find_last_of("/\\");
      if (s != std::string::npos) dir = f.substr(0, s);
   }
   std::map<std::string, int> mat_id;
   const int max_mat = lift ? std::numeric_limits<int>::max() : 100;
   for (int i = 0; i < max_mat; ++i) {
      const std::string p = "material" + std::to_string(i);
      if (!K.has(p + ".name")) break;
      g->mats.push_back(read_material(K, p));
      mat_id[unquote(K.str(p + ".name"))] = (int)g->mats.size() - 1;
   }
   if (g->mats.empty()) throw ParseError{"no materials"};
   const int max_geo = lift ? std::numeric_limits<int>::max() : 1000;
   for (int i = 0; i < max_geo; ++i) {
      const std::string p = "geom" + std::to_string(i);
      if (!K.has(p + ".type")) break;
      const std::string type = unquote(K.str(p + ".type"));
      auto it = mat_id.find(unquote(K.str(p + ".material")));
      if (it == mat_id.end()) continue;  // unknown material: entry skipped
      const int m = it->second;
      const std::string vis = unquote(K.str(p + ".visible", "true"));
      const bool visible = vis != "false" && vis != "0";
      rayon_geometry gd{};
      gd.material_id = m;
      gd.visible = 1;
      if (type == "sphere" || type == "displaced_sphere") {
         gd.type = type == "sphere" ? RAYON_GEOM_SPHERE : RAYON_GEOM_DISPLACED_SPHERE;
         set3(gd.p, K.v(p + ".center"));
         gd.p[3] = (float)K.d(p + ".radius", 1.0);
         if (type == "displaced_sphere") {  // read as the reference reads them; the shading uses fixed values
            (void)K.f(p + ".displacement_scale", 0.2f);
            (void)K.i(p + ".pattern_type", 0);
         }
         g->geoms.push_back(gd);
      } else if (type == "rectangle") {
         gd.type = RAYON_GEOM_RECTANGLE;
         set3(gd.p, K.v(p + ".corner"));
         set3(gd.p + 3, K.v(p + ".u"));
         set3(gd.p + 6, K.v(p + ".v"));
         g->geoms.push_back(gd);
      } else if (type == "triangle") {
         const Vec v[3] = {K.v(p + ".v0"), K.v(p + ".v1"), K.v(p + ".v2")};
         if (K.has(p + ".n0")) {
            const Vec n[3] = {K.v(p + ".n0"), K.v(p + ".n1"), K.v(p + ".n2")};
            add_triangle(g, v, n, m);
         } else {
            add_triangle(g, v, nullptr, m);
         }
      } else if (type == "obj") {
         const std::string file = unquote(K.str(p + ".file"));
         const Vec pos = K.v(p + ".position", {0, 0, 0});
         const Vec scale = K.v(p + ".scale", {1, 1, 1});
         const std::string full = (!file.empty() && file[0] != '/') ? dir + "/" + file : file;
         if (load_obj(g, full, m, pos, scale) < 0) fprintf(stderr, "rayon_load_yaml: cannot open OBJ file %s\n", full.c_str());
      }
      if (!g->geoms.empty()) g->geoms.back().visible = visible ? 1 : 0;
   }
   if (g->geoms.empty()) throw ParseError{"no geometry"};
}

// Shortest decimal that reads back to the same float.
std::string num(double v) {
   char buf[64];
   snprintf(buf, sizeof buf, "%.9g", v);
   return buf;
}

std::string vec3(const float* v) { return "[" + num(v[0]) + ", " + num(v[1]) + ", " + num(v[2]) + "]"; }

const char* type_name(int t) {
   static const char* names[] = {"lambertian", "metal",        "mirror",  "rough_mirror",
                                 "glass",      "dielectric",   "light",   "constant",
                                 "show_normals", nullptr,      "anisotropic_metal", "thin_film",
                                 "clear_coat"};
   return (t >= 0 && t <= RAYON_MAT_CLEAR_COAT) ? names[t] : nullptr;
}

}  // namespace

extern "C" {

rayon_status rayon_load_yaml_ex(const char* path, uint32_t flags, rayon_generated** out) {
   if (!path || !out) {
      ry::set_error("rayon_load_yaml: NULL argument");
      return RAYON_ERR_INVALID;
   }
   auto* g = new rayon_generated();
   try {
      load(g, path, flags);
   } catch (const ParseError& e) {
      delete g;
      ry::set_error("rayon_load_yaml: %s: %s", path, e.msg.c_str());
      return e.msg.rfind("cannot open", 0) == 0 ? RAYON_ERR_IO : RAYON_ERR_PARSE;
   } catch (const std::exception& e) {
      delete g;
      ry::set_error("rayon_load_yaml: %s: %s", path, e.what());
      return RAYON_ERR_PARSE;
   }
   g->finalize();
   *out = g;
   return RAYON_OK;
}

rayon_status rayon_load_yaml(const char* path, rayon_generated** out) { return rayon_load_yaml_ex(path, 0, out); }

rayon_status rayon_save_yaml(const rayon_generated* g, const char* path) {
   if (!g || !path) {
      ry::set_error("rayon_save_yaml: NULL argument");
      return RAYON_ERR_INVALID;
   }
   if (!g->meshes.empty() || !g->sdf.empty()) {
      ry::set_error("rayon_save_yaml: indexed meshes and SDF fields have no form in the scene format");
      return RAYON_ERR_UNSUPPORTED;
   }
   std::ofstream f(path);
   if (!f.is_open()) {
      ry::set_error("rayon_save_yaml: cannot open %s for writing", path);
      return RAYON_ERR_IO;
   }
   const float lf[3] = {(float)g->look_from[0], (float)g->look_from[1], (float)g->look_from[2]};
   const float la[3] = {(float)g->look_at[0], (float)g->look_at[1], (float)g->look_at[2]};
   const float up[3] = {(float)g->vup[0], (float)g->vup[1], (float)g->vup[2]};
   f << "camera:\n  position: " << vec3(lf) << "\n  look_at: " << vec3(la) << "\n  up: " << vec3(up)
     << "\n  fov: " << num(g->vfov) << "\n\n";
   f << "settings:\n  background_color: " << vec3(g->background_color)
     << "\n  background_intensity: " << num(g->background_intensity) << "\n  ambient_light: " << num(g->ambient_light)
     << "\n  use_bvh: " << (g->use_bvh ? "true" : "false")
     << "\n  adaptive_sampling: " << (g->adaptive_sampling ? "true" : "false") << "\n\n";
   f << "materials:\n";
   for (size_t i = 0; i < g->mats.size(); ++i) {
      const rayon_material& m = g->mats[i];
      const char* tn = type_name(m.type);
      if (!tn) {
         ry::set_error("rayon_save_yaml: material %zu has type %d, which the scene format cannot name", i, m.type);
         return RAYON_ERR_UNSUPPORTED;
      }
      // flow form, so the pattern keys stay the material's own (a block "pattern:" would not)
      f << "  - { name: m" << i << ", type: " << tn << ", albedo: " << vec3(m.albedo) << ", emission: " << vec3(m.emission)
        << ", roughness: " << num(m.roughness) << ", metallic: " << num(m.metallic)
        << ", refractive_index: " << num(m.refractive_index) << ", transmission: " << num(m.transmission)
        << ", anisotropy: " << num(m.anisotropy) << ", eta: " << vec3(m.eta) << ", k: " << vec3(m.k)
        << ", film_thickness: " << num(m.film_thickness) << ", film_ior: " << num(m.film_ior);
      if (m.pattern > 0 && m.pattern <= 3) {
         static const char* pn[] = {"", "fibonacci_dots", "checkerboard", "stripes"};
         f << ", pattern.type: " << pn[m.pattern] << ", pattern.color: " << vec3(m.pattern_color)
           << ", pattern.dot_count: " << num(m.pattern_param1) << ", pattern.dot_radius: " << num(m.pattern_param2);
      }
      f << " }\n";
   }
   f << "\ngeometry:\n";
   for (const rayon_geometry& gd : g->geoms) {
      f << "  - { material: m" << gd.material_id << ", visible: " << (gd.visible ? "true" : "false") << ", ";
      switch (gd.type) {
      case RAYON_GEOM_SPHERE:
      case RAYON_GEOM_DISPLACED_SPHERE:
         f << "type: " << (gd.type == RAYON_GEOM_SPHERE ? "sphere" : "displaced_sphere") << ", center: " << vec3(gd.p)
           << ", radius: " << num(gd.p[3]);
         break;
      case RAYON_GEOM_RECTANGLE:
         f << "type: rectangle, corner: " << vec3(gd.p) << ", u: " << vec3(gd.p + 3) << ", v: " << vec3(gd.p + 6);
         break;
      default:
         f << "type: triangle, v0: " << vec3(gd.p) << ", v1: " << vec3(gd.p + 3) << ", v2: " << vec3(gd.p + 6);
         if (gd.has_normals)
            f << ", n0: " << vec3(gd.n) << ", n1: " << vec3(gd.n + 3) << ", n2: " << vec3(gd.n + 6);
         break;
      }
      f << " }\n";
   }
   return f.good() ? RAYON_OK : RAYON_ERR_IO;
}

rayon_status rayon_generated_settings(const rayon_generated* g, float background_color[3], float* background_intensity,
                                      float* ambient_light, int32_t* use_bvh, int32_t* adaptive_sampling) {
   if (!g) {
      ry::set_error("rayon_generated_settings: NULL handle");
      return RAYON_ERR_INVALID;
   }
   if (background_color)
      for (int k = 0; k < 3; ++k) background_color[k] = g->background_color[k];
   if (background_intensity) *background_intensity = g->background_intensity;
   if (ambient_light) *ambient_light = g->ambient_light;
   if (use_bvh) *use_bvh = g->use_bvh;
   if (adaptive_sampling) *adaptive_sampling = g->adaptive_sampling;
   return RAYON_OK;
}

}  // extern "C"
