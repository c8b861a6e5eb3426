// output.cpp — the files a finished render leaves behind, as the reference's command-line
// renderer writes them: the image as PNG (FileUtils::writeImage, utils.hpp:34-48, which
// hands RGB bytes to stb) in a directory created on demand (ensureDirectoryExists,
// utils.hpp:22-31), and the statistics as .txt and .json beside it (writeRenderStats,
// utils.hpp:118-172, with formatDuration, utils.hpp:90-116).
//
// The PNG is written here: per-row filter chosen by the smallest sum of absolute filtered
// bytes (the usual heuristic for photographic images), the filtered rows deflated by zlib,
// and the four chunks (IHDR, IDAT, IEND) with their CRCs. The pixels decode to exactly the
// bytes given; the file is not byte-identical to stb's (its own deflate), so
// image_size_bytes differs from the reference's for the same image.
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "internal.h"

namespace {

void put_be32(std::vector<uint8_t>& o, uint32_t v) {
   o.push_back((uint8_t)(v >> 24));
   o.push_back((uint8_t)(v >> 16));
   o.push_back((uint8_t)(v >> 8));
   o.push_back((uint8_t)v);
}

void put_chunk(std::vector<uint8_t>& o, const char type[4], const uint8_t* data, size_t n) {
   put_be32(o, (uint32_t)n);
   const size_t at = o.size();
   o.insert(o.end(), type, type + 4);
   if (n) o.insert(o.end(), data, data + n);
   const uLong crc = crc32(crc32(0L, Z_NULL, 0), o.data() + at, (uInt)(n + 4));
   put_be32(o, (uint32_t)crc);
}

// Paeth predictor (PNG specification, filter type 4).
inline int paeth(int a, int b, int c) {
   const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
   if (pa <= pb && pa <= pc) return a;
   return pb <= pc ? b : c;
}

// Filter byte + filtered row for each row; the filter of a row is the one whose output has
// the smallest sum of absolute values read as signed bytes.
std::vector<uint8_t> filter_rows(const uint8_t* px, int w, int h, int bpp) {
   const size_t stride = (size_t)w * bpp;
   std::vector<uint8_t> out((stride + 1) * (size_t)h);
   std::vector<uint8_t> cand(5 * stride);
   for (int y = 0; y < h; ++y) {
      const uint8_t* row = px + (size_t)y * stride;
      const uint8_t* up = y > 0 ? row - stride : nullptr;
      long best = -1;
      int best_f = 0;
      for (int f = 0; f < 5; ++f) {
         uint8_t* c = cand.data() + (size_t)f * stride;
         long sum = 0;
         for (size_t i = 0; i < stride; ++i) {
            const int a = i >= (size_t)bpp ? row[i - bpp] : 0;
            const int b = up ? up[i] : 0;
            const int cc = (up && i >= (size_t)bpp) ? up[i - bpp] : 0;
            int pred = 0;
            switch (f) {
            case 1: pred = a; break;
            case 2: pred = b; break;
            case 3: pred = (a + b) >> 1; break;
            case 4: pred = paeth(a, b, cc); break;
            default: break;
            }
            const uint8_t v = (uint8_t)(row[i] - pred);
            c[i] = v;
            sum += std::abs((int)(int8_t)v);
         }
         if (best < 0 || sum < best) {
            best = sum;
            best_f = f;
         }
      }
      uint8_t* dst = out.data() + (size_t)y * (stride + 1);
      dst[0] = (uint8_t)best_f;
      std::memcpy(dst + 1, cand.data() + (size_t)best_f * stride, stride);
   }
   return out;
}

bool ensure_parent_dir(const std::string& path) {
   std::error_code ec;
   const std::filesystem::path parent = std::filesystem::path(path).parent_path();
   if (parent.empty()) return true;
   std::filesystem::create_directories(parent, ec);
   return !ec || std::filesystem::is_directory(parent);
}

// std::filesystem::path::replace_extension of the reference's stats paths.
std::string with_extension(const std::string& path, const char* ext) {
   std::filesystem::path p(path);
   p.replace_extension(ext);
   return p.string();
}

std::string format_duration(int64_t ns) {
   // duration_cast truncates toward zero, as in formatDuration
   const long long ms = ns / 1000000, s = ns / 1000000000, m = ns / 60000000000ll;
   char buf[64];
   if (m > 0)
      std::snprintf(buf, sizeof buf, "%lldm %llds", m, s % 60);
   else if (s >= 10)
      std::snprintf(buf, sizeof buf, "%llds", s);
   else if (s >= 1)
      std::snprintf(buf, sizeof buf, "%.2fs", (double)ms / 1000.0);
   else
      std::snprintf(buf, sizeof buf, "%lldms", ms);
   return buf;
}

}  // namespace

extern "C" {

rayon_status rayon_write_png(const char* path, const uint8_t* pixels, int32_t width, int32_t height,
                             int32_t channels) {
   if (!path || !pixels || width <= 0 || height <= 0 || channels < 1 || channels > 4 ||
       (int64_t)width * height > (1ll << 31)) {
      ry::set_error("rayon_write_png: bad arguments");
      return RAYON_ERR_INVALID;
   }
   static const uint8_t color_type[5] = {0, 0, 4, 2, 6};
   std::vector<uint8_t> filtered = filter_rows(pixels, width, height, channels);
   uLongf zlen = compressBound((uLong)filtered.size());
   std::vector<uint8_t> z(zlen);
   if (compress2(z.data(), &zlen, filtered.data(), (uLong)filtered.size(), Z_DEFAULT_COMPRESSION) != Z_OK) {
      ry::set_error("rayon_write_png: deflate failed");
      return RAYON_ERR_OOM;
   }
   std::vector<uint8_t> file = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
   std::vector<uint8_t> ihdr;
   put_be32(ihdr, (uint32_t)width);
   put_be32(ihdr, (uint32_t)height);
   ihdr.insert(ihdr.end(), {8, color_type[channels], 0, 0, 0});  // depth, colour type, deflate, adaptive, no interlace
   put_chunk(file, "IHDR", ihdr.data(), ihdr.size());
   put_chunk(file, "IDAT", z.data(), zlen);
   put_chunk(file, "IEND", nullptr, 0);
   if (!ensure_parent_dir(path)) {
      ry::set_error("rayon_write_png: cannot create the directory of %s", path);
      return RAYON_ERR_IO;
   }
   FILE* f = std::fopen(path, "wb");
   if (!f) {
      ry::set_error("rayon_write_png: cannot open %s", path);
      return RAYON_ERR_IO;
   }
   const bool ok = std::fwrite(file.data(), 1, file.size(), f) == file.size();
   if (std::fclose(f) != 0 || !ok) {
      ry::set_error("rayon_write_png: write to %s failed", path);
      return RAYON_ERR_IO;
   }
   return RAYON_OK;
}

rayon_status rayon_job_write_png(rayon_ctx* ctx, rayon_job* job, const char* path, float gamma) {
   if (!ctx || !job || !path) {
      ry::set_error("rayon_job_write_png: NULL argument");
      return RAYON_ERR_INVALID;
   }
   const int w = job->params.width, h = job->params.height;
   std::vector<uint8_t> rgb((size_t)w * h * 3);
   const rayon_status st = rayon_job_resolve_u8(ctx, job, rgb.data(), 3, gamma);
   if (st != RAYON_OK) return st;
   return rayon_write_png(path, rgb.data(), w, h, 3);
}

rayon_status rayon_format_duration(int64_t duration_ns, char* buf, int32_t buf_len) {
   if (!buf || buf_len < 1 || duration_ns < 0) {
      ry::set_error("rayon_format_duration: bad arguments");
      return RAYON_ERR_INVALID;
   }
   const std::string s = format_duration(duration_ns);
   if ((int32_t)s.size() >= buf_len) {
      ry::set_error("rayon_format_duration: buffer of %d bytes too small", buf_len);
      return RAYON_ERR_INVALID;
   }
   std::memcpy(buf, s.c_str(), s.size() + 1);
   return RAYON_OK;
}

rayon_status rayon_write_render_stats(const char* image_path, const rayon_render_stats* st) {
   if (!image_path || !st || st->render_time_ns < 0) {
      ry::set_error("rayon_write_render_stats: bad arguments");
      return RAYON_ERR_INVALID;
   }
   int64_t size = st->image_size_bytes;
   if (size < 0) {
      std::error_code ec;
      const auto n = std::filesystem::file_size(image_path, ec);
      size = ec ? 0 : (int64_t)n;
   }
   const double seconds = (double)st->render_time_ns * 1e-9;
   const long long rps = seconds > 0.0 ? std::llround((double)st->rays_traced / seconds) : 0;
   const long long ms = st->render_time_ns / 1000000;
   const std::string pretty = format_duration(st->render_time_ns);
   const std::string name = std::filesystem::path(image_path).filename().string();

   const std::string txt_path = with_extension(image_path, ".txt");
   FILE* f = std::fopen(txt_path.c_str(), "w");
   if (!f) {
      ry::set_error("rayon_write_render_stats: cannot open %s", txt_path.c_str());
      return RAYON_ERR_IO;
   }
   std::fprintf(f,
                "image: %s\nsamples_per_pixel: %d\nresolution: %d x %d\nmax_depth: %d\nrays_traced: %lld\n"
                "image_size_bytes: %lld\nrays_per_second: %lld\nrender_time_ms: %lld\nrender_time_pretty: %s\n",
                name.c_str(), st->samples_per_pixel, st->width, st->height, st->max_depth,
                (long long)st->rays_traced, (long long)size, rps, ms, pretty.c_str());
   if (std::fclose(f) != 0) {
      ry::set_error("rayon_write_render_stats: write to %s failed", txt_path.c_str());
      return RAYON_ERR_IO;
   }

   const std::string json_path = with_extension(image_path, ".json");
   f = std::fopen(json_path.c_str(), "w");
   if (!f) {
      ry::set_error("rayon_write_render_stats: cannot open %s", json_path.c_str());
      return RAYON_ERR_IO;
   }
   std::fprintf(f,
                "{\n  \"image\": \"%s\",\n  \"samples_per_pixel\": %d,\n"
                "  \"resolution\": { \"width\": %d, \"height\": %d },\n  \"max_depth\": %d,\n"
                "  \"rays_traced\": %lld,\n  \"image_size_bytes\": %lld,\n  \"rays_per_second\": %lld,\n"
                "  \"render_time_ms\": %lld,\n  \"render_time_pretty\": \"%s\"\n}\n",
                name.c_str(), st->samples_per_pixel, st->width, st->height, st->max_depth,
                (long long)st->rays_traced, (long long)size, rps, ms, pretty.c_str());
   if (std::fclose(f) != 0) {
      ry::set_error("rayon_write_render_stats: write to %s failed", json_path.c_str());
      return RAYON_ERR_IO;
   }
   return RAYON_OK;
}

}  // extern "C"
