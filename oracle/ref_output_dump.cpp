// ref_output_dump.cpp — TEST INFRASTRUCTURE ONLY. Runs the reference's own output code
// (utils::FileUtils in utils.hpp, with the stb PNG writer it vendors), compiled from the
// reference sources where they lie by oracle/build_ref.sh into oracle/_ref/, so the tests
// can compare rayon_write_png, rayon_write_render_stats and rayon_format_duration with it:
//
//   ref_output_dump dur <ns>                       prints formatDuration(ns)
//   ref_output_dump png <raw> <w> <h> <out.png>    writes the raw RGB bytes with writeImage
//   ref_output_dump stats <image_path> <spp> <w> <h> <rays> <size_bytes> <ns>
//                                                  writes the .txt/.json of writeRenderStats
//                                                  and prints the camera's max_depth
#include <cstdlib>
#include <fstream>
#include <iterator>

#include "utils.hpp"

int main(int argc, char** argv) {
   if (argc >= 3 && std::string(argv[1]) == "dur") {
      std::cout << utils::FileUtils::formatDuration(std::chrono::nanoseconds(std::atoll(argv[2]))) << "\n";
      return 0;
   }
   if (argc == 6 && std::string(argv[1]) == "png") {
      std::ifstream in(argv[2], std::ios::binary);
      std::vector<unsigned char> img((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
      utils::FileUtils::writeImage(img, std::atoi(argv[3]), std::atoi(argv[4]), argv[5]);
      return 0;
   }
   if (argc == 9 && std::string(argv[1]) == "stats") {
      Camera cam(Vec3(0, 0, 0), std::atoi(argv[4]), std::atoi(argv[5]), 3, std::atoi(argv[3]));
      cam.n_rays = std::atoll(argv[6]);
      utils::FileUtils::writeRenderStats(cam, argv[2], (uintmax_t)std::atoll(argv[7]),
                                         std::chrono::nanoseconds(std::atoll(argv[8])));
      std::cout << cam.max_depth << "\n";
      return 0;
   }
   std::cerr << "usage: see the header of ref_output_dump.cpp\n";
   return 2;
}
