#!/bin/bash
# Builds the reference's scene-file loader and its output code from their own sources, where
# they lie under /root/reference, into oracle/_ref/ (git-ignored). Test infrastructure only:
# the tests compare rayon_load_yaml and the PNG/statistics writers against them. Exits 0 without building when the sources are absent.
set -e
REF=${RAYON_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
SRC=$REF/src/rayon
if [ ! -f "$SRC/scenes/yaml_scene_loader.cc" ]; then
  echo "reference sources not found under $REF: oracle/_ref not built"
  exit 0
fi
mkdir -p "$HERE/_ref"
g++ -std=c++17 -O1 -w -I"$SRC/scenes" -I"$SRC/data_structures" -I"$SRC/cpu_renderers/utils" -I"$SRC" \
    "$SRC/scenes/yaml_scene_loader.cc" "$HERE/ref_yaml_dump.cpp" -o "$HERE/_ref/ref_yaml_dump"
echo "$HERE/_ref/ref_yaml_dump"
# the output code of the command-line renderer (utils.hpp, with the stb writer it vendors)
g++ -std=c++17 -O1 -w -I"$SRC" -I"$SRC/camera" -I"$SRC/data_structures" -I"$SRC/cpu_renderers/utils" -I"$REF/src/external" \
    "$HERE/ref_output_dump.cpp" -o "$HERE/_ref/ref_output_dump"
echo "$HERE/_ref/ref_output_dump"
# the reference's offline GPU renderer (renderer_cuda_host.hpp, unchanged) linked against
# librayon_hip.so through the scene shim a RayON maintainer would add
# (rayon_amd/integration/scene_builder_hip.cc); needs the library built first
LIB=$HERE/../rayon_amd/librayon_hip.so
if [ -f "$LIB" ]; then
  g++ -std=c++17 -O1 -w -I"$SRC" -I"$SRC/scenes" -I"$SRC/gpu_renderers" -I"$SRC/data_structures" -I"$SRC/camera" \
      -I"$SRC/cpu_renderers" -I"$SRC/cpu_renderers/utils" -I"$SRC/cpu_renderers/cpu_shapes" -I"$HERE/../include" \
      "$HERE/ref_cuda_host_drive.cpp" "$HERE/../rayon_amd/integration/scene_builder_hip.cc" \
      "$SRC/scenes/yaml_scene_loader.cc" \
      -L"$HERE/../rayon_amd" -lrayon_hip -Wl,-rpath,'$ORIGIN/../../rayon_amd' -pthread \
      -o "$HERE/_ref/ref_cuda_host_drive"
  echo "$HERE/_ref/ref_cuda_host_drive"
fi
