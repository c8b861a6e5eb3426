"""Struct layouts of the C ABI: every ctypes structure of the Python binding has the size
and field offsets that a C compiler gives the matching struct of include/rayon_hip.h
(gcc on a generated translation unit; no GPU)."""
import ctypes as C
import subprocess
from pathlib import Path

import pytest

from rayon_amd import abi

ROOT = Path(__file__).resolve().parent.parent

PAIRS = {
    "Material": "rayon_material", "Geometry": "rayon_geometry", "Mesh": "rayon_mesh",
    "SdfPrim": "rayon_sdf_prim", "SceneDesc": "rayon_scene_desc", "SceneInfo": "rayon_scene_info",
    "Camera": "rayon_camera", "JobParams": "rayon_job_params", "JobStats": "rayon_job_stats",
    "RenderStats": "rayon_render_stats", "WaveStats": "rayon_wave_stats", "ServerConfig": "rayon_server_config",
    "ServerJobInfo": "rayon_server_job_info", "ServerStats": "rayon_server_stats",
    "TuneResult": "rayon_tune_result", "ProgressiveConfig": "rayon_progressive_config",
    "ProgressiveFrame": "rayon_progressive_frame", "GenParams": "rayon_gen_params",
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("layout")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "rayon_hip.h"', "int main(void) {"]
    for py, cname in PAIRS.items():
        st = getattr(abi, py)
        lines.append(f'printf("{py} size %zu\\n", sizeof({cname}));')
        for f in st._fields_:
            lines.append(f'printf("{py} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        py, field, value = line.split()
        got[(py, field)] = int(value)
    return got


@pytest.mark.parametrize("py", sorted(PAIRS))
def test_ctypes_struct_matches_c(c_layout, py):
    st = getattr(abi, py)
    assert C.sizeof(st) == c_layout[(py, "size")], f"{py}: sizeof differs"
    for f in st._fields_:
        assert getattr(st, f[0]).offset == c_layout[(py, f[0])], f"{py}.{f[0]}: offset differs"
