"""Build recipe for librayon_hip.so (HIP for gfx950) and the oracle's liboracle.so.

    python -m rayon_amd.build            # both libraries, in-tree
    python -m rayon_amd.build --product  # only librayon_hip.so

Objects go to rayon_amd/_build/; the libraries land next to the sources that load them
(rayon_amd/librayon_hip.so, oracle/liboracle.so) so they travel with the tree.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "rayon_amd"
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "librayon_hip.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 ops are disabled for every device compile; the compensated sums in the
# accumulation rely on plain IEEE adds, so no fast-math either. Multiply-adds are fused
# only within one source expression (-ffp-contract=on), never across statements after
# inlining: the same device function then rounds the same way in every kernel that
# inlines it, so a job's bits do not depend on which kernel variant (stack size, tree
# layout, counting) a batched pass picks for the company it keeps.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    warn = (proc.stdout + proc.stderr).strip()
    if warn and "packed-fp32-ops" not in warn:
        sys.stderr.write(warn + "\n")


def build_product(jobs: int = 8) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    cmds = []
    objs = []
    for s in srcs:
        o = OBJ / (s.name + ".o")
        objs.append(o)
        lang = ["-x", "hip"] if s.suffix == ".hip" else []
        cmds.append([HIPCC, *HIP_FLAGS, *lang, "-c", str(s), "-o", str(o)])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread", "-lz"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-pthread", "-Wall",
          "-o", str(tmp), str(ORACLE_DIR / "oracle.cpp")])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


CPU_BASE_LIB = ORACLE_DIR / "libcpubase.so"


def build_cpu_baseline() -> Path:
    """The benchmark's CPU baseline: the oracle's restatement of the path built in FP32 with
    the GPU path's tree rule (binned SAH, 16 bins, leaves of at most 4) and the reference CPU
    renderer's schedule (one contiguous chunk of rows per thread)."""
    tmp = CPU_BASE_LIB.with_suffix(".so.tmp")
    _run(["g++", "-std=c++17", "-O3", "-march=x86-64-v2", "-fPIC", "-shared", "-pthread", "-Wall",
          "-DORACLE_REAL=float", "-DORACLE_BVH_BINNED", "-DORACLE_STATIC_ROWS",
          "-o", str(tmp), str(ORACLE_DIR / "oracle.cpp")])
    os.replace(tmp, CPU_BASE_LIB)
    return CPU_BASE_LIB


def build_reference() -> None:
    """The reference's own scene-file loader, compiled from its sources where they lie (only
    where they exist, i.e. not on a machine without the reference), into oracle/_ref/."""
    _run(["bash", str(ORACLE_DIR / "build_ref.sh")])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--product", action="store_true")
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    both = not (a.product or a.oracle)
    if a.product or both:
        print(build_product(a.j))
    if a.oracle or both:
        print(build_oracle())
        print(build_cpu_baseline())
        build_reference()
    return 0


if __name__ == "__main__":
    sys.exit(main())
