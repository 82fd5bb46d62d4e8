"""Build the gfx950 shared library ``libpbrt.so`` in-tree with hipcc.

Every ``csrc/*.hip`` and ``csrc/*.cpp`` file is compiled for gfx950 (object
files under ``_build/``) and linked into ``vulkanpbrt_amd/libpbrt.so``.
Run ``python -m vulkanpbrt_amd.build`` or call :func:`build`.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libpbrt.so"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unknown-pragmas",
    "-Wno-unused-command-line-argument",
    # packed FP32 instructions stay off (v_pk_* f32 forms are not used)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(PKG.parent / "include"),
    "-I", str(CSRC),
]


MODULES = PKG / "renderModules"
APP = PKG / "app"
APP_BIN = PKG / "pbrt_frame"


def _sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")) + list(MODULES.glob("*.cpp")))


# The tracer's shading arithmetic is evaluated as written, one rounding per
# operation like the CPU oracle's FP32 mirror (built with -ffp-contract=off):
# hipcc would otherwise fuse a*b+c into v_fmac_f32, moving per-pixel radiance
# by an ulp at each fused step.  The traversal's slab tests spell their fused
# multiply-adds out with fmaf, so they keep them.
PER_FILE = {
    "trace.hip": ["-ffp-contract=off"],
    "wavefront.hip": ["-ffp-contract=off"],
}


def _compile(src: Path) -> Path:
    obj = BUILD / (src.parent.name + "_" + src.name + ".o")
    cmd = [HIPCC, "-x", "hip", *COMMON, *PER_FILE.get(src.name, []), "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if res.stderr.strip():
        msgs = [ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if msgs:
            sys.stderr.write("\n".join(msgs) + "\n")
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile all sources and link ``libpbrt.so``; returns the library path."""
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(8, max(1, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    # zlib: the PNG reader and writer of the scene front-end (scene_io.cpp)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lz"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    # example application of the render-module API (VulkanPBRT.cpp main loop, headless)
    app_obj = _compile(APP / "pbrt_frame.cpp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-o", str(APP_BIN), str(app_obj), "-L", str(PKG), "-lpbrt",
           "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link of pbrt_frame failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources and {APP_BIN.name}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
