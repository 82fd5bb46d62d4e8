"""The render-module drop-in (vulkanpbrt_amd/app/pbrt_frame.cpp over renderModules/):
every denoiser wiring of add_denoiser_to_commands runs, and replaying the recorded
frame as a hipGraph gives the same bytes as submitting it eagerly."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

APP = Path(__file__).resolve().parents[1] / "vulkanpbrt_amd" / "pbrt_frame"


def run_app(tmp_path, tag, *args):
    out = tmp_path / f"{tag}.bin"
    cmd = [str(APP), "--config", "1", "--frames", "4", "--raw", str(out), *args]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    return np.fromfile(out, np.uint8).reshape(info["height"], info["width"], 4), info


@pytest.mark.parametrize("denoiser,block", [("bmfr", "32"), ("bmfr", "8"), ("bfr", "8"), ("bfr", "16"),
                                            ("bfr", "32"), ("bfr", "blend"), ("bmfr", "blend"), ("learned", "32"),
                                            ("none", "32")])
def test_app_eager_equals_graph(tmp_path, denoiser, block):
    assert APP.exists(), "build the library first (python -m vulkanpbrt_amd.build)"
    args = ["--denoiser", denoiser, "--block", block, "--taa"] if denoiser != "none" else ["--denoiser", "none"]
    eager, info = run_app(tmp_path, "eager", *args)
    graph, _ = run_app(tmp_path, "graph", *args, "--graph")
    print(denoiser, block, "ms per frame", info["ms"])
    assert np.array_equal(eager, graph)
    assert np.all(eager[..., 3] == 255)
    # a lit scene: the final image is not black
    assert eager[..., :3].mean() > 5


def _run(cmd, timeout=120):
    """Runs a child process.  This module never loads the library into the test
    process itself: a process that has loaded the HIP runtime but not yet used it, and
    then runs GPU children, has found no device at its first HIP call on these boxes,
    which would fail every later test of the session."""
    print("running", " ".join(map(str, cmd)), flush=True)
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr
    return r


def test_app_exports_then_denoises_the_files(tmp_path):
    """--export* (VulkanPBRT.cpp:100-120, 457-480) writes each frame's G-buffer,
    illumination and matrices; a second run with --depths/--normals/... takes them as
    external buffers (no tracing) and denoises them to the live run's final image,
    within one code where the normals round through cartesian xyz."""
    d = tmp_path
    live = d / "live.bin"
    _run([APP, "--config", "1", "--frames", "4", "--taa", "--raw", live,
          "--exportDepth", d / "depth_%d.pfm", "--exportPosition", d / "position_%d.pfm",
          "--exportNormal", d / "normal_%d.pfm", "--exportAlbedo", d / "albedo_%d.pfm",
          "--exportMaterial", d / "material_%d.pfm", "--exportIllumination", d / "illum_%d.pfm",
          "--exportMatrices", d / "matrices.json"])
    assert (d / "material_3.pfm").exists() and (d / "matrices.json").exists()
    a = np.fromfile(live, np.uint8).reshape(256, 256, 4)
    for src in ("--depths", "--positions"):
        off = d / f"off{src}.bin"
        pattern = d / ("depth_%d.pfm" if src == "--depths" else "position_%d.pfm")
        r = _run([APP, "-f", "4", "--taa", "--raw", off, src, pattern, "--normals", d / "normal_%d.pfm",
                  "--albedos", d / "albedo_%d.pfm", "--illuminations", d / "illum_%d.pfm",
                  "--matrices", d / "matrices.json"])
        b = np.fromfile(off, np.uint8).reshape(256, 256, 4)
        diff = np.abs(a.astype(int) - b.astype(int)).max(-1)
        print(src, "identical", (diff == 0).mean(), "max", diff.max())
        assert (diff <= 1).mean() >= 0.995 and (diff == 0).mean() >= 0.9
        assert "for 0 triangles" in r.stderr  # nothing traced


def test_app_png_pfm_and_statistics(tmp_path):
    raw, png, pfm = tmp_path / "f.bin", tmp_path / "f.png", tmp_path / "f.pfm"
    r = _run([APP, "--config", "1", "--frames", "3", "--taa", "--raw", raw, "--png", png, "--pfm", pfm])
    img = np.fromfile(raw, np.uint8).reshape(256, 256, 4)
    # read back without loading the library into this process (see _run)
    from PIL import Image
    assert np.array_equal(np.asarray(Image.open(png).convert("RGBA")), img[..., [2, 1, 0, 3]])
    data = pfm.read_bytes()
    head = b"PF\n256 256\n-1.0\n"
    assert data.startswith(head)
    f = np.frombuffer(data[len(head):], "<f4").reshape(256, 256, 3)[::-1]  # bottom row first
    assert np.array_equal(f, img[..., [2, 1, 0]].astype(np.float32) / np.float32(255))
    # the reference GUI's lines (Gui.hpp:39-44) for 5,144 triangles, 256x256, 2 * depth rays/pixel
    assert "ms/frame" in r.stderr and "for 5144 triangles" in r.stderr
    assert "Render size is 256 by 256 with 4 rays/pixel resulting in" in r.stderr
    assert "Samples per pixel: 1" in r.stderr


def test_app_traces_an_obj_file(tmp_path):
    r = _run([sys.executable, "-m", "vulkanpbrt_amd.scene_export", "--config", "1", "--out", tmp_path])
    info = json.loads(r.stdout.strip().splitlines()[-1])
    a, b = tmp_path / "gen.bin", tmp_path / "obj.bin"
    _run([APP, "--config", "1", "--frames", "2", "--raw", a])
    r = _run([APP, "--config", "1", "--frames", "2", "--raw", b, "-i", info["obj"]])
    assert f"for {info['triangles']} triangles" in r.stderr
    ia, ib = (np.fromfile(p, np.uint8).reshape(256, 256, 4)[..., :3].astype(float) for p in (a, b))
    print("final image means", ia.mean(), ib.mean())
    assert abs(ia.mean() - ib.mean()) <= 0.02 * ia.mean()
