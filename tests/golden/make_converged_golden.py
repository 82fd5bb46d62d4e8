"""Writes tests/golden/converged_4096spp.npz: 4096-sample converged images of the FP64
oracle (and of its FP32 mirror) for two fixed cameras, the fixture of
tests/test_gpu_converged.py.

Sample s is traced as frame s (the reference's seed takes the frame number and
ignores sampleNumber, random.glsl:88-96), with the camera held still, so the mean over
s = 0..4095 is the 4096-spp estimate of ptRaygen.rgen:24-104 for that camera.  Stored
per case: the FP64 mean and per-pixel second moment of the demodulated illumination
(RGB), the FP32 mirror's mean over the same frames, and the primary hit's albedo and
normal in both precisions (the camera is still and unjittered, so every sample has the
same primary hit; pixels whose FP32 and FP64 primary hits differ look at an edge or a
corner where two surfaces meet at the same FP32 t).

Run from the repository root: python tests/golden/make_converged_golden.py
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle  # noqa: E402
from vulkanpbrt_amd import camera, capi  # noqa: E402

SAMPLES = 4096
CASES = {
    # name: (generated scene config and seed, width, height, max depth)
    "cornell_64x64": (1, 64, 64, 2),
    "atrium_64x36": (2, 64, 36, 4),
}


def push_constants(gen, w, h, frame):
    view, proj = camera.pose_matrices(gen.camera(0), w, h)
    return camera.push_constants(view, proj, view, frame, 0)


def unit(spherical):
    """G-buffer normal (theta = acos z, phi = atan(y, x), ptRaygen.rgen:51-53) as a unit vector."""
    t, p = spherical[..., 0].astype(np.float64), spherical[..., 1].astype(np.float64)
    return np.stack([np.cos(p) * np.sin(t), np.sin(p) * np.sin(t), np.cos(t)], -1)


def same_surface(r64, material, albedo, normal):
    """Pixels whose primary hit lies on the FP64 oracle's surface: same material, normal
    within 0.01 rad, G-buffer albedo within 4 of 255 (a texel-filtered albedo rounds
    differently when the hit's texture coordinates differ in the last bits)."""
    da = np.abs(r64["albedo"].astype(np.int32) - np.asarray(albedo).astype(np.int32)).max(-1)
    return (np.all(r64["material"] == material, -1) & (da <= 4) &
            (np.sum(unit(r64["normal"]) * unit(normal), -1) > np.cos(0.01)))


def relmse(a, b, eps=1e-2):
    return float(np.mean((a - b) ** 2 / (b * b + eps)))


def main():
    out, meta = {}, {}
    for name, (cfg, w, h, depth) in CASES.items():
        gen = capi.GeneratedScene(cfg, cfg)
        osc = oracle.OracleScene(gen.desc)
        s64 = np.zeros((h, w, 3))
        q64 = np.zeros((h, w, 3))
        s32 = np.zeros((h, w, 3))
        t0 = time.time()
        for f in range(SAMPLES):
            pc = push_constants(gen, w, h, f)
            iv = np.array(pc.view_inverse[:], np.float32)
            ip = np.array(pc.proj_inverse[:], np.float32)
            r64 = osc.trace(w, h, iv, ip, frame=f, max_depth=depth, precision=64, threads=8)
            r32 = osc.trace(w, h, iv, ip, frame=f, max_depth=depth, precision=32, threads=8)
            c64 = r64["illumination"][..., :3].astype(np.float64)
            s64 += c64
            q64 += c64 * c64
            s32 += r32["illumination"][..., :3]
            if f == 0:
                out[f"{name}_albedo"] = r64["albedo"]
                out[f"{name}_normal"] = r64["normal"]
                out[f"{name}_material"] = r64["material"]
                out[f"{name}_same_primary"] = same_surface(r64, r32["material"], r32["albedo"], r32["normal"])
        m64, m32 = s64 / SAMPLES, s32 / SAMPLES
        out[f"{name}_mean64"] = m64.astype(np.float32)
        out[f"{name}_m2_64"] = (q64 / SAMPLES).astype(np.float32)
        out[f"{name}_mean32"] = m32.astype(np.float32)
        var = np.maximum(q64 / SAMPLES - m64 * m64, 0.0)
        keep = out[f"{name}_same_primary"]
        meta[name] = dict(config=cfg, seed=cfg, width=w, height=h, max_depth=depth, samples=SAMPLES,
                          camera="generated camera path, pose 0, held still; sample s traced as frame s",
                          edge_pixels=int((~keep).sum()),
                          relmse_mirror32_vs_fp64=relmse(m32[keep], m64[keep]),
                          expected_relmse_two_independent=float(np.mean((2 * var / SAMPLES / (m64 * m64 + 1e-2))[keep])),
                          seconds=round(time.time() - t0, 1))
        print(name, meta[name], flush=True)
    out["meta"] = np.frombuffer(json.dumps(meta).encode(), np.uint8)
    np.savez_compressed(ROOT / "tests/golden/converged_4096spp.npz", **out)


if __name__ == "__main__":
    main()
