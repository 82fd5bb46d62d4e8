"""Write a scene description as Wavefront OBJ + MTL + PNG textures, the input format of
``pbrt_load_obj`` (include/pbrt_io.h), so that a generated workload can be re-read
through the front-end the reference uses for real assets (assimp, source/VulkanPBRT.cpp:199-210).

Instances are flattened into world space (positions in float64 by the object-to-world
matrix, normals by its inverse transpose).  Per mesh:
  * packed material -> MTL: Ka / Kd / Ks / Ke / Tf / Ni, Ns = 2 / roughness^2 - 2 (the
    inverse of RayTracingVisitor.cpp:292; a material whose specular colour is not zero
    keeps Ns >= 0.01, below which vsgXchange would zero it, assimp.cpp:694-698);
  * meshes with a zero first normal get the area-weighted vertex normals of
    RayTracingVisitor.cpp:52-118, and meshes without texture coordinates the default
    per-triangle assignment of RayTracingVisitor.cpp:122-159, written explicitly
    (the OBJ reader duplicates vertices per face corner, which would change both);
  * texture coordinates are written with v -> 1 - v, undone by the reader's FlipUVs.
A material whose illum is 7 with a transmittance of exactly 1 (not expressible in MTL)
and whose alpha-mask fields differ from vsg's defaults is reported in the return value.

    python -m vulkanpbrt_amd.scene_export --config 2 --out /tmp/atrium
"""
from __future__ import annotations

import argparse
import io
import json
import time
from pathlib import Path

import numpy as np


def _area_weighted_normals(pos: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """Running area-weighted mean of the unit face normals per vertex, in triangle order
    (RayTracingVisitor.cpp:63-85), not renormalised."""
    n = np.zeros_like(pos, dtype=np.float32)
    wsum = np.zeros(len(pos), np.float32)
    p = pos.astype(np.float32)
    tri = idx.reshape(-1, 3)
    for a, b, c in tri:
        fn = np.cross(p[b] - p[a], p[c] - p[a]).astype(np.float32)
        w = np.float32(np.sqrt(np.float32(fn @ fn)))
        fn = fn / w
        for i in (a, b, c):
            nw = np.float32(w + wsum[i])
            n[i] = n[i] * (wsum[i] / nw) + fn * (w / nw)
            wsum[i] = nw
    return n


def _default_uvs(nv: int, idx: np.ndarray) -> np.ndarray:
    uv = np.zeros((nv, 2), np.float32)
    for a, b, c in idx.reshape(-1, 3):
        uv[a], uv[b], uv[c] = (0, 0), (0, 1), (1, 0)
    return uv


def _fmt(a: np.ndarray) -> list[str]:
    """Rows of numbers with 9 significant digits (enough to round-trip float32)."""
    buf = io.StringIO()
    np.savetxt(buf, np.asarray(a, np.float64).reshape(len(a), -1), fmt="%.9g")
    return buf.getvalue().splitlines()


def export(arrays: dict, out_dir: Path, name: str = "scene", write_png=None) -> dict:
    """Write `name`.obj / .mtl and tex<i>.png under out_dir from a scene description's
    host arrays (capi.desc_arrays); returns counts and any material it could not express."""
    from vulkanpbrt_amd import capi
    write_png = write_png or capi.write_png
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    t0 = time.perf_counter()
    meshes, instances, textures = arrays["meshes"], arrays["instances"], arrays["textures"]
    for i, tex in enumerate(textures):
        write_png(out_dir / f"tex{i}.png", tex)
    mtl, inexact = [], []
    for mi, m in enumerate(meshes):
        f = np.frombuffer(m["material"].tobytes()[:80], np.float32)
        rough = float(f[3])
        ns = 2.0 / max(rough, 1e-4) ** 2 - 2.0
        if np.any(f[8:11] != 0) and ns < 0.01:
            ns = 0.01
            inexact.append((mi, f"roughness {rough:.6g} with a specular colour (Ns 0.01: roughness 0.9975)"))
        tf = f[12:15]
        if f[15] == 7.0 and np.all(tf == 1.0):
            inexact.append((mi, "illum 7 with transmittance 1"))
        if f[11] != 1.0 or f[19] != 0.5:
            inexact.append((mi, "alpha mask fields"))
        lines = [f"newmtl m{mi}", "Ka " + _fmt(f[None, 0:3])[0], "Kd " + _fmt(f[None, 4:7])[0],
                 "Ks " + _fmt(f[None, 8:11])[0], "Ke " + _fmt(f[None, 16:19])[0], "Tf " + _fmt(tf[None])[0],
                 "Ni " + _fmt(f[None, 7:8])[0], f"Ns {ns:.9g}"]
        for key, slot in (("map_Kd", 0), ("norm", 1), ("map_Ks", 2), ("map_Ke", 3)):
            if m["textures"][slot] >= 0:
                lines.append(f"{key} tex{int(m['textures'][slot])}.png")
        mtl.append("\n".join(lines))
    (out_dir / f"{name}.mtl").write_text("\n\n".join(mtl) + "\n")
    # per mesh: object-space normals and uvs as the packing would build them
    prepared = []
    for m in meshes:
        pos = m["positions"].reshape(-1, 3).astype(np.float64)
        idx = m["indices"].astype(np.int64)
        nor = m["normals"]
        nor = np.zeros_like(pos, np.float32) if nor is None else nor.reshape(-1, 3)
        if not np.any(nor[0]):
            nor = _area_weighted_normals(pos.astype(np.float32), idx)
        uv = m["uvs"].reshape(-1, 2) if m["uvs"] is not None else _default_uvs(len(pos), idx)
        prepared.append((pos, nor.astype(np.float64), uv, idx))
    obj = [f"# {len(instances)} instances of {len(meshes)} meshes, flattened", f"mtllib {name}.mtl"]
    base = 1
    tris = 0
    for ii, (xf, mesh) in enumerate(instances):
        pos, nor, uv, idx = prepared[mesh]
        M = np.array(xf, np.float64).reshape(4, 4).T   # column-major storage
        wp = pos @ M[:3, :3].T + M[:3, 3]
        nm = np.linalg.inv(M[:3, :3]).T
        wn = nor @ nm.T
        obj.append(f"o i{ii}")
        obj.append(f"usemtl m{mesh}")
        obj += ["v " + s for s in _fmt(wp)]
        obj += ["vt " + s for s in _fmt(np.stack([uv[:, 0], 1.0 - uv[:, 1]], 1))]
        obj += ["vn " + s for s in _fmt(wn)]
        t = idx.reshape(-1, 3) + base
        obj += [f"f {a}/{a}/{a} {b}/{b}/{b} {c}/{c}/{c}" for a, b, c in t]
        base += len(pos)
        tris += len(t)
    (out_dir / f"{name}.obj").write_text("\n".join(obj) + "\n")
    return dict(obj=str(out_dir / f"{name}.obj"), triangles=tris, meshes=len(meshes), instances=len(instances),
                textures=len(textures), inexact=inexact, seconds=round(time.perf_counter() - t0, 2))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--config", type=int, default=2, help="generated workload (1, 2 or 5)")
    ap.add_argument("--seed", type=int, default=None, help="scene seed (default: the config number)")
    ap.add_argument("--out", required=True, help="output folder")
    args = ap.parse_args()
    from vulkanpbrt_amd import capi
    gen = capi.GeneratedScene(args.config, args.seed if args.seed is not None else args.config)
    info = export(gen.arrays(), Path(args.out), f"config{args.config}")
    print(json.dumps(info))


if __name__ == "__main__":
    main()
