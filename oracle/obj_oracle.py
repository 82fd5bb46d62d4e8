"""Test infrastructure: a CPU restatement of the reference's scene front-end for
Wavefront OBJ/MTL, the checker of the product loader (``pbrt_load_obj``).  Only
``tests/`` import it.

The reference reads scenes through assimp (5.0/5.1, not vendored:
external/vsgXchange/src/assimp/build_vars.cmake:1-4) with the import flags
Triangulate | FlipUVs | OptimizeMeshes | SortByPType | GenUVCoords
(external/vsgXchange/src/assimp/assimp.cpp:316), copies each aiMesh into vertex,
normal, texcoord and index arrays (assimp.cpp:405-452), converts materials on the
Phong path (assimp.cpp:672-745) and packs them for the tracer
(source/scene/RayTracingVisitor.cpp:279-298).  assimp's OBJ reader is restated from
its published behaviour: one vertex per face corner, quads split at a reflex
corner, FlipUVs v -> 1 - v, MTL defaults Kd 0.6 / Tf 1 / Ni 1, faces without a
usemtl (or with an unknown name) on the default material.  Polygons with more than
four corners are split as a fan from corner 0 (assimp clips ears: the same surface
for a planar convex polygon, other triangles).

Pure Python over small files: plain loops, float32 where the product stores float32.
"""
from __future__ import annotations

import math
import struct
from pathlib import Path

import numpy as np

F32 = np.float32


def _mtl_defaults():
    return dict(Ka=[0.0, 0.0, 0.0], Kd=[0.6, 0.6, 0.6], Ks=[0.0, 0.0, 0.0], Ke=[0.0, 0.0, 0.0],
                Tf=[1.0, 1.0, 1.0], Ns=0.0, Ni=1.0, map_Kd="", map_Ks="", map_Ke="", map_Kn="")


def _texture_name(tokens):
    """Last statement word after the -option arguments (the MTL texture syntax)."""
    i = 0
    while i < len(tokens) and tokens[i].startswith("-"):
        opt = tokens[i]
        i += 1
        if opt in ("-imfchan", "-type"):
            i += 1
            continue
        while i < len(tokens):
            try:
                float(tokens[i])
                i += 1
            except ValueError:
                break
        if i < len(tokens) and tokens[i] in ("on", "off"):
            i += 1
    return " ".join(tokens[i:])


def _logical_lines(text):
    out, cur = [], ""
    for raw in text.split("\n"):
        s = raw.rstrip("\r \t")
        if s.endswith("\\"):
            cur += s[:-1] + " "
            continue
        out.append(cur + raw)
        cur = ""
    if cur:
        out.append(cur)
    return out


def parse_mtl(text):
    mats, cur = {}, None
    for line in _logical_lines(text):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        k = t[0]
        if k == "newmtl":
            cur = _mtl_defaults()
            mats[" ".join(t[1:])] = cur
            continue
        if cur is None:
            continue
        if k in ("Ka", "Kd", "Ks", "Ke", "Tf"):
            if t[1] in ("spectral", "xyz"):
                continue
            v = [float(x) for x in t[1:4]]
            if len(v) == 1:
                v = v * 3
            elif len(v) == 2:
                v = [v[0], v[1], v[1]]
            cur[k] = v
        elif k in ("Ns", "Ni"):
            cur[k] = float(t[1])
        elif k == "map_Kd":
            cur["map_Kd"] = _texture_name(t[1:])
        elif k == "map_Ks":
            cur["map_Ks"] = _texture_name(t[1:])
        elif k in ("map_Ke", "map_emissive"):
            cur["map_Ke"] = _texture_name(t[1:])
        elif k in ("map_Kn", "norm"):
            cur["map_Kn"] = _texture_name(t[1:])
    return mats


def packed_phong(ambient, diffuse, specular, emissive, transmissive, shininess, ior, alpha_mask=1.0,
                 cutoff=0.5, category=0):
    """WaveFrontMaterialPacked of a vsg::PhongMaterial (RayTracingVisitor.cpp:279-298) as
    96 bytes, and whether the mesh is emissive."""
    f = np.zeros(20, F32)
    f[0:3], f[3] = ambient[:3], np.sqrt(F32(2.0) / (F32(shininess) + F32(2.0)))  # float32 throughout
    f[4:7], f[7] = diffuse[:3], ior
    f[8:11], f[11] = specular[:3], alpha_mask
    tr = np.array(transmissive, F32)
    f[12:15], f[15] = tr, (7.0 if np.any(tr != 1.0) else 0.0)
    f[16:19], f[19] = emissive[:3], cutoff
    raw = f.tobytes() + struct.pack("<I", category) + b"\0" * 12
    em = float(np.float32(emissive[0]) + np.float32(emissive[1]) + np.float32(emissive[2])) != 0.0
    return np.frombuffer(raw, np.uint8).copy(), em


def phong_from_mtl(m):
    """vsgXchange's Phong conversion of an OBJ material (assimp.cpp:672-702): Ns below
    0.01 zeroes the shininess and the specular colour."""
    ns = F32(m["Ns"])
    spec = list(m["Ks"])
    if ns < F32(0.01):
        ns, spec = F32(0.0), [0.0, 0.0, 0.0]
    return dict(ambient=m["Ka"], diffuse=m["Kd"], specular=spec, emissive=m["Ke"], transmissive=m["Tf"],
                shininess=float(ns), ior=m["Ni"])


def _quad_start(p):
    """Corner whose two edge angles to the opposite diagonal add up to more than pi."""
    for i in range(4):
        o = p[i]
        def unit(q):
            d = [F32(q[k]) - F32(o[k]) for k in range(3)]
            n = F32(math.sqrt(float(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))
            return [F32(x / n) for x in d]
        left, diag, right = unit(p[(i + 3) % 4]), unit(p[(i + 2) % 4]), unit(p[(i + 1) % 4])
        dl = float(left[0] * diag[0] + left[1] * diag[1] + left[2] * diag[2])
        dr = float(right[0] * diag[0] + right[1] * diag[1] + right[2] * diag[2])
        ang = math.acos(max(-1.0, min(1.0, dl))) + math.acos(max(-1.0, min(1.0, dr)))
        if ang > 3.14159265:
            return i
    return 0


def load_obj(path):
    """Meshes of an OBJ file as the product's loader must produce them: a list of
    dicts (positions, normals, uvs, indices, index_stride, material bytes, emissive,
    texture file names) in order of first use of their material."""
    path = Path(path)
    v, vt, vn = [], [], []
    mtl = {}
    meshes = {}   # material key -> corners
    order = []
    cur = None    # None: assimp's default material

    def idx(i, n):
        i = int(i)
        r = i - 1 if i > 0 else n + i
        if i == 0 or not 0 <= r < n:
            raise ValueError("index out of range")
        return r

    for line in _logical_lines(path.read_text()):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        k = t[0]
        if k == "v":
            v.append([float(x) for x in t[1:4]])
        elif k == "vt":
            vt.append([float(t[1]), float(t[2]) if len(t) > 2 else 0.0])
        elif k == "vn":
            vn.append([float(x) for x in t[1:4]])
        elif k == "mtllib":
            for name in t[1:]:
                mtl.update(parse_mtl((path.parent / name).read_text()))
        elif k == "usemtl":
            name = " ".join(t[1:])
            cur = name if name in mtl else None
        elif k == "f":
            corners = []
            for tok in t[1:]:
                parts = tok.split("/")
                c = [idx(parts[0], len(v)),
                     idx(parts[1], len(vt)) if len(parts) > 1 and parts[1] else -1,
                     idx(parts[2], len(vn)) if len(parts) > 2 and parts[2] else -1]
                corners.append(c)
            if len(corners) < 3:
                continue
            if len(corners) == 4:
                s = _quad_start([v[c[0]] for c in corners])
                tri = [corners[(s + j) % 4] for j in (0, 1, 2, 0, 2, 3)]
            else:
                tri = []
                for j in range(1, len(corners) - 1):
                    tri += [corners[0], corners[j], corners[j + 1]]
            if cur not in meshes:
                meshes[cur] = []
                order.append(cur)
            meshes[cur] += tri
    out = []
    for key in order:
        cs = meshes[key]
        m = mtl[key] if key is not None else _mtl_defaults()
        n = len(cs)
        pos = np.array([v[c[0]] for c in cs], F32).reshape(-1)
        nor = np.array([vn[c[2]] if c[2] >= 0 else [0.0, 0.0, 0.0] for c in cs], F32).reshape(-1)
        uv = np.array([[vt[c[1]][0], 1.0 - F32(vt[c[1]][1])] if c[1] >= 0 else [0.0, 0.0] for c in cs],
                      F32).reshape(-1)
        stride = 2 if n < 65535 else 4
        ph = phong_from_mtl(m)
        mat, em = packed_phong(**ph)
        out.append(dict(positions=pos, normals=nor, uvs=uv,
                        indices=np.arange(n, dtype=np.uint16 if stride == 2 else np.uint32),
                        index_stride=stride, material=mat, emissive=em,
                        textures=(m["map_Kd"], m["map_Kn"], m["map_Ks"], m["map_Ke"])))
    return out
