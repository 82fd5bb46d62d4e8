"""The scene front-end (pbrt_load_obj and the image codecs, include/pbrt_io.h) on the
CPU: images against Pillow's independent codecs, materials against closed forms of
RayTracingVisitor.cpp:253-298 and vsgXchange's Phong conversion (assimp.cpp:672-745),
OBJ files against hand-derived arrays and against the restatement in
oracle/obj_oracle.py."""
import ctypes as C
import math
import struct
import zlib

import numpy as np
import pytest
from PIL import Image

from oracle import obj_oracle
from vulkanpbrt_amd import capi

F32 = np.float32


# ---------------------------------------------------------------------------
# images
# ---------------------------------------------------------------------------
def _rgba(im: Image.Image) -> np.ndarray:
    return np.asarray(im.convert("RGBA"))


@pytest.fixture(scope="module")
def pattern():
    rng = np.random.default_rng(3)
    a = rng.integers(0, 256, (37, 53, 4), dtype=np.uint8)
    a[5:9, 7:20, 3] = 0
    return a


@pytest.mark.parametrize("mode", ["RGBA", "RGB", "L", "LA", "P", "1"])
def test_png_decode_matches_pillow(tmp_path, pbrt_lib, pattern, mode):
    src = Image.fromarray(pattern, "RGBA")
    im = src.convert(mode) if mode != "P" else src.convert("RGB").quantize(37)
    f = tmp_path / f"t_{mode}.png"
    if mode == "P":
        im.info["transparency"] = bytes([0, 128] + [255] * 35)
        im.save(f, transparency=im.info["transparency"])
    else:
        im.save(f)
    got = capi.load_image(f)
    want = _rgba(Image.open(f))
    assert got.shape == want.shape
    assert np.array_equal(got, want)


def _png_bytes(rows: np.ndarray, color_type: int, depth: int, filters) -> bytes:
    """A PNG with explicit per-row filter types (ISO/IEC 15948 section 9)."""
    h = rows.shape[0]
    raw_rows = rows.reshape(h, -1).astype(np.int32)
    bpp = max(1, {0: 1, 2: 3, 4: 2, 6: 4}[color_type] * depth // 8)
    out = bytearray()
    prev = np.zeros(raw_rows.shape[1], np.int32)
    for y in range(h):
        cur = raw_rows[y]
        ft = filters[y % len(filters)]
        f = np.zeros_like(cur)
        for i in range(len(cur)):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if ft == 0:
                pred = 0
            elif ft == 1:
                pred = a
            elif ft == 2:
                pred = b
            elif ft == 3:
                pred = (a + b) // 2
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            f[i] = (cur[i] - pred) & 0xFF
        out.append(ft)
        out += bytes(f.astype(np.uint8))
        prev = cur

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    w = rows.shape[1]
    ihdr = struct.pack(">IIBBBBB", w, h, depth, color_type, 0, 0, 0)
    return b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(bytes(out))) + \
        chunk(b"IEND", b"")


@pytest.mark.parametrize("filters", [[0], [1], [2], [3], [4], [0, 1, 2, 3, 4]])
def test_png_row_filters(tmp_path, pbrt_lib, pattern, filters):
    f = tmp_path / "f.png"
    f.write_bytes(_png_bytes(pattern, 6, 8, filters))
    assert np.array_equal(capi.load_image(f), pattern)
    # RGB, 3 bytes per pixel
    f.write_bytes(_png_bytes(pattern[..., :3].copy(), 2, 8, filters))
    got = capi.load_image(f)
    assert np.array_equal(got[..., :3], pattern[..., :3]) and np.all(got[..., 3] == 255)


def test_png_16_bit_keeps_the_high_byte(tmp_path, pbrt_lib, pattern):
    """16-bit samples become their high byte (stb_image's conversion)."""
    v = pattern[..., :3].astype(np.uint16) * 257 + np.arange(3, dtype=np.uint16)
    be = v.astype(">u2").view(np.uint8).reshape(v.shape[0], v.shape[1], 6)
    f = tmp_path / "d16.png"
    f.write_bytes(_png_bytes(be, 2, 16, [4, 1]))
    got = capi.load_image(f)
    assert np.array_equal(got[..., :3], (v >> 8).astype(np.uint8))


def test_png_write_round_trips_through_pillow(tmp_path, pbrt_lib, pattern):
    f = tmp_path / "w.png"
    capi.write_png(f, pattern)
    assert np.array_equal(_rgba(Image.open(f)), pattern)
    bgra = pattern[..., [2, 1, 0, 3]].copy()
    capi.write_png(f, bgra, bgra=True)
    assert np.array_equal(_rgba(Image.open(f)), pattern)


@pytest.mark.parametrize("mode", ["RGBA", "RGB", "L"])
@pytest.mark.parametrize("rle", [False, True])
@pytest.mark.parametrize("orientation", [1, -1])
def test_tga_decode_matches_pillow(tmp_path, pbrt_lib, pattern, mode, rle, orientation):
    im = Image.fromarray(pattern, "RGBA").convert(mode)
    f = tmp_path / "t.tga"
    im.save(f, compression="tga_rle" if rle else None, orientation=orientation)
    got = capi.load_image(f)
    assert np.array_equal(got, _rgba(Image.open(f)))
    assert np.array_equal(got, _rgba(im))


@pytest.mark.parametrize("mode", ["RGB", "L"])
def test_pnm_decode_matches_pillow(tmp_path, pbrt_lib, pattern, mode):
    im = Image.fromarray(pattern, "RGBA").convert(mode)
    f = tmp_path / ("t.ppm" if mode == "RGB" else "t.pgm")
    im.save(f)
    assert np.array_equal(capi.load_image(f), _rgba(im))


def test_pam_with_alpha_and_16_bit_pnm(tmp_path, pbrt_lib, pattern):
    h, w = pattern.shape[:2]
    f = tmp_path / "t.pam"
    f.write_bytes(f"P7\nWIDTH {w}\nHEIGHT {h}\nDEPTH 4\nMAXVAL 255\nTUPLTYPE RGB_ALPHA\nENDHDR\n".encode()
                  + pattern.tobytes())
    assert np.array_equal(capi.load_image(f), pattern)
    # maxval 1023: scaled to 8 bits with rounding
    v = (pattern[..., 0].astype(np.uint32) * 4) % 1024
    f = tmp_path / "t16.pgm"
    f.write_bytes(f"P5\n# comment\n{w} {h}\n1023\n".encode() + v.astype(">u2").tobytes())
    got = capi.load_image(f)
    assert np.array_equal(got[..., 0], ((v * 255 + 511) // 1023).astype(np.uint8))


def test_image_errors(tmp_path, pbrt_lib):
    f = tmp_path / "x.jpg"
    f.write_bytes(b"\xff\xd8\xff\xe0" + b"\0" * 32)
    with pytest.raises(capi.PbrtError) as e:
        capi.load_image(f)
    assert e.value.status == -2 and "JPEG" in str(e.value)
    f = tmp_path / "x.png"
    f.write_bytes(b"\x89PNG\r\n\x1a\n" + b"\0" * 8)
    with pytest.raises(capi.PbrtError):
        capi.load_image(f)
    with pytest.raises(capi.PbrtError):
        capi.load_image(tmp_path / "missing.png")


@pytest.mark.parametrize("header", [
    # w * h * depth * bytes and w * h * 4 both wrap to 0 in 64 bits
    b"P7\nWIDTH 2147483648\nHEIGHT 2147483648\nDEPTH 4\nMAXVAL 255\nENDHDR\n",
    # a width that does not fit 32 bits
    b"P7\nWIDTH 99999999999\nHEIGHT 1\nDEPTH 4\nMAXVAL 255\nENDHDR\n",
    b"P6\n99999999999 1\n255\n",
    # over the 2^28-pixel cap without wrapping
    b"P5\n65536 65536\n255\n",
])
def test_pnm_oversized_header_is_rejected(tmp_path, pbrt_lib, header):
    """A PNM header whose dimensions overflow (or exceed the 2^28-pixel cap PNG also
    has) is an invalid argument, never a short allocation written past its end."""
    f = tmp_path / "big.pam"
    f.write_bytes(header + b"\0" * 64)
    with pytest.raises(capi.PbrtError) as e:
        capi.load_image(f)
    assert e.value.status == -1


# ---------------------------------------------------------------------------
# materials
# ---------------------------------------------------------------------------
def _floats(m: capi.Material) -> np.ndarray:
    return np.frombuffer(bytes(m), np.float32)


def test_pack_phong_closed_form(pbrt_lib):
    ph = capi.PhongMaterial()
    ph.ambient[:] = [0.1, 0.2, 0.3, 1.0]
    ph.diffuse[:] = [0.5, 0.25, 0.125, 1.0]
    ph.specular[:] = [0.04, 0.04, 0.04, 1.0]
    ph.emissive[:] = [0.0, 0.0, 0.0, 1.0]
    ph.transmissive[:] = [1.0, 1.0, 1.0]
    ph.shininess = 14.0
    ph.alpha_mask, ph.alpha_mask_cutoff, ph.index_of_refraction, ph.category_id = 1.0, 0.5, 1.5, 9
    out, em = capi.Material(), C.c_uint32(7)
    capi.check(pbrt_lib.pbrt_pack_phong_material(C.byref(ph), C.byref(out), C.byref(em)))
    f = _floats(out)
    assert em.value == 0
    # ambient_roughness: ambient, sqrt(2 / (14 + 2)) = sqrt(1/8)
    assert np.array_equal(f[:4], F32([0.1, 0.2, 0.3, math.sqrt(0.125)]))
    assert np.array_equal(f[4:8], F32([0.5, 0.25, 0.125, 1.5]))       # diffuse, ior
    assert np.array_equal(f[8:12], F32([0.04, 0.04, 0.04, 1.0]))      # specular, alpha mask
    assert np.array_equal(f[12:16], F32([1, 1, 1, 0]))                # transmissive, illum 0
    assert np.array_equal(f[16:20], F32([0, 0, 0, 0.5]))              # emission, cutoff
    assert out.category_id == 9
    # any transmissive component != 1 turns on refraction (illum 7); emission sum != 0
    ph.transmissive[:] = [1.0, 0.5, 1.0]
    ph.emissive[:] = [0.0, 0.0, 2.0, 1.0]
    ph.shininess = 0.0
    capi.check(pbrt_lib.pbrt_pack_phong_material(C.byref(ph), C.byref(out), C.byref(em)))
    f = _floats(out)
    assert f[15] == 7.0 and em.value == 1 and f[3] == 1.0  # shininess 0: roughness 1


def test_pack_pbr_copies_base_color_into_ambient(pbrt_lib):
    pb = capi.PbrMaterial()
    pb.base_color[:] = [0.9, 0.8, 0.7, 0.6]
    pb.emissive[:] = [0, 0, 0, 1]
    pb.diffuse[:] = [0.3, 0.3, 0.3, 1]
    pb.specular[:] = [0.5, 0.5, 0.5, 0.25]
    pb.transmission[:] = [1, 1, 1]
    pb.metallic, pb.roughness, pb.alpha_mask, pb.alpha_mask_cutoff, pb.index_of_refraction = 1, 0.3, 1, 0.5, 1.33
    out, em = capi.Material(), C.c_uint32()
    capi.check(pbrt_lib.pbrt_pack_pbr_material(C.byref(pb), C.byref(out), C.byref(em)))
    f = _floats(out)
    assert np.array_equal(f[:4], F32([0.9, 0.8, 0.7, 0.3]))  # base colour, then roughness
    assert np.array_equal(f[4:8], F32([0.3, 0.3, 0.3, 1.33]))
    assert np.array_equal(f[8:12], F32([0.5, 0.5, 0.5, 1.0]))
    assert f[15] == 0.0 and em.value == 0


# ---------------------------------------------------------------------------
# OBJ
# ---------------------------------------------------------------------------
def _load(path, **kw):
    sc = capi.LoadedScene(path, **kw)
    return sc, sc.arrays()


def test_textured_quad_known_answer(tmp_path, pbrt_lib):
    """A unit quad with texture coordinates, no normals, no material library: one
    mesh of two triangles on the default material, v flipped, normals zero."""
    (tmp_path / "q.obj").write_text(
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\nf 1/1 2/2 3/3 4/4\n")
    sc, a = _load(tmp_path / "q.obj")
    assert len(a["meshes"]) == 1 and len(a["instances"]) == 1
    m = a["meshes"][0]
    assert np.array_equal(m["positions"].reshape(-1, 3),
                          F32([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 0, 0], [1, 1, 0], [0, 1, 0]]))
    assert np.array_equal(m["uvs"].reshape(-1, 2), F32([[0, 1], [1, 1], [1, 0], [0, 1], [1, 0], [0, 0]]))
    assert np.all(m["normals"] == 0)
    assert m["indices"].dtype == np.uint16 and np.array_equal(m["indices"], np.arange(6))
    f = np.frombuffer(m["material"].tobytes(), np.float32)
    assert np.array_equal(f[4:7], F32([0.6, 0.6, 0.6]))  # assimp's default Kd
    assert f[3] == 1.0 and np.all(f[8:11] == 0)          # Ns 0: roughness 1, no specular
    assert np.array_equal(a["instances"][0][0], np.eye(4, dtype=F32).ravel())
    st = sc.stats()
    assert st.triangle_count == 2 and st.vertex_count == 6 and st.mesh_count == 1
    assert list(st.bounds_min) == [0, 0, 0] and list(st.bounds_max) == [1, 1, 0]


def test_concave_quad_splits_at_the_reflex_corner(tmp_path, pbrt_lib):
    """Corner 2 of this dart is reflex: the split is (2,3,0), (2,0,1), the only one
    whose triangles stay inside the quad."""
    (tmp_path / "d.obj").write_text("v 0 0 0\nv 4 0 0\nv 1 1 0\nv 0 4 0\nf 1 2 3 4\n")
    _, a = _load(tmp_path / "d.obj")
    p = a["meshes"][0]["positions"].reshape(-1, 3)
    assert np.array_equal(p, F32([[1, 1, 0], [0, 4, 0], [0, 0, 0], [1, 1, 0], [0, 0, 0], [4, 0, 0]]))
    assert obj_oracle._quad_start([[0, 0, 0], [4, 0, 0], [1, 1, 0], [0, 4, 0]]) == 2
    assert obj_oracle._quad_start([[0, 0, 0], [4, 0, 0], [4, 4, 0], [0, 4, 0]]) == 0


def _write_scene(d, rng):
    """A scene exercising the OBJ/MTL grammar: two material libraries, texture options,
    negative indices, faces with and without vt/vn, a default-material group, an unknown
    material, a line continuation, points and lines, polygons of 3..6 corners."""
    tex = rng.integers(0, 256, (8, 8, 4), dtype=np.uint8)
    tex[0:3, 0:3, 3] = 0
    Image.fromarray(tex, "RGBA").save(d / "albedo.png")
    Image.fromarray(tex[..., :3].copy(), "RGB").save(d / "spec.tga")
    (d / "a.mtl").write_text(
        "# library one\nnewmtl red\nKd 0.75 0.125 0.0625\nKs 0.5 0.5 0.5\nNs 30\nKa 0.25\n"
        "map_Kd -s 1 1 1 -clamp on albedo.png\n"
        "newmtl glass\nKd 1 1 1\nTf 0.9 0.9 1.0\nNi 1.5\nNs 0.005\nKs 0.2 0.2 0.2\n")
    (d / "b.mtl").write_text(
        "newmtl lamp\nKe 4 4 3.5\nKd 0 0 0\nmap_Ks -imfchan r spec.tga\nillum 1\n"
        "newmtl shiny\nKs 0.9 0.8 0.7\nNs 200\nKd 0.1 0.1 0.1\nnorm albedo.png\n")
    lines = ["# generated", "mtllib a.mtl b.mtl", "o thing"]
    nv = 40
    for _ in range(nv):
        x, y, z = (rng.integers(-64, 64, 3) / 16.0)
        lines.append(f"v {x} {y} {z}")
    for _ in range(12):
        lines.append(f"vt {rng.integers(0, 33) / 32} {rng.integers(0, 33) / 32}")
    for _ in range(9):
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(F32)
        lines.append("vn " + " ".join(np.format_float_positional(x, unique=True) for x in n))

    def face(k, style):
        ids = rng.choice(nv, k, replace=False) + 1
        toks = []
        for i in ids:
            vi = i if rng.random() < 0.7 else i - nv - 1   # negative: relative to the end
            if style == 0:
                toks.append(f"{vi}")
            elif style == 1:
                toks.append(f"{vi}/{rng.integers(1, 13)}")
            elif style == 2:
                toks.append(f"{vi}//{rng.integers(1, 10)}")
            else:
                toks.append(f"{vi}/{rng.integers(1, 13)}/{-int(rng.integers(1, 10))}")
        return "f " + " ".join(toks)

    lines += [face(3, 0), face(3, 3), "g walls", "usemtl red", face(4, 3), face(3, 1), face(5, 2),
              "usemtl glass", face(3, 2), face(6, 3), "l 1 2 3", "p 4",
              "usemtl lamp", face(3, 0), face(4, 0), "usemtl nowhere", face(3, 1),
              "usemtl red", face(3, 3) + " \\", "", "usemtl shiny", face(4, 2), face(3, 3)]
    # a continued face: split one face over two lines
    cont = face(4, 1).split(" ")
    lines += [" ".join(cont[:3]) + " \\", " ".join(cont[3:])]
    (d / "s.obj").write_text("\n".join(lines) + "\n")
    return tex


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_obj_matches_the_restatement(tmp_path, pbrt_lib, seed):
    rng = np.random.default_rng(seed)
    tex = _write_scene(tmp_path, rng)
    sc, a = _load(tmp_path / "s.obj")
    ref = obj_oracle.load_obj(tmp_path / "s.obj")
    assert len(a["meshes"]) == len(ref)
    names = {}
    for got, want in zip(a["meshes"], ref):
        for k in ("positions", "normals", "uvs", "indices"):
            assert got[k].dtype == want[k].dtype and np.array_equal(got[k], want[k]), k
        assert np.array_equal(got["material"], want["material"])
        for slot, fname in enumerate(want["textures"]):
            ti = int(got["textures"][slot])
            assert (ti < 0) == (fname == "")
            if fname:
                names.setdefault(fname, ti)
                assert names[fname] == ti  # one texture per file
    st = sc.stats()
    assert st.emissive_mesh_count == sum(m["emissive"] for m in ref) == 1
    assert st.texture_count == len(names) == 2
    pix = a["textures"][names["albedo.png"]]
    assert np.array_equal(pix, tex)
    assert np.array_equal(a["textures"][names["spec.tga"]][..., :3], tex[..., :3])


def test_obj_instances_and_scene_creation_inputs(tmp_path, pbrt_lib):
    rng = np.random.default_rng(5)
    _write_scene(tmp_path, rng)
    t = np.stack([np.eye(4, dtype=F32).ravel(), np.eye(4, dtype=F32).ravel()])
    t[1, 12:15] = [5, 0, 0]  # column-major translation
    sc, a = _load(tmp_path / "s.obj", transforms=t, bvh_builder=capi.PBRT_BVH_LBVH_GPU)
    nm = len(a["meshes"])
    assert len(a["instances"]) == 2 * nm
    assert [m for _, m in a["instances"]] == list(range(nm)) * 2
    assert np.array_equal(a["instances"][nm][0], t[1])
    assert sc.desc.bvh_builder == capi.PBRT_BVH_LBVH_GPU
    assert sc.stats().instance_count == 2 * nm


def test_obj_large_mesh_uses_32_bit_indices(tmp_path, pbrt_lib):
    n = 22000  # 66000 corners
    lines = ["v 0 0 0", "v 1 0 0", "v 0 1 0"] + ["f 1 2 3"] * n
    (tmp_path / "big.obj").write_text("\n".join(lines))
    _, a = _load(tmp_path / "big.obj")
    m = a["meshes"][0]
    assert m["indices"].dtype == np.uint32 and len(m["indices"]) == 3 * n


@pytest.mark.parametrize("body,msg", [("v 0 0 0\nf 1 2 3\n", "out of range"),
                                      ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 x\n", "bad face index"),
                                      ("v 0 0 0\n", "no triangles"),
                                      ("mtllib missing.mtl\n", "cannot open")])
def test_obj_errors(tmp_path, pbrt_lib, body, msg):
    (tmp_path / "e.obj").write_text(body)
    with pytest.raises(capi.PbrtError) as e:
        capi.LoadedScene(tmp_path / "e.obj")
    assert msg in str(e.value) and e.value.status == -1
