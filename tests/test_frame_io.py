"""Offline frame files (include/pbrt_io.h, vulkanpbrt_amd/frame_io.py) on the CPU:
MatrixIO's JSON schema and plain lists (source/io/RenderIO.cpp:593-716), PFM images,
and the GBufferIO conversions (RenderIO.cpp:160-382) against numpy restatements in
float32 and closed forms (round trips, known pixels)."""
import json
import struct

import numpy as np
import pytest

from vulkanpbrt_amd import camera, capi, frame_io

F32 = np.float32


def _mat(a):
    return np.array(a[:], np.float32)


def test_matrices_json_known_answer(tmp_path, pbrt_lib):
    """A file in the reference's schema: 16 numbers per matrix in vsg::mat4 storage
    (column) order; proj / invProj only for "ModelView+Projection"."""
    view = list(range(1, 17))
    doc = {"amtOfFrames": 2, "matrices": [
        {"type": "ModelView+Projection", "storageType": "ColumnMajor", "view": view,
         "invView": [v * 2 for v in view], "proj": [v * 3 for v in view], "invProj": [v * 4 for v in view]},
        {"type": "ModelViewProjection", "storageType": "ColumnMajor", "view": [0.5] * 16, "invView": [-1] * 16}]}
    f = tmp_path / "m.json"
    f.write_text(json.dumps(doc))
    m = frame_io.read_matrices(f)
    assert len(m) == 2
    assert np.array_equal(_mat(m[0].view), np.arange(1, 17)) and np.array_equal(_mat(m[0].inv_view), 2 * np.arange(1, 17))
    assert m[0].has_proj == 1 and np.array_equal(_mat(m[0].inv_proj), 4 * np.arange(1, 17))
    assert m[1].has_proj == 0 and np.all(_mat(m[1].proj) == 0) and np.all(_mat(m[1].view) == 0.5)


def test_matrices_write_schema_and_round_trip(tmp_path, pbrt_lib):
    rng = np.random.default_rng(1)
    mats = [frame_io.camera_matrices(*camera.pose_matrices(_pose(rng), 320, 200))]
    mats.append(frame_io.camera_matrices(rng.normal(size=(4, 4)) + 4 * np.eye(4)))
    f = tmp_path / "out.json"
    frame_io.write_matrices(f, mats)
    doc = json.loads(f.read_text())
    assert doc["amtOfFrames"] == 2 and len(doc["matrices"]) == 2
    m0 = doc["matrices"][0]
    assert m0["type"] == "ModelView+Projection" and m0["storageType"] == "ColumnMajor"
    assert set(m0) == {"type", "storageType", "view", "invView", "proj", "invProj"}
    assert doc["matrices"][1]["type"] == "ModelViewProjection" and "proj" not in doc["matrices"][1]
    assert np.array_equal(np.array(m0["view"], np.float32), _mat(mats[0].view))
    back = frame_io.read_matrices(f)
    for a, b in zip(mats, back):
        assert bytes(a) == bytes(b)


def _pose(rng):
    p = capi.CameraPose()
    p.eye[:] = rng.uniform(-3, 3, 3).tolist()
    p.center[:] = [0.0, 0.0, 0.5]
    p.up[:] = [0.0, 0.0, 1.0]
    p.fovy_degrees, p.near_distance, p.far_distance = 45.0, 0.1, 100.0
    return p


def test_matrices_plain_list_inverts(tmp_path, pbrt_lib):
    rng = np.random.default_rng(2)
    a = rng.normal(size=(4, 4)) + 3 * np.eye(4)
    b = rng.normal(size=(4, 4)) + 3 * np.eye(4)
    text = "{" + ", ".join(f"{x:.6f}" for x in a.ravel()) + "}\n" + " ".join(f"{x:.6f}" for x in b.ravel()) + "\n"
    f = tmp_path / "m.txt"
    f.write_text(text)
    m = frame_io.read_matrices(f)
    assert len(m) == 2
    for mm, src in zip(m, (a, b)):
        v = _mat(mm.view).reshape(4, 4)
        assert np.array_equal(v.ravel(), np.array([float(f"{x:.6f}") for x in src.ravel()], np.float32))
        # column-major storage: the product of the stored matrices is the identity
        prod = _mat(mm.inv_view).reshape(4, 4).T.astype(np.float64) @ v.T.astype(np.float64)
        assert np.allclose(prod, np.eye(4), atol=1e-5)


def test_pfm_layout_and_round_trip(tmp_path, pbrt_lib):
    img = np.arange(2 * 3 * 3, dtype=np.float32).reshape(2, 3, 3) / 7
    f = tmp_path / "a.pfm"
    frame_io.write_pfm(f, img)
    raw = f.read_bytes()
    head = b"PF\n3 2\n-1.0\n"
    assert raw[:len(head)] == head
    body = np.frombuffer(raw[len(head):], "<f4").reshape(2, 3, 3)
    assert np.array_equal(body[::-1], img)          # bottom row first
    assert np.array_equal(frame_io.read_pfm(f), img)
    # one channel, from a 4-channel array
    rgba = np.random.default_rng(0).random((5, 4, 4), dtype=np.float32)
    frame_io.write_pfm(f, rgba[..., :1])
    assert np.array_equal(frame_io.read_pfm(f), rgba[..., 0])
    frame_io.write_pfm(f, rgba)                      # first three channels
    assert np.array_equal(frame_io.read_pfm(f), rgba[..., :3])
    # a big-endian file (positive scale)
    g = tmp_path / "be.pfm"
    g.write_bytes(b"Pf\n2 1\n1.0\n" + struct.pack(">2f", 1.5, -2.25))
    assert np.array_equal(frame_io.read_pfm(g), np.array([[1.5, -2.25]], np.float32))
    with pytest.raises(capi.PbrtError):
        frame_io.read_pfm(tmp_path / "missing.pfm")


def test_normal_conversions(pbrt_lib):
    rng = np.random.default_rng(3)
    n = rng.normal(size=(500, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    xyzw = np.concatenate([n, np.ones((500, 1))], 1).astype(np.float32)
    tp = frame_io.normals_to_spherical(xyzw)
    ref = np.stack([np.arccos(xyzw[:, 2]), np.arctan2(xyzw[:, 1], xyzw[:, 0])], 1)
    assert np.allclose(tp, ref, atol=6e-7, rtol=0)  # 2 ulp at pi (libm vs numpy float32)
    back = frame_io.spherical_to_normals(tp)
    assert np.all(back[:, 3] == 1.0)
    assert np.allclose(back[:, :3], xyzw[:, :3], atol=1e-6)
    # poles and the axes in closed form
    tp = np.array([[0, 0], [np.pi, 0], [np.pi / 2, np.pi / 2]], np.float32)
    assert np.allclose(frame_io.spherical_to_normals(tp)[:, :3], [[0, 0, 1], [0, 0, -1], [0, 1, 0]], atol=1e-7)


def test_albedo_conversions(pbrt_lib):
    codes = np.arange(256, dtype=np.uint8)
    rgba8 = np.stack([codes, codes[::-1], codes, np.full(256, 255, np.uint8)], 1)
    f = frame_io.unorm_to_float(rgba8)
    assert np.array_equal(f, rgba8.astype(np.float32) / F32(255))
    back = frame_io.albedo_from_float(f)
    # truncation of 255 * (c / 255) in float32: exact except where the product rounds below c
    expect = np.clip(np.floor(f * F32(255)), 0, 255).astype(np.uint8)
    assert np.array_equal(back, expect)
    lost = np.nonzero(back[:, 0] != codes)[0]
    print("codes that come back one lower:", lost.tolist())
    assert np.all(back[:, 0] >= codes.astype(int) - 1)
    assert np.array_equal(frame_io.albedo_from_float(np.array([[-0.5, 1.5, 0.999, 0.0]], np.float32)),
                          np.array([[0, 255, 254, 0]], np.uint8))
    # 16-bit floats: normal numbers convert exactly (the reference's bit conversion)
    h = np.array([[0.0, 0.25, 0.5, 1.0], [0.75, 0.125, 0.9, 0.33]], np.float16)
    got = frame_io.albedo_from_half(h.view(np.uint16))
    v = h.astype(np.float32)
    v[0, 0] = F32(2.0 ** -15)   # zero bits map to 2^-15 in that conversion
    assert np.array_equal(got, np.floor(v * F32(255)).astype(np.uint8))


def _camera():
    p = capi.CameraPose()
    p.eye[:], p.center[:], p.up[:] = [1.0, -4.0, 1.5], [0.0, 0.0, 0.8], [0.0, 0.0, 1.0]
    p.fovy_degrees, p.near_distance, p.far_distance = 40.0, 0.05, 50.0
    return p


def test_depth_position_round_trip(pbrt_lib):
    w, h = 40, 24
    view, proj = camera.pose_matrices(_camera(), w, h)
    m = frame_io.camera_matrices(view, proj)
    depth = np.random.default_rng(4).uniform(0.5, 20.0, (h, w)).astype(np.float32)
    pos = frame_io.depth_to_position(depth, m)
    assert np.all(pos[..., 3] == 1.0)
    # the position lies on the tracer's camera ray of the pixel centre (createRay)
    from tests import helpers
    pc = camera.push_constants(view, proj, view, 0)
    rays = helpers.camera_rays(pc, w, h).reshape(h, w, 8).astype(np.float64)
    ref = rays[..., :3] + rays[..., 4:7] * depth[..., None]
    assert np.allclose(pos[..., :3], ref, rtol=0, atol=2e-5 * depth.max())
    back = frame_io.position_to_depth(pos, m)
    assert np.allclose(back, depth, rtol=2e-6)
    with pytest.raises(capi.PbrtError):
        frame_io.depth_to_position(depth, frame_io.camera_matrices(view))  # no projection matrix


def test_g_buffer_export_import(tmp_path, pbrt_lib):
    rng = np.random.default_rng(5)
    w, h = 16, 8
    view, proj = camera.pose_matrices(_camera(), w, h)
    mats = [frame_io.camera_matrices(view, proj)] * 2
    frames = []
    for _ in range(2):
        n = rng.normal(size=(h, w, 3))
        n /= np.linalg.norm(n, axis=-1, keepdims=True)
        g = dict(depth=rng.uniform(1, 9, (h, w)).astype(np.float32),
                 normal=np.stack([np.arccos(n[..., 2]), np.arctan2(n[..., 1], n[..., 0])], -1).astype(np.float32),
                 albedo=rng.integers(0, 256, (h, w, 4), dtype=np.uint8),
                 material=rng.integers(0, 256, (h, w, 4), dtype=np.uint8))
        g["albedo"][..., 3] = 255
        frames.append(g)
    fm = frame_io.formats_in(tmp_path, depth="d_%d.pfm", position="p_%d.pfm", normal="n_%d.pfm",
                             albedo="a_%d.pfm", material="m_%d.pfm")
    frame_io.export_g_buffer(fm, frames, mats)
    assert (tmp_path / "m_1.pfm").exists()
    for kind in ("depth", "position"):
        sel = {k: v for k, v in fm.items() if k != ("position" if kind == "depth" else "depth")}
        back = frame_io.import_g_buffer(sel, 2, mats if kind == "position" else None)
        for a, b in zip(frames, back):
            if kind == "depth":
                assert np.array_equal(a["depth"], b["depth"])
            else:
                assert np.allclose(a["depth"], b["depth"], rtol=3e-6)
            assert np.allclose(a["normal"][..., 0], b["normal"][..., 0], atol=1e-3)
            assert np.all(np.abs(a["albedo"].astype(int) - b["albedo"]) <= 1)
    ill = [rng.random((h, w, 4), dtype=np.float32) for _ in range(2)]
    frame_io.export_illumination(str(tmp_path / "i_%d.pfm"), ill)
    back = frame_io.import_illumination(str(tmp_path / "i_%d.pfm"), 2)
    for a, b in zip(ill, back):
        assert np.array_equal(a[..., :3], b[..., :3]) and np.all(b[..., 3] == 1.0)
