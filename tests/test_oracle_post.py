"""Pins of the oracle's screen-space passes (CPU only).

* BMFR noise: Jenkins hash + addRandom against tests/golden/bmfr_noise_kat.json.
* BMFR FP64 least squares: against numpy.linalg.lstsq (LAPACK) on the same block.
* BMFR FP32 Householder (the reference algorithm): fitted values close to FP64.
* accumulator / TAA / format converter: closed-form cases.
"""
import json
from pathlib import Path

import numpy as np
import pytest

NOISE = json.loads((Path(__file__).parent / "golden" / "bmfr_noise_kat.json").read_text())
rng = np.random.default_rng(1234)


def half(a):
    return np.asarray(a, np.float32).astype(np.float16).view(np.uint16)


def unhalf(a):
    return np.asarray(a, np.uint16).view(np.float16).astype(np.float32)


@pytest.mark.parametrize("pb", [64, 256, 1024])
def test_bmfr_noise_kat(oracle_lib, pb):
    bs = int(round(pb ** 0.5))
    w = h = 2 * bs
    wp, hp, _, _ = oracle_lib.bmfr_geometry(w, h, bs)
    feats = np.full((13, hp, wp), 0x3C00, np.uint16)  # every feature == 1.0
    for case in (c for c in NOISE if c["pb"] == pb):
        A = oracle_lib.bmfr_block_matrix(w, h, bs, case["frame"], feats, 0, 0)
        got = A[case["i"], case["col"]].view(np.uint32)
        assert got == case["noised_one_bits"], case
    A = oracle_lib.bmfr_block_matrix(w, h, bs, 3, feats, 0, 0)
    assert np.all(A[:, 10:] == 1.0)           # colour columns carry no noise
    assert np.all(np.abs(A[:, :10].astype(np.float64) - 1.0) <= 1e-4 + 2.4e-7)  # noise + 2 ulp


def synthetic_gbuffer(w, h, seed=0):
    r = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    depth = (2.0 + 0.5 * np.sin(xx / 7.0) + 0.3 * np.cos(yy / 5.0)).astype(np.float32)
    theta = (0.6 + 0.3 * np.sin(xx / 11.0 + yy / 13.0)).astype(np.float32)
    phi = (1.0 + 0.8 * np.cos(yy / 9.0)).astype(np.float32)
    normal = np.stack([theta, phi], -1)
    base = np.stack([0.5 + 0.3 * np.sin(xx / 6), 0.4 + 0.2 * np.cos(yy / 8), 0.3 + 0.1 * np.sin((xx + yy) / 10)], -1)
    noisy = np.concatenate([base + 0.2 * r.standard_normal(base.shape), np.ones((h, w, 1))], -1)
    return depth, normal.astype(np.float32), half(np.clip(noisy, 0, None))


def test_bmfr_lsq64_matches_lapack(oracle_lib):
    w = h = 96
    bs = 32
    depth, normal, noisy = synthetic_gbuffer(w, h)
    feats = oracle_lib.bmfr_pre(w, h, bs, 5, depth, normal, noisy)
    W64 = oracle_lib.bmfr_lsq64(w, h, bs, 5, feats)
    for (bx, by) in [(0, 0), (1, 2), (3, 1), (4, 4)]:
        A = oracle_lib.bmfr_block_matrix(w, h, bs, 5, feats, bx, by).astype(np.float64)
        sol = np.linalg.lstsq(A[:, :10], A[:, 10:], rcond=None)[0]  # 10 x 3
        ref = sol.reshape(-1)  # feature-major, rgb inner
        fit_ref = A[:, :10] @ sol
        fit = A[:, :10] @ W64[by, bx].reshape(10, 3)
        assert np.max(np.abs(fit - fit_ref)) <= 1e-9 * max(1.0, np.abs(fit_ref).max())
        # the weights themselves agree to the conditioning of the block
        assert np.allclose(W64[by, bx], ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())


def test_bmfr_householder_fp32_close_to_fp64(oracle_lib):
    w = h = 96
    bs = 32
    depth, normal, noisy = synthetic_gbuffer(w, h, 3)
    feats = oracle_lib.bmfr_pre(w, h, bs, 2, depth, normal, noisy)
    W32 = oracle_lib.bmfr_fit(w, h, bs, 2, feats)
    W64 = oracle_lib.bmfr_lsq64(w, h, bs, 2, feats)
    _, _, nbx, nby = oracle_lib.bmfr_geometry(w, h, bs)
    worst = 0.0
    for by in range(nby):
        for bx in range(nbx):
            A = oracle_lib.bmfr_block_matrix(w, h, bs, 2, feats, bx, by).astype(np.float64)
            f32 = A[:, :10] @ W32[by, bx].reshape(10, 3).astype(np.float64)
            f64 = A[:, :10] @ W64[by, bx].reshape(10, 3)
            worst = max(worst, np.max(np.abs(f32 - f64)) / max(1e-3, np.abs(f64).max()))
    assert worst < 5e-3


def _static_camera_consts(w, h):
    from vulkanpbrt_amd import camera
    view = camera.look_at([0, -3, 1], [0, 0, 1], [0, 0, 1])
    proj = camera.perspective(60, w / h, 0.1, 1000)
    inv_view = camera.colmajor(np.linalg.inv(view))
    inv_proj = camera.colmajor(np.linalg.inv(proj))
    prev_view = camera.colmajor(view)
    prev_origin = np.array([0, -3, 1, 1], np.float32)
    return inv_proj, inv_view, prev_view, prev_origin


def test_accumulator_frame0_passthrough(oracle_lib):
    w, h = 40, 24
    consts = _static_camera_consts(w, h)
    illum = rng.uniform(0, 3, (h, w, 4)).astype(np.float32)
    depth = np.full((h, w), 2.5, np.float32)
    z = np.zeros((h, w), np.float32)
    motion, spp, acc = oracle_lib.accumulate(w, h, 0, *consts, illum, depth, z, np.zeros((h, w), np.uint8),
                                             np.zeros((h, w, 4), np.uint16))
    assert np.all(unhalf(motion) == -1.0)
    assert np.all(spp == 1)  # round(255 / 256)
    assert np.array_equal(acc[..., :3], half(illum[..., :3]))
    assert np.all(unhalf(acc[..., 3]) == 1.0)


def test_accumulator_static_camera_blend_and_reject(oracle_lib):
    w, h = 40, 24
    consts = _static_camera_consts(w, h)
    illum = rng.uniform(0, 3, (h, w, 4)).astype(np.float32)
    # constant depth (a sphere around the eye): the reprojected depth matches exactly
    # wherever the bilinear tap lands; a depth gradient would be rejected near the right
    # and bottom borders because of the size / (size - 0.5) uv scale (accumulator.comp:70)
    yy, xx = np.mgrid[0:h, 0:w]
    depth = np.full((h, w), 3.0, np.float32)
    prev_acc = np.zeros((h, w, 4), np.float32)
    prev_acc[..., :3] = [0.25, 0.5, 0.75]
    motion, spp, acc = oracle_lib.accumulate(w, h, 1, *consts, illum, depth, depth, np.full((h, w), 1, np.uint8),
                                             half(prev_acc))
    m = unhalf(motion)
    inner = (slice(2, h - 2), slice(2, w - 2))
    assert np.all(m[inner] >= 0)  # the last column may round to uv > 1 and be rejected
    expect_u = (xx + 0.5) / (w - 0.5)  # static camera: ((clip + 1) / 2) * W / (W - 0.5)
    assert np.allclose(m[inner][..., 0], expect_u[inner], atol=2e-3)
    alpha = max(1.0 / ((1 / 256 + 1 / 255) * 256.0), 0.1)
    expect = prev_acc[..., :3] * (1 - alpha) + illum[..., :3] * alpha
    assert np.allclose(unhalf(acc[..., :3])[inner], expect[inner], rtol=2e-3, atol=2e-3)
    # depth dissimilarity > 1% rejects the history
    motion2, spp2, acc2 = oracle_lib.accumulate(w, h, 1, *consts, illum, depth, depth * 1.05,
                                                np.full((h, w), 1, np.uint8), half(prev_acc))
    assert np.all(unhalf(motion2) == -1.0) and np.all(spp2 == 1)
    assert np.array_equal(acc2[..., :3], half(illum[..., :3]))


def test_taa_cases(oracle_lib):
    w, h = 16, 12
    den = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    hist = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    motion = half(np.full((h, w, 2), 0.5, np.float32))
    # frame 0: pass-through, and the history becomes a raw copy of the final image
    f0, h0 = oracle_lib.taa(w, h, 0, motion, den, hist)
    assert np.array_equal(f0[..., :3], den[..., :3]) and np.array_equal(h0, f0)
    # invalid motion: pass-through
    f1, _ = oracle_lib.taa(w, h, 3, half(np.full((h, w, 2), -1.0, np.float32)), den, hist)
    assert np.array_equal(f1[..., :3], den[..., :3])
    # constant image with history equal to it (history is read as RGBA bytes of a
    # BGRA image: a grey value is swap-invariant): inside the box => 0.4c + 0.6c = c
    grey = np.full((h, w, 4), 120, np.uint8)
    f2, _ = oracle_lib.taa(w, h, 3, motion, grey, grey)
    assert np.array_equal(f2[..., :3], grey[..., :3])
    # a history far outside the neighbourhood box is rejected
    f3, _ = oracle_lib.taa(w, h, 3, motion, grey, np.full((h, w, 4), 250, np.uint8))
    assert np.array_equal(f3[..., :3], grey[..., :3])


def test_convert_closed_form(oracle_lib):
    w, h = 9, 7
    src = rng.uniform(-0.2, 1.2, (h, w, 4)).astype(np.float32)
    out = oracle_lib.convert(w, h, src)
    ref = np.rint(np.clip(src, 0, 1) * 255).astype(np.uint8)
    assert np.array_equal(out, ref[..., [2, 1, 0, 3]])
