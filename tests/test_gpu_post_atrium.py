"""The post chain on atrium frames (instanced, textured, alpha-masked geometry, four
bounces): two consecutive 128x72 frames of the camera spline, traced by the oracle's
FP32 mirror, run through accumulate -> BMFR (FP64-MFMA Gram fit, 32-pixel blocks) ->
TAA on the GPU and through the oracle's accumulator, bmfrPre, FP64 least squares,
bmfrPost and TAA on the CPU, each stage compared (accumulator.comp, bmfrPre/Fit/Post.comp,
taa.comp)."""
import ctypes as C

import numpy as np
import pytest

from vulkanpbrt_amd import camera, capi
from tests import helpers

pytestmark = pytest.mark.gpu

W, H, BS = 128, 72, 32


def unhalf(a):
    return np.asarray(a, np.uint16).view(np.float16).astype(np.float32)


def dev(cuda, a):
    return cuda.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.fixture(scope="module")
def atrium_frames(oracle_lib, pbrt_lib):
    gen = capi.GeneratedScene(2, 2)
    osc = oracle_lib.OracleScene(gen.desc)
    out = []
    for f in (7, 8):
        view, proj = camera.pose_matrices(gen.camera(f), W, H)
        pc = camera.push_constants(view, proj, view, f)
        r = osc.trace(W, H, np.array(pc.view_inverse[:], np.float32), np.array(pc.proj_inverse[:], np.float32),
                      frame=f, max_depth=4, precision=32)
        r["view"], r["proj"] = view, proj
        out.append(r)
    return out


def consts(f, prev):
    o = np.linalg.inv(prev["view"])[:, 3].astype(np.float32)
    o[3] = 1
    return (camera.colmajor(np.linalg.inv(f["proj"])), camera.colmajor(np.linalg.inv(f["view"])),
            camera.colmajor(prev["view"]), o)


def gpu_accumulate(cuda, frame, f, prev, prev_depth, prev_spp, prev_acc):
    p = capi.AccumulateParams(W, H, frame, 0)
    c = consts(f, prev)
    p.inv_proj[:], p.inv_view[:], p.prev_view[:], p.prev_origin[:] = (x.tolist() for x in c)
    t = [dev(cuda, a) for a in (f["illumination"], f["depth"], prev_depth, prev_spp, prev_acc)]
    motion = cuda.zeros((H, W, 2), dtype=cuda.int16, device="cuda")
    spp = cuda.zeros((H, W), dtype=cuda.uint8, device="cuda")
    acc = cuda.zeros((H, W, 4), dtype=cuda.int16, device="cuda")
    im = capi.AccumulateImages(*(capi.ptr(x) for x in (*t, motion, spp, acc)))
    capi.check(capi.lib().pbrt_accumulate(C.byref(p), C.byref(im), helpers.stream(cuda)))
    cuda.cuda.synchronize()
    return motion.cpu().numpy().view(np.uint16), spp.cpu().numpy(), acc.cpu().numpy().view(np.uint16)


def test_atrium_post_chain_matches_oracle(cuda, oracle_lib, atrium_frames):
    f0, f1 = atrium_frames
    zeros = (np.zeros((H, W), np.float32), np.zeros((H, W), np.uint8), np.zeros((H, W, 4), np.uint16))
    # accumulator: frame 0 from empty history, frame 1 reprojected onto frame 0
    g0 = gpu_accumulate(cuda, 0, f0, f0, *zeros)
    r0 = oracle_lib.accumulate(W, H, 0, *consts(f0, f0), f0["illumination"], f0["depth"], *zeros)
    for g, r in zip(g0, r0):
        assert np.array_equal(g, r)
    g1 = gpu_accumulate(cuda, 1, f1, f0, f0["depth"], g0[1], g0[2])
    r1 = oracle_lib.accumulate(W, H, 1, *consts(f1, f0), f1["illumination"], f1["depth"], f0["depth"], r0[1], r0[2])
    motion_ok = np.all(np.abs(unhalf(g1[0]) - unhalf(r1[0])) <= 1e-3, axis=-1)
    acc_ok = np.all(np.isclose(unhalf(g1[2]), unhalf(r1[2]), rtol=2e-3, atol=1e-4), axis=-1)
    print(f"atrium accumulate frame 1: motion {motion_ok.mean():.5f} accumulated {acc_ok.mean():.5f}, "
          f"reprojected {(unhalf(r1[0])[..., 0] >= 0).mean():.3f}")
    assert motion_ok.mean() > 0.999 and acc_ok.mean() > 0.999 and (unhalf(r1[0])[..., 0] >= 0).mean() > 0.5

    # BMFR on the GPU's accumulator output, frame 1
    frame = 1
    motion, spp, noisy = g1
    denoised_in = np.zeros((2, H, W, 4), np.uint16)
    p = capi.BmfrParams(W, H, BS, frame, capi.BMFR_FIT_GRAM_F64)
    n = C.c_size_t()
    capi.check(capi.lib().pbrt_bmfr_workspace_size(C.byref(p), C.byref(n)))
    ws = cuda.zeros(n.value, dtype=cuda.uint8, device="cuda")
    t = {k: dev(cuda, v) for k, v in dict(depth=f1["depth"], normal=f1["normal"], albedo=f1["albedo"],
                                          motion=motion, spp=spp, noisy=noisy).items()}
    den = dev(cuda, denoised_in)
    bfinal = cuda.zeros((H, W, 4), dtype=cuda.uint8, device="cuda")
    im = capi.BmfrImages(*(capi.ptr(x) for x in (t["depth"], t["normal"], t["albedo"], t["motion"], t["spp"],
                                                  t["noisy"], den, bfinal, ws)), n.value)
    capi.check(capi.lib().pbrt_bmfr(C.byref(p), C.byref(im), helpers.stream(cuda)))
    cuda.cuda.synchronize()
    wp, hp = (W // BS + 2) * BS, (H // BS + 2) * BS
    fbytes = (13 * wp * hp * 2 + 255) // 256 * 256
    nw = 30 * (hp // BS) * (wp // BS)
    weights = ws.cpu().numpy()[fbytes:fbytes + 4 * nw].view(np.float32).reshape(hp // BS, wp // BS, 30)
    feats = oracle_lib.bmfr_pre(W, H, BS, frame, f1["depth"], f1["normal"], noisy)
    W64 = oracle_lib.bmfr_lsq64(W, H, BS, frame, feats)
    _, _, nbx, nby = oracle_lib.bmfr_geometry(W, H, BS)
    errs = []
    for by in range(nby):
        for bx in range(nbx):
            A = oracle_lib.bmfr_block_matrix(W, H, BS, frame, feats, bx, by).astype(np.float64)[:, :10]
            f64 = A @ W64[by, bx].reshape(10, 3)
            errs.append(np.abs(A @ weights[by, bx].reshape(10, 3).astype(np.float64) - f64).max()
                        / max(1e-2, np.abs(f64).max()))
    print(f"atrium BMFR fitted-value error vs FP64 LSQ: max {max(errs):.3e}")
    assert max(errs) < 5e-5
    # bmfrPost with the oracle's FP64 weights against the GPU's
    oden, ofinal = oracle_lib.bmfr_post(W, H, BS, frame, f1["depth"], f1["normal"], f1["albedo"], motion, spp, noisy,
                                        W64.astype(np.float32), denoised_in)
    gfinal = bfinal.cpu().numpy()
    gden = den.cpu().numpy().view(np.uint16)
    layer = (frame & 1) ^ 1
    d_ok = np.isclose(unhalf(gden[layer]), unhalf(oden[layer]), rtol=2e-3, atol=1e-3).all(-1)
    f_ok = (np.abs(gfinal.astype(int) - ofinal.astype(int)) <= 1).all(-1)
    print(f"atrium BMFR post: denoised {d_ok.mean():.5f} final {f_ok.mean():.5f}")
    assert d_ok.mean() > 0.999 and f_ok.mean() > 0.999

    # TAA on frame 1 with a history from frame 0's BMFR output (the oracle's)
    hist = ofinal.copy()
    rf, rh = oracle_lib.taa(W, H, frame, motion, ofinal, hist)
    tm, td, th = dev(cuda, motion), dev(cuda, ofinal), dev(cuda, hist)
    out = cuda.zeros((H, W, 4), dtype=cuda.uint8, device="cuda")
    ti = capi.TaaImages(*(capi.ptr(x) for x in (tm, td, th, out)))
    capi.check(capi.lib().pbrt_taa(W, H, frame, C.byref(ti), helpers.stream(cuda)))
    cuda.cuda.synchronize()
    g = out.cpu().numpy()
    ok = (np.abs(g.astype(int) - rf.astype(int)) <= 1).all(-1)
    print(f"atrium TAA frame {frame}: agreement {ok.mean():.5f}")
    assert ok.mean() > 0.999
    assert np.array_equal(th.cpu().numpy(), g)
