"""pbrt_bmfr_noalloc (BMFR with the library's per-device scratch array, no allocation)
gives the same bytes as pbrt_bmfr with a caller workspace, through the C ABI, on two
calls in a row (the second on a smaller frame) for each fit."""
import ctypes as C

import numpy as np
import pytest

from vulkanpbrt_amd import capi
from tests import helpers

pytestmark = pytest.mark.gpu


def _inputs(cuda, w, h, seed):
    """A seeded synthetic G-buffer and accumulator output of w x h pixels."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    depth = (4.0 + 0.01 * xx + 0.02 * yy + 0.1 * rng.random((h, w))).astype(np.float32)
    normal = np.stack([np.arccos(np.clip(rng.uniform(-1, 1, (h, w)), -1, 1)),
                       rng.uniform(-np.pi, np.pi, (h, w))], -1).astype(np.float32)
    albedo = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    motion = np.stack([(xx + 0.5) / w, (yy + 0.5) / h], -1).astype(np.float16).view(np.uint16)
    spp = rng.integers(1, 40, (h, w), dtype=np.uint8)
    noisy = (rng.random((h, w, 4)) * np.array([1.0, 0.8, 0.6, 1.0])).astype(np.float16).view(np.uint16)
    den = (rng.random((2, h, w, 4))).astype(np.float16).view(np.uint16)

    def d(a):
        return cuda.from_numpy(np.ascontiguousarray(a)).cuda()

    return dict(depth=d(depth), normal=d(normal), albedo=d(albedo), motion=d(motion.view(np.int16)),
                spp=d(spp), noisy=d(noisy.view(np.int16)), den=d(den.view(np.int16)))


def _run(cuda, entry, t, w, h, frame, fit, workspace):
    den = t["den"].clone()
    final = cuda.zeros((h, w, 4), dtype=cuda.uint8, device="cuda")
    p = capi.BmfrParams(w, h, 32, frame, fit)
    ws_ptr, ws_bytes = (capi.ptr(workspace), workspace.numel()) if workspace is not None else (None, 0)
    im = capi.BmfrImages(capi.ptr(t["depth"]), capi.ptr(t["normal"]), capi.ptr(t["albedo"]), capi.ptr(t["motion"]),
                         capi.ptr(t["spp"]), capi.ptr(t["noisy"]), capi.ptr(den), capi.ptr(final), ws_ptr, ws_bytes)
    capi.check(entry(C.byref(p), C.byref(im), helpers.stream(cuda)))
    cuda.cuda.synchronize()
    return den.cpu().numpy(), final.cpu().numpy()


@pytest.mark.parametrize("fit,sizes", [
    (capi.BMFR_FIT_GRAM_F64, [(1920, 1080), (256, 192)]),
    (capi.BMFR_FIT_GRAM_F16, [(640, 360), (200, 130)]),
    (capi.BMFR_FIT_HOUSEHOLDER, [(512, 384), (96, 80)]),
])
def test_noalloc_matches_workspace_entry_point(cuda, pbrt_lib, fit, sizes):
    for call, (w, h) in enumerate(sizes):
        print(f"bmfr noalloc fit {fit} call {call} {w}x{h}", flush=True)
        t = _inputs(cuda, w, h, seed=100 + call)
        frame = 5 + call
        n = C.c_size_t()
        capi.check(pbrt_lib.pbrt_bmfr_workspace_size(C.byref(capi.BmfrParams(w, h, 32, frame, fit)), C.byref(n)))
        ws = cuda.zeros(n.value, dtype=cuda.uint8, device="cuda")
        den_ref, fin_ref = _run(cuda, pbrt_lib.pbrt_bmfr, t, w, h, frame, fit, ws)
        den, fin = _run(cuda, pbrt_lib.pbrt_bmfr_noalloc, t, w, h, frame, fit, None)
        assert np.array_equal(fin, fin_ref), f"final image differs at {w}x{h}"
        assert np.array_equal(den, den_ref), f"denoised layers differ at {w}x{h}"
        assert fin_ref[..., :3].max() > 0


def test_noalloc_rejects_frames_beyond_its_scratch(cuda, pbrt_lib):
    t = _inputs(cuda, 64, 64, seed=1)
    # 7680x4320 Householder needs ~900 MB of features: rejected before any launch
    p = capi.BmfrParams(7680, 4320, 32, 0, capi.BMFR_FIT_HOUSEHOLDER)
    im = capi.BmfrImages(capi.ptr(t["depth"]), capi.ptr(t["normal"]), capi.ptr(t["albedo"]), capi.ptr(t["motion"]),
                         capi.ptr(t["spp"]), capi.ptr(t["noisy"]), capi.ptr(t["den"]), capi.ptr(t["den"]), None, 0)
    assert pbrt_lib.pbrt_bmfr_noalloc(C.byref(p), C.byref(im), helpers.stream(cuda)) == capi.PBRT_ERR_INVALID_ARG
    assert b"scratch" in pbrt_lib.pbrt_last_error()
