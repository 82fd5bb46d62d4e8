"""The device builders (LBVH, PBRT_BVH_LBVH_GPU; PLOC, PBRT_BVH_PLOC_GPU) and the host
spatial-split tree against the host SAH tree: all are acceleration structures over the
same triangles, so every closest-hit query must return the same triangle and the same t
(up to equal-t ties, where the traversal order picks), every shadow query the same
answer, and a traced frame the same pixels.  The device radix sort the GPU builders use
is checked against numpy's stable sort."""
import ctypes as C

import numpy as np
import pytest

from vulkanpbrt_amd import capi
from tests import helpers

pytestmark = pytest.mark.gpu


def intersect(torch, scene, rays):
    n = rays.shape[0]
    r = torch.from_numpy(rays).cuda()
    hits = torch.zeros((n, 4), dtype=torch.int32, device="cuda")
    bary = torch.zeros((n, 2), dtype=torch.float32, device="cuda")
    occ = torch.zeros(n, dtype=torch.uint8, device="cuda")
    capi.check(capi.lib().pbrt_intersect(scene.handle, capi.ptr(r), n, capi.ptr(hits), capi.ptr(bary),
                                         helpers.stream(torch)))
    capi.check(capi.lib().pbrt_occluded(scene.handle, capi.ptr(r), n, capi.ptr(occ), helpers.stream(torch)))
    torch.cuda.synchronize()
    return hits.cpu().numpy().view(np.uint32), occ.cpu().numpy()


def scene_rays(gen, w, h, n_random, seed):
    pc, _, _ = helpers.frame_push_constants(gen, w, h, 0)
    cam = helpers.camera_rays(pc, w, h)
    rng = np.random.default_rng(seed)
    o = cam[0, :3]
    d = rng.normal(size=(n_random, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rnd = np.zeros((n_random, 8), np.float32)
    rnd[:, :3] = o + rng.normal(scale=0.5, size=(n_random, 3))
    rnd[:, 3] = 1e-3
    rnd[:, 4:7] = d
    rnd[:, 7] = 1e4
    return np.concatenate([cam, rnd]), pc


BUILDERS = {"lbvh": capi.PBRT_BVH_LBVH_GPU, "sbvh": capi.PBRT_BVH_SBVH_HOST, "ploc": capi.PBRT_BVH_PLOC_GPU}


@pytest.mark.parametrize("n,bits", [(1, 30), (4095, 30), (4096, 32), (4097, 8), (100003, 13), (1 << 20, 30)])
def test_radix_sort_matches_stable_argsort(cuda, pbrt_lib, n, bits):
    """pbrt_radix_sort_pairs: keys with many repeats (so stability shows), values = input
    index; equal to numpy's stable argsort on the masked key bits, for tile-aligned and
    ragged sizes and for an odd number of 8-bit passes (13 bits)."""
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    keys[rng.random(n) < 0.5] &= np.uint32(0xFFFF00FF)  # repeats in the middle digits
    masked = keys & np.uint32((1 << bits) - 1) if bits < 32 else keys
    kd = cuda.from_numpy(masked.view(np.int32).copy()).cuda()
    vd = cuda.arange(n, dtype=cuda.int32, device="cuda")
    k2, v2 = cuda.empty_like(kd), cuda.empty_like(vd)
    nb = C.c_size_t()
    capi.check(pbrt_lib.pbrt_radix_sort_scratch_size(n, C.byref(nb)))
    scratch = cuda.empty(nb.value, dtype=cuda.uint8, device="cuda")
    capi.check(pbrt_lib.pbrt_radix_sort_pairs(capi.ptr(kd), capi.ptr(vd), capi.ptr(k2), capi.ptr(v2), n, bits,
                                              capi.ptr(scratch), nb.value, helpers.stream(cuda)))
    cuda.cuda.synchronize()
    order = np.argsort(masked, kind="stable")
    assert np.array_equal(vd.cpu().numpy().view(np.uint32), order.astype(np.uint32))
    assert np.array_equal(kd.cpu().numpy().view(np.uint32), masked[order])


@pytest.mark.parametrize("builder", ["lbvh", "sbvh", "ploc"])
@pytest.mark.parametrize("config", [1, 2])
def test_lbvh_queries_match_sah(cuda, pbrt_lib, config, builder):
    """The same for the spatial-split tree (PBRT_BVH_SBVH_HOST), whose leaves may hold a
    triangle more than once: hits report the same (instance, primitive)."""
    gen = capi.GeneratedScene(config, config)
    gen.desc.bvh_builder = capi.PBRT_BVH_SAH_HOST
    sah = capi.Scene(gen.desc)
    kind = BUILDERS[builder]
    gen.desc.bvh_builder = kind
    lb = capi.Scene(gen.desc)
    s, l = sah.info(), lb.info()
    print(f"config {config}: {l.triangle_count} triangles; SAH {s.bvh_node_count} nodes depth {s.bvh_depth} "
          f"{s.bvh_build_ms:.1f} ms (host) | {builder} {l.bvh_node_count} nodes depth {l.bvh_depth} "
          f"{l.bvh_build_ms:.2f} ms")
    assert s.bvh_builder == capi.PBRT_BVH_SAH_HOST and l.bvh_builder == kind
    assert l.triangle_count == s.triangle_count and 1 < l.bvh_depth <= 40  # the traversal stack
    if builder in ("lbvh", "ploc"):
        assert l.bvh_node_count == l.triangle_count - 1
    rays, pc = scene_rays(gen, 320, 180, 40000, config)
    hs, os_ = intersect(cuda, sah, rays)
    hl, ol = intersect(cuda, lb, rays)
    ts, tl = hs[:, 0].view(np.float32), hl[:, 0].view(np.float32)
    same_tri = (hs[:, 1] == hl[:, 1]) & (hs[:, 2] == hl[:, 2])
    hit = hs[:, 1] != 0xFFFFFFFF
    tie = hit & (hl[:, 1] != 0xFFFFFFFF) & (np.abs(ts - tl) <= 1e-6 * np.abs(ts))
    print(f"rays {len(rays)}: same triangle {same_tri.mean():.6f}, differing at equal t {int((~same_tri & tie).sum())}")
    assert np.all(same_tri | tie)
    assert np.array_equal(ts[same_tri], tl[same_tri])
    assert np.array_equal(os_, ol)
    # a traced frame (megakernel) is the same image up to tie pixels; trees of depth <= 32
    # run the megakernel built with a 32-entry stack for five waves per SIMD, deeper ones
    # (the atrium's PLOC and LBVH trees) the 40-entry build, so config 2 compares the two
    print(f"megakernel stacks: SAH tree {32 if s.bvh_depth <= 32 else 40} entries, "
          f"{builder} tree {32 if l.bvh_depth <= 32 else 40} entries")
    imgs = []
    for sc in (sah, lb):
        img = helpers.gbuffer(cuda, 320, 180)
        helpers.trace(cuda, sc, pc, img, 320, 180)
        cuda.cuda.synchronize()
        imgs.append(img["illumination"].cpu().numpy())
    eq = np.all(imgs[0] == imgs[1], axis=-1).mean()
    # pixels whose path meets an equal-t tie (coplanar instanced boxes and walls in the
    # atrium) shade the other surface; every other pixel is bit-identical
    print(f"traced frame: identical pixels {eq:.5f}")
    assert eq > 0.995
    sah.close()
    lb.close()


def test_ploc_city_queries_match_sah(cuda, pbrt_lib):
    """PLOC on the 4.9M-triangle city: the tree fits the 40-entry traversal stack and
    answers closest-hit and shadow queries like the host SAH tree."""
    gen = capi.GeneratedScene(5, 5)
    gen.desc.bvh_builder = capi.PBRT_BVH_SAH_HOST
    sah = capi.Scene(gen.desc)
    gen.desc.bvh_builder = capi.PBRT_BVH_PLOC_GPU
    pl = capi.Scene(gen.desc)
    s, p = sah.info(), pl.info()
    print(f"city: {p.triangle_count} triangles; SAH depth {s.bvh_depth} {s.bvh_build_ms:.0f} ms (host) | "
          f"PLOC depth {p.bvh_depth} {p.bvh_build_ms:.1f} ms (device)")
    assert p.bvh_builder == capi.PBRT_BVH_PLOC_GPU and 1 < p.bvh_depth <= 40
    rays, _ = scene_rays(gen, 320, 180, 40000, 5)
    hs, os_ = intersect(cuda, sah, rays)
    hp, op = intersect(cuda, pl, rays)
    ts, tp = hs[:, 0].view(np.float32), hp[:, 0].view(np.float32)
    same_tri = (hs[:, 1] == hp[:, 1]) & (hs[:, 2] == hp[:, 2])
    tie = (hs[:, 1] != 0xFFFFFFFF) & (hp[:, 1] != 0xFFFFFFFF) & (np.abs(ts - tp) <= 1e-6 * np.abs(ts))
    print(f"rays {len(rays)}: same triangle {same_tri.mean():.6f}")
    assert np.all(same_tri | tie)
    assert np.array_equal(ts[same_tri], tp[same_tri])
    assert np.array_equal(os_, op)
    sah.close()
    pl.close()
