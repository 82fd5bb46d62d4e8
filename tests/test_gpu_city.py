"""The city stress workload (about 4.9 M triangles from 4096 instanced towers; 65,536
emissive window quads = 131,072 triangle lights, so light sampling is UNIFORM,
PBRTPipeline.cpp:92-96 and lighting.glsl:170-227; 8 bounces with Russian roulette from
the first bounce, lighting.glsl:274-280).

* The city's camera at 96x54 against the oracle's FP32 mirror: both tracers find the
  same primary hits, the same albedo, the mirror's radiance on most pixels, and trace
  the oracle's number of rays to 1 %.
* One full 3840x2160 frame (trace, accumulator, BMFR, TAA) captured into a hipGraph
  equals the same frame launched eagerly, byte for byte, over two frames; the output is
  finite and non-negative.
"""
import ctypes as C

import numpy as np
import pytest

from vulkanpbrt_amd import camera, capi
from vulkanpbrt_amd.frame import Frame
from tests import helpers

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def city(oracle_lib, pbrt_lib):
    gen = capi.GeneratedScene(5, 5)
    gen.desc.bvh_builder = capi.PBRT_BVH_SAH_HOST
    scene = capi.Scene(gen.desc, keepalive=gen)
    osc = oracle_lib.OracleScene(gen.desc)
    assert scene.info().light_count > 800  # UNIFORM sampling
    return gen, scene, osc


@pytest.mark.parametrize("tracer", [capi.TRACER_MEGAKERNEL, capi.TRACER_PERSISTENT, capi.TRACER_WAVEFRONT])
def test_city_matches_fp32_mirror(cuda, city, tracer):
    gen, scene, osc = city
    w, h, frame = 96, 54, 3
    pc, _, _ = helpers.frame_push_constants(gen, w, h, frame)
    img = helpers.gbuffer(cuda, w, h)
    cnt = cuda.zeros(6, dtype=cuda.int64, device="cuda")
    helpers.trace(cuda, scene, pc, img, w, h, max_depth=8, tracer=tracer, counters=cnt)
    cuda.cuda.synchronize()
    gpu = helpers.to_np(img)
    ref = osc.trace(w, h, np.array(pc.view_inverse[:], np.float32), np.array(pc.proj_inverse[:], np.float32),
                    frame=frame, max_depth=8, precision=32, light_mode=capi.LIGHT_SAMPLE_UNIFORM)
    d_ok = np.isclose(gpu["depth"], ref["depth"], rtol=1e-4, atol=1e-6)
    alb_ok = np.all(gpu["albedo"] == ref["albedo"], axis=-1)
    gi, ri = gpu["illumination"][..., :3], ref["illumination"][..., :3]
    rel = np.abs(gi - ri) / np.maximum(np.maximum(np.abs(gi), np.abs(ri)), 1e-3)
    rad_ok = np.all(rel <= 1e-4, axis=-1)
    exact = np.all(gi == ri, axis=-1)
    c = cnt.cpu().numpy()
    print(f"city tracer {tracer}: depth {d_ok.mean():.4f} albedo {alb_ok.mean():.4f} radiance {rad_ok.mean():.4f} "
          f"(bit-identical {exact.mean():.4f}); "
          f"rays gpu {c[0]} + {c[1]} shadow, oracle {ref['rays']} + {ref['shadow_rays']}; "
          f"mean radiance gpu {gi[ri < 999].mean():.5f} oracle {ri[ri < 999].mean():.5f}")
    assert d_ok.mean() >= 0.999 and alb_ok.mean() >= 0.999
    assert rad_ok.mean() >= 0.999 and exact.mean() >= 0.999
    assert abs(int(c[0]) - ref["rays"]) <= 0.01 * ref["rays"]
    assert abs(int(c[1]) - ref["shadow_rays"]) <= 0.01 * ref["shadow_rays"]


def test_city_2160p_graph_frame_equals_eager(cuda, city):
    torch = cuda
    gen, scene, _ = city
    w, h = 3840, 2160
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        eager = Frame(torch, capi, camera, gen, scene, w, h, 8, "wavefront", "gram")
        graphed = Frame(torch, capi, camera, gen, scene, w, h, 8, "wavefront", "gram")
        g = C.c_void_p()
        capi.check(capi.lib().pbrt_frame_graph_create(C.byref(g)))
        try:
            for f in range(2):
                eager.run(f)
                torch.cuda.synchronize()
                want = {k: v.cpu().numpy().copy() for k, v in (("final", eager.final),
                                                                ("illumination", eager.img["illumination"]))}
                graphed.run(f, graph=g)
                torch.cuda.synchronize()
                for k, v in (("final", graphed.final), ("illumination", graphed.img["illumination"])):
                    assert np.array_equal(want[k], v.cpu().numpy()), f"frame {f}: {k} differs"
                ill = want["illumination"][..., :3]
                assert np.isfinite(ill).all() and (ill >= 0).all()
                print(f"city 2160p frame {f}: graph == eager; final mean {want['final'][..., :3].mean():.2f}, "
                      f"kernel nodes {graphed.graph_nodes}")
            assert want["final"][..., :3].mean() > 1
        finally:
            capi.check(capi.lib().pbrt_frame_graph_destroy(g))
