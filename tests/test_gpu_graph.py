"""Whole-frame hipGraph (pbrt_frame_graph_*): a frame captured from the trace,
accumulator, denoiser and TAA entry points and launched as one graph gives the same
bytes as the same frame launched eagerly, frame after frame (each graph launch
is checked against its eager twin, with eager work and synchronisations between
launches), and the executable graph is updated in place rather than rebuilt."""
import ctypes as C

import numpy as np
import pytest

from vulkanpbrt_amd import camera, capi
from vulkanpbrt_amd.frame import Frame

pytestmark = pytest.mark.gpu


def make_frame(torch, tracer, denoiser, gen, scene):
    return Frame(torch, capi, camera, gen, scene, 256, 256, 2, tracer, "gram", denoiser)


def snapshot(fr):
    return {k: v.cpu().numpy().copy() for k, v in (("final", fr.final), ("illumination", fr.img["illumination"]),
                                                     ("bmfr_final", fr.bmfr_final))}


@pytest.mark.parametrize("tracer,denoiser", [("wavefront", "bmfr"), ("megakernel", "bmfr"), ("persistent", "bmfr"), ("megakernel", "cnnc"), ("wavefront", "cnn"),
                                             ("wavefront", "cnn2")])
def test_frame_graph_equals_eager(cuda, tracer, denoiser):
    torch = cuda
    gen = capi.GeneratedScene(1, 1)
    scene = capi.Scene(gen.desc)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        eager = make_frame(torch, tracer, denoiser, gen, scene)
        graphed = make_frame(torch, tracer, denoiser, gen, scene)
        g = C.c_void_p()
        capi.check(capi.lib().pbrt_frame_graph_create(C.byref(g)))
        try:
            for f in range(5):
                eager.run(f)
                torch.cuda.synchronize()
                want = snapshot(eager)
                graphed.run(f, graph=g)
                torch.cuda.synchronize()
                got = snapshot(graphed)
                for k in want:
                    assert np.array_equal(want[k], got[k]), f"frame {f}: {k} differs"
            print(tracer, denoiser, "graph kernel nodes", graphed.graph_nodes, "instantiations",
                  graphed.graph_instantiations)
            assert graphed.graph_nodes >= 4
            # frame 0 and 1 differ in topology only if the ping-pong history changes the
            # launches; the camera and frame number alone are argument updates
            assert graphed.graph_instantiations <= 2
            assert want["final"][..., :3].mean() > 5
        finally:
            capi.check(capi.lib().pbrt_frame_graph_destroy(g))


def test_frame_graph_errors(cuda):
    torch = cuda
    lib = capi.lib()
    g = C.c_void_p()
    capi.check(lib.pbrt_frame_graph_create(C.byref(g)))
    assert lib.pbrt_frame_graph_begin(g, None) != 0  # the null stream cannot be captured
    assert lib.pbrt_frame_graph_end(g, C.c_void_p(torch.cuda.Stream().cuda_stream), None, None) != 0  # no capture
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    capi.check(lib.pbrt_frame_graph_begin(g, C.c_void_p(s1.cuda_stream)))
    assert lib.pbrt_frame_graph_begin(g, C.c_void_p(s1.cuda_stream)) != 0  # already open
    assert lib.pbrt_frame_graph_end(g, C.c_void_p(s2.cuda_stream), None, None) != 0  # other stream: abandoned
    # the stream is usable again after the abandoned capture
    x = torch.ones(16, device="cuda")
    with torch.cuda.stream(s1):
        y = x * 2
    torch.cuda.synchronize()
    assert y.sum().item() == 32
    capi.check(lib.pbrt_frame_graph_destroy(g))


def test_frame_failing_inside_capture_keeps_the_graph(cuda):
    """A frame whose launches fail on the host while its capture is open (here TAA
    rejects a zero-size image) raises, closes the capture with pbrt_frame_graph_abort
    and leaves the handle with the caller: the next frame captures and launches
    through the same graph, and the caller's single destroy succeeds."""
    torch = cuda
    lib = capi.lib()
    gen = capi.GeneratedScene(1, 1)
    scene = capi.Scene(gen.desc)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        eager = make_frame(torch, "megakernel", "bmfr", gen, scene)
        graphed = make_frame(torch, "megakernel", "bmfr", gen, scene)
        g = C.c_void_p()
        capi.check(lib.pbrt_frame_graph_create(C.byref(g)))
        try:
            good_taa = graphed.taa

            def bad_taa(frame):
                images = capi.TaaImages(None, None, None, None)
                capi.check(lib.pbrt_taa(0, 0, frame, C.byref(images), graphed.stream()))

            graphed.taa = bad_taa
            with pytest.raises(capi.PbrtError):
                graphed.run(0, graph=g)
            graphed.taa = good_taa
            assert lib.pbrt_frame_graph_abort(g) == 0  # nothing open any more: a no-op
            for f in range(2):
                eager.run(f)
                graphed.run(f, graph=g)
            torch.cuda.synchronize()
            assert np.array_equal(snapshot(eager)["final"], snapshot(graphed)["final"])
        finally:
            assert lib.pbrt_frame_graph_destroy(g) == 0


def history(fr):
    out = {k: v.cpu().numpy().copy() for k, v in (("final", fr.final), ("bmfr_final", fr.bmfr_final),
                                                   ("denoised", fr.denoised), ("taa_hist", fr.taa_hist),
                                                   ("motion", fr.motion), ("illumination", fr.img["illumination"]),
                                                   ("depth", fr.img["depth"]))}
    for i in range(2):
        out[f"acc{i}"], out[f"spp{i}"] = fr.acc[i].cpu().numpy().copy(), fr.spp[i].cpu().numpy().copy()
    return out


@pytest.mark.parametrize("tracer,denoiser", [("megakernel", "bmfr"), ("wavefront", "bmfr"), ("megakernel", "cnn2")])
def test_pipelined_frames_equal_one_stream_frames(cuda, tracer, denoiser):
    """Frame.run_pipelined (trace of frame f+1 beside the post chain of frame f on a
    second stream, trace outputs double-buffered) leaves the same bytes as run() frame
    after frame on one stream: the final and denoised images and every history."""
    torch = cuda
    gen = capi.GeneratedScene(1, 1)
    scene = capi.Scene(gen.desc)
    side, post = torch.cuda.Stream(), torch.cuda.Stream()
    with torch.cuda.stream(side):
        one = make_frame(torch, tracer, denoiser, gen, scene)
        two = make_frame(torch, tracer, denoiser, gen, scene)
        for f in range(6):
            one.run(f)
        two.run_pipelined(0, 6, post)
        torch.cuda.synchronize()
        want, got = history(one), history(two)
        for k in want:
            assert np.array_equal(want[k], got[k]), f"after 6 frames: {k} differs"
        # a second call continues the sequence (the set written last is reused first)
        for f in range(6, 9):
            one.run(f)
        two.run_pipelined(6, 3, post)
        torch.cuda.synchronize()
        want, got = history(one), history(two)
        for k in want:
            assert np.array_equal(want[k], got[k]), f"after 9 frames: {k} differs"
        assert want["final"][..., :3].mean() > 5
