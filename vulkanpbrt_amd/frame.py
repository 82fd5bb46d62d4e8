"""The per-frame render pipeline over the C ABI, as the reference's main loop runs
it (source/VulkanPBRT.cpp:551-620: trace -> accumulator -> BMFR -> TAA, with the
accumulator's ping-pong history): device images held by torch, every stage one
C-ABI call on torch's current stream, optionally the whole frame captured into
one hipGraph (pbrt_frame_graph_*)."""
from __future__ import annotations

import ctypes as C

import numpy as np


# learned denoisers by frame option: 6 layers, the fused 2-layer and the compact network
CNN_KINDS = {"cnn": 6, "cnn2": 2, "cnnc": "c"}


class Frame:
    """Device images and the per-frame launch sequence (all on torch's current stream)."""

    def __init__(self, torch, capi, camera, gen, scene, w, h, depth, tracer, bmfr_fit, denoiser="bmfr"):
        self.torch, self.capi, self.camera = torch, capi, camera
        self.denoiser = denoiser
        self.gen, self.scene, self.w, self.h, self.depth = gen, scene, w, h, depth
        self.tracer = {"wavefront": capi.TRACER_WAVEFRONT, "persistent": capi.TRACER_PERSISTENT}.get(
            tracer, capi.TRACER_MEGAKERNEL)
        self.block_threads, self.queue_paths = 0, 0
        self.bmfr_threads = 0  # Gram-fit workgroup size (0: the library default)
        self.fit = capi.BMFR_FIT_GRAM_F64 if bmfr_fit == "gram" else capi.BMFR_FIT_HOUSEHOLDER
        d = "cuda"
        t = torch
        self.img = dict(
            depth=t.zeros((h, w), dtype=t.float32, device=d),
            normal=t.zeros((h, w, 2), dtype=t.float32, device=d),
            material=t.zeros((h, w, 4), dtype=t.uint8, device=d),
            albedo=t.zeros((h, w, 4), dtype=t.uint8, device=d),
            illumination=t.zeros((h, w, 4), dtype=t.float32, device=d),
        )
        # ping-pong history: index = frame parity (replaces the copy-back image copies)
        self.depth_hist = [t.zeros((h, w), dtype=t.float32, device=d) for _ in range(2)]
        self.spp = [t.zeros((h, w), dtype=t.uint8, device=d) for _ in range(2)]
        self.acc = [t.zeros((h, w, 4), dtype=t.int16, device=d) for _ in range(2)]
        self.motion = t.zeros((h, w, 2), dtype=t.int16, device=d)
        self.denoised = t.zeros((2, h, w, 4), dtype=t.int16, device=d)
        self.bmfr_final = t.zeros((h, w, 4), dtype=t.uint8, device=d)
        self.taa_hist = t.zeros((h, w, 4), dtype=t.uint8, device=d)
        self.final = t.zeros((h, w, 4), dtype=t.uint8, device=d)
        bp = capi.BmfrParams(w, h, 32, 0, self.fit)
        n = C.c_size_t()
        capi.check(capi.lib().pbrt_bmfr_workspace_size(C.byref(bp), C.byref(n)))
        self.ws = t.zeros(n.value, dtype=t.uint8, device=d)
        self.ws_bytes = n.value
        tn = C.c_size_t()
        capi.check(capi.lib().pbrt_trace_workspace_size(w, h, C.byref(tn)))
        self.trace_ws = t.zeros(tn.value, dtype=t.uint8, device=d)
        self.prev_view = None
        self.img_sets = None  # trace outputs, double-buffered by run_pipelined
        self.graph_nodes, self.graph_instantiations = 0, 0
        self.nets = {}
        if denoiser in CNN_KINDS:
            self.ensure_cnn(CNN_KINDS[denoiser])

    def ensure_cnn(self, layers=6):
        """A learned denoiser with its committed trained parameters: the 6-layer network,
        the fused 2-layer one or the compact one ("c")."""
        if layers in self.nets:
            return self.nets[layers]
        from vulkanpbrt_amd.denoiser_train import Network, weights_for
        net = Network(1, layers)
        net.load(weights_for(layers))
        self.nets[layers] = net
        if not hasattr(self, "cnn_ws"):
            cn = C.c_size_t()
            self.capi.check(self.capi.lib().pbrt_cnn_workspace_size(self.w, self.h, 1, 0, C.byref(cn)))
            self.cnn_ws = self.torch.zeros(cn.value, dtype=self.torch.uint8, device="cuda")
        return net

    def stream(self):
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def push_constants(self, frame):
        pose = self.gen.camera(frame)
        view, proj = self.camera.pose_matrices(pose, self.w, self.h)
        prev = self.prev_view if self.prev_view is not None else view
        return self.camera.push_constants(view, proj, prev, frame), view, proj

    def trace_params(self):
        return self.capi.TraceParams(self.w, self.h, self.depth, 0, self.capi.LIGHT_SAMPLE_AUTO,
                                     self.capi.OUTPUT_DEMOD_ILLUMINATION_FLOAT, 1, self.tracer, self.block_threads,
                                     self.queue_paths)

    def trace_images(self):
        capi = self.capi
        return capi.TraceImages(*(capi.ptr(self.img[k]) for k in ("depth", "normal", "material", "albedo",
                                                                   "illumination")),
                                capi.ptr(self.trace_ws), self.trace_ws.numel())

    def autotune(self, repeats=3):
        """pbrt_autotune_trace on the first frame; adopts the fastest launch shape."""
        capi = self.capi
        pc, _, _ = self.push_constants(0)
        tun = capi.TraceTuning()
        capi.check(capi.lib().pbrt_autotune_trace(self.scene.handle, C.byref(self.trace_params()), C.byref(pc),
                                                  C.byref(self.trace_images()), repeats, C.byref(tun), self.stream()))
        self.tracer, self.block_threads, self.queue_paths = tun.best.tracer, tun.best.block_threads, \
            tun.best.queue_paths
        names = {capi.TRACER_MEGAKERNEL: "megakernel", capi.TRACER_WAVEFRONT: "wavefront",
                 capi.TRACER_PERSISTENT: "persistent"}
        out = {"best": {"tracer": names[tun.best.tracer], "block_threads": tun.best.block_threads,
                        "queue_paths": tun.best.queue_paths},
               "candidates": [{"tracer": names[tun.tracer[i]], "block_threads": tun.block_threads[i],
                               "queue_paths": tun.queue_paths[i], "ms": round(tun.ms[i], 3)}
                              for i in range(tun.candidates)]}
        if self.fit != capi.BMFR_FIT_HOUSEHOLDER:
            # the Gram fit's tile shape, on a traced and accumulated first frame
            self.run(0)
            bt = capi.BmfrTuning()
            capi.check(capi.lib().pbrt_autotune_bmfr(C.byref(self.bmfr_params(0)), C.byref(self.bmfr_images(0)),
                                                     repeats, C.byref(bt), self.stream()))
            self.bmfr_threads = bt.best.fit_threads
            self.reset_history()
            out["bmfr"] = {"best_fit_threads": bt.best.fit_threads,
                           "candidates": [{"fit_threads": bt.fit_threads[i], "ms": round(bt.ms[i], 4)}
                                          for i in range(bt.candidates)]}
        return out

    def trace(self, pc, counters=None):
        capi = self.capi
        p = self.trace_params()
        ti = self.trace_images()
        cnt = capi.ptr(counters) if counters is not None else None
        capi.check(capi.lib().pbrt_trace(self.scene.handle, C.byref(p), C.byref(pc), C.byref(ti), cnt,
                                         self.stream()))

    def accumulate(self, frame, view, proj):
        capi, cam = self.capi, self.camera
        cur, prv = frame & 1, (frame & 1) ^ 1
        p = capi.AccumulateParams(self.w, self.h, frame, 0)
        p.inv_proj[:] = cam.colmajor(np.linalg.inv(proj)).tolist()
        p.inv_view[:] = cam.colmajor(np.linalg.inv(view)).tolist()
        pv = self.prev_view if self.prev_view is not None else view
        p.prev_view[:] = cam.colmajor(pv).tolist()
        o = np.linalg.inv(pv)[:, 3].astype(np.float32)
        o[3] = 1.0
        p.prev_origin[:] = o.tolist()
        # the G-buffer depth of this frame is written into depth_hist[cur] after the pass
        im = capi.AccumulateImages(capi.ptr(self.img["illumination"]), capi.ptr(self.img["depth"]),
                                   capi.ptr(self.depth_hist[prv]), capi.ptr(self.spp[prv]), capi.ptr(self.acc[prv]),
                                   capi.ptr(self.motion), capi.ptr(self.spp[cur]), capi.ptr(self.acc[cur]))
        capi.check(capi.lib().pbrt_accumulate(C.byref(p), C.byref(im), self.stream()))
        capi.check(capi.lib().pbrt_copy_back(self.w, self.h, capi.ptr(self.img["depth"]), None, None,
                                             capi.ptr(self.depth_hist[cur]), None, None, self.stream()))

    def reset_history(self):
        """Forget every previous frame (accumulation, BMFR and TAA history)."""
        for t in (*self.depth_hist, *self.spp, *self.acc, self.denoised, self.taa_hist, self.motion):
            t.zero_()
        self.prev_view = None

    def bmfr_params(self, frame):
        return self.capi.BmfrParams(self.w, self.h, 32, frame, self.fit, self.bmfr_threads)

    def bmfr_images(self, frame):
        capi = self.capi
        cur = frame & 1
        return capi.BmfrImages(capi.ptr(self.img["depth"]), capi.ptr(self.img["normal"]), capi.ptr(self.img["albedo"]),
                               capi.ptr(self.motion), capi.ptr(self.spp[cur]), capi.ptr(self.acc[cur]),
                               capi.ptr(self.denoised), capi.ptr(self.bmfr_final), capi.ptr(self.ws), self.ws_bytes)

    def bmfr(self, frame):
        capi = self.capi
        capi.check(capi.lib().pbrt_bmfr(C.byref(self.bmfr_params(frame)), C.byref(self.bmfr_images(frame)),
                                        self.stream()))

    def cnn(self, frame, layers=6, denoised=None):
        """The learned denoiser on this frame's accumulated illumination; the final image
        goes where BMFR's does, the demodulated radiance (RGBA32F) into `denoised` if given."""
        capi = self.capi
        cur = frame & 1
        im = capi.CnnImages(self.w, self.h, 1, capi.ptr(self.acc[cur]), capi.ptr(self.img["depth"]),
                            capi.ptr(self.img["normal"]), capi.ptr(self.img["albedo"]), capi.ptr(self.spp[cur]),
                            capi.ptr(self.motion), capi.ptr(denoised), capi.ptr(self.bmfr_final), None,
                            capi.ptr(self.cnn_ws), self.cnn_ws.numel())
        capi.check(capi.lib().pbrt_cnn_denoise(self.nets[layers].handle, C.byref(im), self.stream()))

    def denoise(self, frame):
        if self.denoiser in CNN_KINDS:
            self.cnn(frame, CNN_KINDS[self.denoiser])
        else:
            self.bmfr(frame)

    def taa(self, frame):
        capi = self.capi
        im = capi.TaaImages(capi.ptr(self.motion), capi.ptr(self.bmfr_final), capi.ptr(self.taa_hist),
                            capi.ptr(self.final))
        capi.check(capi.lib().pbrt_taa(self.w, self.h, frame, C.byref(im), self.stream()))

    def run_offline(self, frame, gbuffer, illumination, view, proj):
        """The reference's offline path (use_external_buffers, VulkanPBRT.cpp:224-246 and
        OfflineGBuffer::upload_to_g_buffer_command): the frame's G-buffer and illumination
        come from files (vulkanpbrt_amd.frame_io) instead of the tracer; accumulation,
        denoising and TAA then run as in run().  `view` / `proj` are the frame's camera
        matrices (row-major)."""
        t = self.torch
        for k in ("depth", "normal", "albedo"):
            self.img[k].copy_(t.from_numpy(np.ascontiguousarray(gbuffer[k])))
        self.img["illumination"].copy_(t.from_numpy(np.ascontiguousarray(illumination, dtype=np.float32)))
        self.accumulate(frame, view, proj)
        self.denoise(frame)
        self.taa(frame)
        self.prev_view = view

    def run_pipelined(self, first, count, post_stream):
        """Frames first .. first + count - 1, each frame's accumulation, denoising and TAA
        on `post_stream` while the next frame is traced on the current stream, so the
        post chain of one frame runs beside the trace of the next.  The trace outputs
        are double-buffered (frame f writes set f & 1) and frame f's trace waits until
        the post chain of frame f - 2 is done with that set; the post chain stays in
        order on its stream, which orders the accumulation, BMFR and TAA histories.
        The images equal those of run() frame by frame; on return the current stream
        is ordered after the last frame's post chain."""
        t = self.torch
        trace_stream = t.cuda.current_stream()
        if self.img_sets is None:
            self.img_sets = [self.img, {k: t.zeros_like(v) for k, v in self.img.items()}]
        post_stream.wait_stream(trace_stream)
        posted = []
        for k in range(count):
            f = first + k
            self.img = self.img_sets[f & 1]
            if k >= 2:
                trace_stream.wait_event(posted[k - 2])
            pc, view, proj = self.push_constants(f)
            self.trace(pc)
            traced = t.cuda.Event()
            traced.record(trace_stream)
            with t.cuda.stream(post_stream):
                post_stream.wait_event(traced)
                self.accumulate(f, view, proj)
                self.denoise(f)
                self.taa(f)
                done = t.cuda.Event()
                done.record(post_stream)
            posted.append(done)
            self.prev_view = view
        trace_stream.wait_stream(post_stream)

    def run(self, frame, events=None, graph=None):
        """One frame; with `graph` (a pbrt_frame_graph handle) its launches are captured
        into one hipGraph and launched together (no per-stage events then)."""
        if graph is not None:
            lib = self.capi.lib()
            self.capi.check(lib.pbrt_frame_graph_begin(graph, self.stream()))
            try:
                self.run(frame)
            except BaseException:
                # close the open capture; the handle stays the caller's to destroy
                lib.pbrt_frame_graph_abort(graph)
                raise
            nodes, inst = C.c_uint32(), C.c_uint32()
            self.capi.check(lib.pbrt_frame_graph_end(graph, self.stream(), C.byref(nodes), C.byref(inst)))
            self.graph_nodes, self.graph_instantiations = nodes.value, self.graph_instantiations + inst.value
            return
        pc, view, proj = self.push_constants(frame)
        if events:
            events[0].record()
        self.trace(pc)
        if events:
            events[1].record()
        self.accumulate(frame, view, proj)
        if events:
            events[2].record()
        self.denoise(frame)
        if events:
            events[3].record()
        self.taa(frame)
        if events:
            events[4].record()
        self.prev_view = view
