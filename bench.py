#!/usr/bin/env python3
"""Frame benchmark: 1080p trace + denoise on one MI355X.

A step is one frame of the atrium workload (seed-generated ~267k triangles,
1920x1080, 1 spp, 4 bounces, SURFACE_STRENGTH light sampling): path tracing,
temporal accumulation, BMFR (32x32 blocks) and TAA, with the camera moving along
the scene's 64-frame spline.  Inputs (scene, BVH, textures, images) are resident
in HBM before the timed region.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--config 2] [--tracer megakernel|wavefront]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 5])
    ap.add_argument("--tracer", default="wavefront", choices=["megakernel", "wavefront"])
    ap.add_argument("--bmfr-fit", default="gram", choices=["gram", "householder"])
    ap.add_argument("--autotune", action="store_true",
                    help="pick the tracer launch shape (pbrt_autotune_trace) on the first frame before timing")
    ap.add_argument("--bvh", default=None, choices=["sah", "lbvh"],
                    help="BVH builder: binned SAH on the host or the Morton-code LBVH on the GPU "
                         "(default: SAH, LBVH for the 5M-triangle city)")
    ap.add_argument("--denoiser", default="bmfr", choices=["bmfr", "cnn"],
                    help="denoiser of the timed frame: BMFR 32x32 or the learned kernel-predicting network")
    ap.add_argument("--no-denoise-eval", action="store_true",
                    help="skip the learned-vs-BMFR relMSE evaluation on held-out Cornell-box scenes")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB three times (the reference kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes) and exit")
    return ap.parse_args()


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summaries (lines `traffic_bytes_per_launch: <kernel> <bytes>` under profiles/)."""
    import re
    for path in sorted((ROOT / "profiles").glob("*.md")):
        for line in path.read_text().splitlines():
            m = re.match(r"traffic_bytes_per_launch:\s+(\S+)\s+(\d+)", line)
            if m and m.group(1) == kernel:
                return int(m.group(2)), path.name
    return None, None


def copy_calibration(torch, mib):
    n = mib << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n, "copy_calibration_bytes_written": n, "copies": 3}))


WORKLOADS = {
    1: dict(name="cornell box + 16 icospheres (5.1k tris), 256x256, 1 spp, 2 bounces, accumulator + BMFR + TAA",
            width=256, height=256, depth=2, seed=1),
    2: dict(name="atrium (~267k tris, 8 textures, 3 area quads + sun), 1920x1080, 1 spp, 4 bounces, "
                 "accumulator + BMFR 32x32 + TAA over the 64-frame camera spline",
            width=1920, height=1080, depth=4, seed=2),
    5: dict(name="city stress (~4.9M tris: 4096 towers, 65,536 emissive window quads = 131k triangle lights, "
                 "UNIFORM light sampling), 3840x2160, 1 spp, 8 bounces with Russian roulette, accumulator + "
                 "BMFR + TAA over the 128-frame camera path",
            width=3840, height=2160, depth=8, seed=5, bvh="lbvh"),
}


class Frame:
    """Device images and the per-frame launch sequence (all on torch's current stream)."""

    def __init__(self, torch, capi, camera, gen, scene, w, h, depth, tracer, bmfr_fit, denoiser="bmfr"):
        self.torch, self.capi, self.camera = torch, capi, camera
        self.denoiser = denoiser
        self.gen, self.scene, self.w, self.h, self.depth = gen, scene, w, h, depth
        self.tracer = capi.TRACER_WAVEFRONT if tracer == "wavefront" else capi.TRACER_MEGAKERNEL
        self.block_threads, self.queue_paths = 0, 0
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
        self.net = None
        if denoiser == "cnn":
            self.ensure_cnn()

    def ensure_cnn(self):
        """The learned denoiser with the committed trained parameters."""
        if self.net is not None:
            return
        from vulkanpbrt_amd.denoiser_train import DEFAULT_WEIGHTS, Network
        self.net = Network()
        self.net.load(DEFAULT_WEIGHTS)
        cn = C.c_size_t()
        self.capi.check(self.capi.lib().pbrt_cnn_workspace_size(self.w, self.h, 1, 0, C.byref(cn)))
        self.cnn_ws = self.torch.zeros(cn.value, dtype=self.torch.uint8, device="cuda")

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
        names = {capi.TRACER_MEGAKERNEL: "megakernel", capi.TRACER_WAVEFRONT: "wavefront"}
        return {"best": {"tracer": names[tun.best.tracer], "block_threads": tun.best.block_threads,
                         "queue_paths": tun.best.queue_paths},
                "candidates": [{"tracer": names[tun.tracer[i]], "block_threads": tun.block_threads[i],
                                "queue_paths": tun.queue_paths[i], "ms": round(tun.ms[i], 3)}
                               for i in range(tun.candidates)]}

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

    def bmfr(self, frame):
        capi = self.capi
        cur = frame & 1
        p = capi.BmfrParams(self.w, self.h, 32, frame, self.fit)
        im = capi.BmfrImages(capi.ptr(self.img["depth"]), capi.ptr(self.img["normal"]), capi.ptr(self.img["albedo"]),
                             capi.ptr(self.motion), capi.ptr(self.spp[cur]), capi.ptr(self.acc[cur]),
                             capi.ptr(self.denoised), capi.ptr(self.bmfr_final), capi.ptr(self.ws), self.ws_bytes)
        capi.check(capi.lib().pbrt_bmfr(C.byref(p), C.byref(im), self.stream()))

    def cnn(self, frame):
        capi = self.capi
        cur = frame & 1
        im = capi.CnnImages(self.w, self.h, 1, capi.ptr(self.acc[cur]), capi.ptr(self.img["depth"]),
                            capi.ptr(self.img["normal"]), capi.ptr(self.img["albedo"]), capi.ptr(self.spp[cur]),
                            capi.ptr(self.motion), None, capi.ptr(self.bmfr_final), None, capi.ptr(self.cnn_ws),
                            self.cnn_ws.numel())
        capi.check(capi.lib().pbrt_cnn_denoise(self.net.handle, C.byref(im), self.stream()))

    def denoise(self, frame):
        if self.denoiser == "cnn":
            self.cnn(frame)
        else:
            self.bmfr(frame)

    def taa(self, frame):
        capi = self.capi
        im = capi.TaaImages(capi.ptr(self.motion), capi.ptr(self.bmfr_final), capi.ptr(self.taa_hist),
                            capi.ptr(self.final))
        capi.check(capi.lib().pbrt_taa(self.w, self.h, frame, C.byref(im), self.stream()))

    def run(self, frame, events=None):
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


def denoise_comparison(torch, fr, frame, runs, evaluate):
    """Both denoisers on the same 1080p frame (the last timed one), `runs` launches each,
    timed with events on the launch stream; with `evaluate`, their relMSE against
    256-spp references on 16 held-out Cornell-box scenes (128x128, first frame)."""
    fr.ensure_cnn()
    out = {}
    for name, fn in (("bmfr", fr.bmfr), ("cnn", fr.cnn)):
        fn(frame)  # warm-up
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        for _ in range(runs):
            fn(frame)
        b.record()
        torch.cuda.synchronize()
        out[f"{name}_ms"] = round(a.elapsed_time(b) / runs, 4)
    out["cnn_speedup_vs_bmfr"] = round(out["bmfr_ms"] / out["cnn_ms"], 3)
    if evaluate:
        from vulkanpbrt_amd import denoiser_train as dt
        pairs = dt.render_pairs(torch, np.arange(90_000, 90_016), 128, 128, 256)
        res = dt.evaluate(torch, fr.net, pairs)
        out.update({k: round(v, 5) for k, v in res.items()})
        out["eval_set"] = ("16 held-out seed-generated Cornell-box scenes (seeds 90000-90015), 128x128, 1-spp "
                           "accumulator input, 256-spp references; relMSE of the remodulated radiance "
                           "(and of the demodulated radiance)")
    return out


def cpu_baseline(gen, cfg, budget_s, frames):
    """The oracle's FP32 restatement on the host cores, over bands of 1/8 of the image
    rows (from the middle, then the rest) of the timed camera frames, until the time
    budget is spent (at least one band)."""
    from oracle import oracle
    from vulkanpbrt_amd import camera
    osc = oracle.OracleScene(gen.desc)
    w, h = cfg["width"], cfg["height"]
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
    band = max(1, h // 8)
    starts = [(4 + k) % 8 * band for k in range(8)]
    total_rays, total_t, px, done = 0, 0.0, 0, set()
    for f in frames:
        view, proj = camera.pose_matrices(gen.camera(f), w, h)
        pc = camera.push_constants(view, proj, view, f)
        for y0 in starts:
            y1 = min(h, y0 + band)
            t0 = time.perf_counter()
            r = osc.trace(w, h, np.array(pc.view_inverse[:], np.float32), np.array(pc.proj_inverse[:], np.float32),
                          frame=f, max_depth=cfg["depth"], rows=(y0, y1), threads=threads)
            total_t += time.perf_counter() - t0
            total_rays += r["rays"] + r["shadow_rays"]
            px += (y1 - y0) * w
            done.add(f)
            if total_t >= budget_s:
                break
        if total_t >= budget_s:
            break
    return dict(value=round(total_rays / total_t / 1e6, 4), unit="Mrays/s", cores=threads, kind="port",
                sample=f"{px / (w * h):.2f} frames' worth of {w}x{h} pixels in row bands of camera frames "
                       f"{min(done)}..{max(done)} of the timed run, FP32 oracle restatement with FP64 BVH queries, "
                       f"{total_rays / 1e6:.1f} M rays in {total_t:.1f} s")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported in this build "
              "(multi-GPU sharding is design-only).", file=sys.stderr)
        sys.exit(2)
    import torch
    from vulkanpbrt_amd import camera, capi
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(torch, args.copy_calibration)
        return
    cfg = WORKLOADS[args.config]
    w, h = cfg["width"], cfg["height"]
    t_setup = time.perf_counter()
    gen = capi.GeneratedScene(args.config, cfg["seed"])
    bvh = args.bvh or cfg.get("bvh", "sah")
    gen.desc.bvh_builder = capi.PBRT_BVH_LBVH_GPU if bvh == "lbvh" else capi.PBRT_BVH_SAH_HOST
    scene = capi.Scene(gen.desc)
    info = scene.info()
    fr = Frame(torch, capi, camera, gen, scene, w, h, cfg["depth"], args.tracer, args.bmfr_fit, args.denoiser)
    setup_s = time.perf_counter() - t_setup

    tuning = fr.autotune() if args.autotune else None
    for f in range(args.warmup):
        fr.run(f)
    torch.cuda.synchronize()

    K = args.steps
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(K)]
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    lib = capi.lib()
    torch.cuda.synchronize()
    # per-stage tracer time: HIP events recorded by the library on the launch stream
    capi.check(lib.pbrt_trace_timing_begin(scene.handle))
    start.record()
    for k in range(K):
        fr.run(args.warmup + k, ev[k])
    end.record()
    torch.cuda.synchronize()
    timing = capi.TraceTiming()
    capi.check(lib.pbrt_trace_timing_end(scene.handle, C.byref(timing)))
    total_ms = start.elapsed_time(end)
    stage = np.array([[ev[k][i].elapsed_time(ev[k][i + 1]) for i in range(4)] for k in range(K)])
    stage_ms = stage.mean(axis=0)
    ms_per_step = total_ms / K

    # rays and traversal work of exactly the timed frames: the counting variant of the
    # tracer replays each timed frame after the timed region (same camera and frame
    # index, so the same path decisions)
    counters = torch.zeros(6, dtype=torch.int64, device="cuda")
    for k in range(K):
        pc, _, _ = fr.push_constants(args.warmup + k)
        fr.trace(pc, counters)
    torch.cuda.synchronize()
    cnt = counters.cpu().numpy().astype(np.float64)
    rays = cnt[0] + cnt[1]
    rays_per_frame = rays / K
    mrays = rays / (total_ms / 1e3) / 1e6
    nominal = 1000.0 / ms_per_step * (2 * cfg["depth"]) * w * h / 1e6  # Gui.hpp:41-43 formula

    # roofline of the dominant kernel.  Algorithmic bytes: 48 B ray I/O per traced
    # ray + 64 B per BVH node visit + 48 B per triangle record tested (+ 36 B/px of
    # G-buffer and illumination output for the megakernel, which also shades).
    ext_bytes = 48.0 * cnt[0] + 64.0 * cnt[2] + 48.0 * cnt[3]
    sh_bytes = 48.0 * cnt[1] + 64.0 * cnt[4] + 48.0 * cnt[5]
    kernels = {}
    if fr.tracer == capi.TRACER_WAVEFRONT:
        for name, b, ms, n in (("k_wf_extend", ext_bytes, timing.extend_ms, timing.extend_launches),
                               ("k_wf_shadow", sh_bytes, timing.shadow_ms, timing.shadow_launches)):
            kernels[name] = dict(ms_per_frame=round(ms / K, 3), launches=int(n),
                                 avg_launch_us=round(ms / max(1, n) * 1e3, 1),
                                 gbs=round(b / (ms / 1e3) / 1e9, 1))
        kernels["k_wf_shade"] = dict(ms_per_frame=round(timing.shade_ms / K, 3), launches=int(timing.shade_launches))
        kernels["other"] = dict(ms_per_frame=round(timing.other_ms / K, 3))
        dom, dom_bytes, dom_ms = "k_wf_extend", ext_bytes, timing.extend_ms
    else:
        dom, dom_bytes, dom_ms = "pt_megakernel", ext_bytes + sh_bytes + 36.0 * w * h * K, timing.megakernel_ms
        kernels["pt_megakernel"] = dict(ms_per_frame=round(dom_ms / K, 3), launches=int(timing.megakernel_launches))
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    traffic, traffic_src = measured_traffic(dom)
    n_launch = timing.extend_launches if fr.tracer == capi.TRACER_WAVEFRONT else timing.megakernel_launches
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, kernel=dom,
                    algorithmic_bytes_per_launch=round(dom_bytes / max(1, n_launch)),
                    avg_launch_us=round(dom_ms / max(1, n_launch) * 1e3, 1),
                    traffic_source=f"profiles/{traffic_src}" if traffic_src else None)

    out = {
        "metric": "1080p trace+denoise ms & Mrays/s; learned vs BMFR denoise ms at equal relMSE",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seed-generated scene, random-free camera spline)",
        "config": {"workload": cfg["name"], "width": w, "height": h, "spp": 1, "max_depth": cfg["depth"],
                   "triangles": info.triangle_count, "lights": info.light_count, "tracer": args.tracer,
                   "denoiser": args.denoiser, "bmfr_fit": args.bmfr_fit, "seed": cfg["seed"]},
        "stage_ms": {"trace": round(stage_ms[0], 3), "accumulate": round(stage_ms[1], 3),
                     "denoise": round(stage_ms[2], 3), "taa": round(stage_ms[3], 3)},
        "rays_per_frame": int(rays_per_frame),
        "trace_kernels": kernels,
        "nominal_mrays_reference_formula": round(nominal, 2),
        "bvh": {"builder": bvh, "build_ms": round(info.bvh_build_ms, 2),
                "nodes": info.bvh_node_count, "depth": info.bvh_depth,
                "node_visits_per_closest_ray": round(cnt[2] / max(1.0, cnt[0]), 2),
                "tri_tests_per_closest_ray": round(cnt[3] / max(1.0, cnt[0]), 2),
                "node_visits_per_shadow_ray": round(cnt[4] / max(1.0, cnt[1]), 2),
                "tri_tests_per_shadow_ray": round(cnt[5] / max(1.0, cnt[1]), 2)},
        "setup_s": round(setup_s, 2),
        "roofline": roofline,
    }
    if tuning:
        out["autotune"] = tuning
        out["config"]["tracer"] = tuning["best"]["tracer"]
    out["denoise_comparison"] = denoise_comparison(torch, fr, args.warmup + K - 1, max(K, 10),
                                                   not args.no_denoise_eval)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(gen, cfg, args.cpu_seconds, range(args.warmup, args.warmup + K))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
