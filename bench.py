#!/usr/bin/env python3
"""Frame benchmark: 1080p trace + denoise on one MI355X.

A step is one frame of the atrium workload (seed-generated ~267k triangles,
1920x1080, 1 spp, 4 bounces, SURFACE_STRENGTH light sampling): path tracing,
temporal accumulation, BMFR (32x32 blocks) and TAA, with the camera moving along
the scene's 64-frame spline.  Inputs (scene, BVH, textures, images) are resident
in HBM before the timed region.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--config 2] [--tracer megakernel|persistent|wavefront]
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
NODE_BYTES = 32.0  # one quantized BVH2 node


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 5])
    ap.add_argument("--tracer", default="wavefront", choices=["megakernel", "persistent", "wavefront"],
                    help="tracer timed with --no-autotune")
    ap.add_argument("--bmfr-fit", default="gram", choices=["gram", "householder"])
    ap.add_argument("--autotune", dest="autotune", action="store_true", default=True,
                    help="(default) pick the tracer and its launch shape (pbrt_autotune_trace) and BMFR's Gram "
                         "tile shape (pbrt_autotune_bmfr) on the first frame before timing")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="time --tracer with the library's default launch shapes")
    ap.add_argument("--bvh", default=None, choices=["sah", "sbvh", "lbvh", "ploc"],
                    help="BVH builder: binned SAH on the host, SAH with spatial splits on the host, or the "
                         "Morton-code LBVH on the GPU (default: spatial splits; binned SAH for the 5M-triangle "
                         "city, whose small triangles gain nothing from spatial splits)")
    ap.add_argument("--denoiser", default="bmfr", choices=["bmfr", "cnn", "cnn2", "cnnc"],
                    help="denoiser of the timed frame: BMFR 32x32, the 6-layer learned kernel-predicting network "
                         "or the fused 2-layer one")
    ap.add_argument("--no-denoise-eval", action="store_true",
                    help="skip the learned-vs-BMFR relMSE evaluation on held-out Cornell-box scenes")
    ap.add_argument("--denoise-quality", action="store_true",
                    help="instead of timing: relMSE of BMFR (with its temporal history) and of the learned "
                         "denoisers over the workload's camera sequence at its full size, against references "
                         "of --quality-spp samples per pixel; prints one JSON line")
    ap.add_argument("--quality-frames", type=int, default=16, help="frames of the --denoise-quality sequence")
    ap.add_argument("--quality-score", default="0,15", help="frames scored by --denoise-quality (comma list)")
    ap.add_argument("--quality-spp", type=int, default=4096, help="samples per pixel of the references")
    ap.add_argument("--graph", dest="graph", action="store_true", default=None,
                    help="capture each frame (trace, accumulate, denoise, TAA) into one hipGraph and launch it "
                         "(default for the atrium and the city; the per-kernel breakdown then comes from an eager pass over "
                         "the same frames after the timed region)")
    ap.add_argument("--eager", dest="graph", action="store_false", help="launch the frame's kernels one by one")
    ap.add_argument("--pipeline", dest="pipeline", action="store_true", default=False,
                    help="trace each frame on one stream while the previous frame's accumulation, denoising and "
                         "TAA run on a second one (trace outputs double-buffered; eager launches only); the "
                         "per-stage breakdown then comes from a one-stream pass over the same frames.  Off by "
                         "default: on the atrium the post chain beside the megakernel costs more than it hides")
    ap.add_argument("--no-pipeline", dest="pipeline", action="store_false",
                    help="(default) run each frame's stages in order on one stream")
    ap.add_argument("--scene", default=None, metavar="OBJ",
                    help="trace a Wavefront OBJ/MTL file (pbrt_load_obj) instead of the generated scene, with the "
                         "workload's camera path, image size and explicit lights")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB three times (the reference kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes) and exit")
    return ap.parse_args()


def profile_value(key, kernel, workload):
    """A figure for `kernel` on `workload` (e.g. "atrium-1080p-sbvh") recorded in a committed
    rocprofv3 summary under profiles/ as a line `<key>: <kernel> <workload> <value>`;
    returns (value, file name) or (None, None)."""
    import re
    for path in sorted((ROOT / "profiles").glob("*.md")):
        for line in path.read_text().splitlines():
            m = re.match(rf"{key}:\s+(\S+)\s+(\S+)\s+([0-9.eE+-]+)\s*$", line)
            if m and m.group(1) == kernel and m.group(2) == workload:
                return float(m.group(3)), path.name
    return None, None


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summaries (`traffic_bytes_per_launch:` lines)."""
    v, src = profile_value("traffic_bytes_per_launch", kernel, workload)
    return (int(v), src) if v is not None else (None, None)


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
            width=256, height=256, depth=2, seed=1, tag="cornell-256"),
    2: dict(name="atrium (~262k tris, 8 textures, 3 area quads + sun), 1920x1080, 1 spp, 4 bounces, "
                 "accumulator + BMFR 32x32 + TAA over the 64-frame camera spline",
            width=1920, height=1080, depth=4, seed=2, graph=True, tag="atrium-1080p"),
    5: dict(name="city stress (~4.9M tris: 4096 towers, 65,536 emissive window quads = 131k triangle lights, "
                 "UNIFORM light sampling), 3840x2160, 1 spp, 8 bounces with Russian roulette, accumulator + "
                 "BMFR + TAA over the 128-frame camera path",
            width=3840, height=2160, depth=8, seed=5, bvh="sah", graph=True, tag="city-2160p"),
}


def denoise_comparison(torch, fr, frame, runs, evaluate):
    """BMFR and the learned denoisers on the same frame (the last timed one), `runs`
    launches each, timed with events on the launch stream; with `evaluate`, their
    relMSE on the timed workload itself: the first 8 frames of its camera path through
    accumulate -> denoise, frames 0 and 7 scored against 1024-spp references (BMFR
    with its temporal history, the networks on the same accumulated input; the full
    4096-spp evaluation is `--denoise-quality`)."""
    from vulkanpbrt_amd import denoiser_train as dt
    kinds = [6, 2] + (["c"] if dt.COMPACT_WEIGHTS.exists() else [])
    for k in kinds:
        fr.ensure_cnn(k)
    out = {}
    for name, fn in [("bmfr", fr.bmfr)] + [(f"cnn{k}", lambda f, k=k: fr.cnn(f, k)) for k in kinds]:
        fn(frame)  # warm-up
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        for _ in range(runs):
            fn(frame)
        b.record()
        torch.cuda.synchronize()
        out[f"{name}_ms"] = round(a.elapsed_time(b) / runs, 4)
    for k in kinds:
        out[f"cnn{k}_speedup_vs_bmfr"] = round(out["bmfr_ms"] / out[f"cnn{k}_ms"], 3)
    if evaluate:
        q = dt.sequence_quality(torch, fr, 8, (0, 7), 1024, layers=tuple(kinds))
        names = ["accumulated", "bmfr"] + [f"cnn{k}" for k in kinds]
        out["relmse"] = {k: round(v, 5) for k, v in q["mean"].items()
                         if k in names or k in [n + "_demodulated" for n in names[1:]]}
        for k in kinds:
            out[f"cnn{k}_relmse_vs_bmfr"] = round(q["mean"][f"cnn{k}"] / q["mean"]["bmfr"], 3)
            out[f"cnn{k}_at_or_below_bmfr_relmse"] = bool(q["mean"][f"cnn{k}"] <= q["mean"]["bmfr"])
        out["eval_set"] = ("the timed workload's first 8 camera-path frames, frames 0 and 7 scored against 1024-spp "
                           "references; relMSE of the remodulated radiance (and of the demodulated radiance), "
                           "mean of the two frames")
    return out


def cpu_baseline(gen, desc, cfg, budget_s, frames):
    """The oracle's FP32 restatement on the host cores, over bands of 1/8 of the image
    rows (from the middle, then the rest) of the timed camera frames, until the time
    budget is spent (at least one band)."""
    from oracle import oracle
    from vulkanpbrt_amd import camera
    osc = oracle.OracleScene(desc)
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
                       f"{min(done)}..{max(done)} of the timed run, FP32 oracle restatement (the device's FP32 ray/triangle "
                       f"test over the oracle's median-split BVH), "
                       f"{total_rays / 1e6:.1f} M rays in {total_t:.1f} s")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported in this build "
              "(multi-GPU sharding is design-only).", file=sys.stderr)
        sys.exit(2)
    import torch
    from vulkanpbrt_amd import camera, capi
    from vulkanpbrt_amd.frame import Frame
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
    bvh = args.bvh or cfg.get("bvh", "sbvh")
    builder = {"lbvh": capi.PBRT_BVH_LBVH_GPU, "sbvh": capi.PBRT_BVH_SBVH_HOST,
               "ploc": capi.PBRT_BVH_PLOC_GPU}.get(bvh, capi.PBRT_BVH_SAH_HOST)
    gen.desc.bvh_builder = builder
    desc, loaded, load_info = gen.desc, None, None
    if args.scene:
        t_load = time.perf_counter()
        loaded = capi.LoadedScene(args.scene, bvh_builder=builder)
        st = loaded.stats()
        desc = loaded.desc
        # a file carries no light sources but its emissive meshes (the reference does not
        # visit vsg::Light nodes): the workload's explicit lights (the sun) are added
        desc.light_count, desc.lights = gen.desc.light_count, gen.desc.lights
        load_info = {"file": str(args.scene), "meshes": st.mesh_count, "triangles": st.triangle_count,
                     "textures": st.texture_count, "emissive_meshes": st.emissive_mesh_count,
                     "file_mb": round(st.file_bytes / 1e6, 1), "parse_ms": round(st.parse_ms, 1),
                     "load_s": round(time.perf_counter() - t_load, 2)}
    scene = capi.Scene(desc, keepalive=(loaded, gen))
    info = scene.info()
    fr = Frame(torch, capi, camera, gen, scene, w, h, cfg["depth"], args.tracer, args.bmfr_fit, args.denoiser)
    setup_s = time.perf_counter() - t_setup

    tuning = fr.autotune() if args.autotune else None
    if args.denoise_quality:
        from vulkanpbrt_amd import denoiser_train as dt
        score = sorted({int(s) for s in args.quality_score.split(",")})
        t0 = time.perf_counter()
        kinds = (6, 2) + (("c",) if dt.COMPACT_WEIGHTS.exists() else ())
        q = dt.sequence_quality(torch, fr, args.quality_frames, score, args.quality_spp, layers=kinds,
                                progress=True)
        print(json.dumps({"denoise_quality": {
            "workload": cfg["name"], "frames": args.quality_frames, "scored_frames": score,
            "reference_spp": args.quality_spp, "metric": "relMSE of the remodulated radiance "
            "(albedo x demodulated), mean over pixels and rgb of (x - ref)^2 / (ref^2 + 0.01); "
            "*_demodulated: of the demodulated radiance",
            "networks": "6-layer, fused 2-layer and compact (3x3 then pointwise) kernel-predicting networks "
                        "(cnn6, cnn2, cnnc), committed parameters trained on Cornell-box and atrium scenes of "
                        "other seeds", "seconds": round(time.perf_counter() - t0, 1),
            "mean": {n: round(v, 5) for n, v in q["mean"].items()},
            "per_frame": {str(f): {n: round(v, 5) for n, v in row.items()} for f, row in q["frames"].items()}}}))
        return
    use_graph = (cfg.get("graph", False) and not args.pipeline) if args.graph is None else args.graph
    graph = None
    if use_graph:
        graph = C.c_void_p()
        capi.check(capi.lib().pbrt_frame_graph_create(C.byref(graph)))
        # a capture needs a stream other than the null stream
        side = torch.cuda.Stream()
        torch.cuda.set_stream(side)
    for f in range(args.warmup):
        fr.run(f, graph=graph)
    torch.cuda.synchronize()

    K = args.steps
    lib = capi.lib()

    def eager_pass(first):
        """K frames launched one by one with per-stage events and the library's
        per-kernel tracer timing (HIP events on the launch stream)."""
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(K)]
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        capi.check(lib.pbrt_trace_timing_begin(scene.handle))
        start.record()
        for k in range(K):
            fr.run(first + k, ev[k])
        end.record()
        torch.cuda.synchronize()
        tm = capi.TraceTiming()
        capi.check(lib.pbrt_trace_timing_end(scene.handle, C.byref(tm)))
        st = np.array([[ev[k][i].elapsed_time(ev[k][i + 1]) for i in range(4)] for k in range(K)])
        return start.elapsed_time(end), st.mean(axis=0), tm

    graph_info = pipe_info = None
    if use_graph:
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record()
        for k in range(K):
            fr.run(args.warmup + k, graph=graph)
        end.record()
        torch.cuda.synchronize()
        total_ms = start.elapsed_time(end)
        inst_timed = fr.graph_instantiations
        # the same frames again, eagerly, for the per-stage and per-kernel breakdown
        # (the accumulator history differs; the traced rays do not depend on it)
        eager_ms, stage_ms, timing = eager_pass(args.warmup)
        graph_info = {"frames_per_graph": 1, "kernel_nodes": fr.graph_nodes,
                      "instantiations": inst_timed, "eager_ms_per_step": round(eager_ms / K, 3)}
    elif args.pipeline:
        post = torch.cuda.Stream()
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record()
        fr.run_pipelined(args.warmup, K, post)
        end.record()
        torch.cuda.synchronize()
        total_ms = start.elapsed_time(end)
        eager_ms, stage_ms, timing = eager_pass(args.warmup)
        pipe_info = {"streams": 2, "one_stream_ms_per_step": round(eager_ms / K, 3)}
    else:
        total_ms, stage_ms, timing = eager_pass(args.warmup)
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
    # ray + 32 B per (quantized) BVH node visit + 48 B per triangle record tested
    # (+ 36 B/px of G-buffer and illumination output for the megakernel, which also shades).
    ext_bytes = 48.0 * cnt[0] + NODE_BYTES * cnt[2] + 48.0 * cnt[3]
    sh_bytes = 48.0 * cnt[1] + NODE_BYTES * cnt[4] + 48.0 * cnt[5]
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
        dom = "pt_persistent" if fr.tracer == capi.TRACER_PERSISTENT else "pt_megakernel"
        dom_bytes, dom_ms = ext_bytes + sh_bytes + 36.0 * w * h * K, timing.megakernel_ms
        kernels[dom] = dict(ms_per_frame=round(dom_ms / K, 3), launches=int(timing.megakernel_launches))
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    workload = f"{cfg['tag']}-{bvh}"
    traffic, traffic_src = measured_traffic(dom, workload)
    valu_busy, valu_src = profile_value("valu_busy_frac", dom, workload)
    lanes, _ = profile_value("valu_lane_util", dom, workload)
    n_launch = timing.extend_launches if fr.tracer == capi.TRACER_WAVEFRONT else timing.megakernel_launches
    avg_us = dom_ms / max(1, n_launch) * 1e3
    # `achieved` / `frac`: algorithmic bytes (ray I/O, node and triangle fetches, output) over
    # the HBM peak.  Most node and triangle fetches hit L1/L2/MALL, so this is gather
    # throughput through the cache hierarchy; what reaches HBM is `traffic` (rocprofv3
    # FETCH_SIZE + WRITE_SIZE per launch, from the named profile), and what binds the walk
    # is VALU issue under divergence (`valu_busy_frac`: SIMD cycles issuing VALU over the
    # kernel's cycles; `valu_lane_util`: live lanes per VALU instruction).
    roofline = dict(bound="valu_issue", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), frac_of="algorithmic bytes / HBM peak",
                    traffic=traffic, kernel=dom,
                    algorithmic_bytes_per_launch=round(dom_bytes / max(1, n_launch)),
                    avg_launch_us=round(avg_us, 1),
                    hbm_gbs=round(traffic / avg_us / 1e3, 1) if traffic else None,
                    hbm_frac=round(traffic / avg_us / 1e3 / HBM_PEAK_GBS, 4) if traffic else None,
                    valu_busy_frac=valu_busy, valu_lane_util=lanes,
                    traffic_source=f"profiles/{traffic_src}" if traffic_src else None,
                    valu_source=f"profiles/{valu_src}" if valu_src else None)

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
        "data": ("synthetic (seed-generated scene, random-free camera spline)" if not args.scene else
                 "scene file read by pbrt_load_obj, the workload's camera spline"),
        "config": {"workload": cfg["name"], "width": w, "height": h, "spp": 1, "max_depth": cfg["depth"],
                   "triangles": info.triangle_count, "lights": info.light_count, "tracer": args.tracer,
                   "denoiser": args.denoiser, "bmfr_fit": args.bmfr_fit, "seed": cfg["seed"],
                   "launch": "hipgraph per frame" if use_graph else
                             ("eager, trace and post chain on two streams" if args.pipeline else "eager")},
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
    if load_info:
        out["config"]["workload"] += " (geometry and materials from the scene file)"
        out["scene_file"] = load_info
    if graph_info:
        out["hipgraph"] = graph_info
    if pipe_info:
        out["pipeline"] = pipe_info
    if tuning:
        out["autotune"] = tuning
        out["config"]["tracer"] = tuning["best"]["tracer"]
        if "bmfr" in tuning:
            out["config"]["bmfr_fit_threads"] = tuning["bmfr"]["best_fit_threads"]
    out["denoise_comparison"] = denoise_comparison(torch, fr, args.warmup + K - 1, max(K, 10),
                                                   not args.no_denoise_eval)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(gen, desc, cfg, args.cpu_seconds, range(args.warmup, args.warmup + K))
    print(json.dumps(out))
    if graph is not None:
        capi.check(lib.pbrt_frame_graph_destroy(graph))


if __name__ == "__main__":
    main()
