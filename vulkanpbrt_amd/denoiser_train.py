"""Training pairs from the path tracer, Adam training of the learned denoiser, and the
relMSE comparison with BMFR, all through the C ABI of libpbrt.so (torch only holds
device memory).

Pairs: for every seed, the Cornell-box scene of that seed (config 1) is traced at
1 sample per pixel with its G-buffer and passed through the accumulator's first
frame (the input BMFR and the network read), and traced again `target_spp` times
with other frame numbers (independent random streams); the running mean of those
demodulated frames is the reference.  Atrium pairs (render_sequence_pairs) are crops
of the frame pipeline's own denoiser input on atrium scenes of other seeds than the
timed one: a few camera-path frames traced and accumulated with reprojection, and a
reference at the last camera.

    python -m vulkanpbrt_amd.denoiser_train --train-scenes 256 --eval-scenes 16 \\
        --size 128 --target-spp 512 --atrium-seeds 16 --atrium-spp 1024 \\
        --steps 30000 --lr 1e-3 --layers 6 2 --out kpn6.bin kpn2.bin
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import time
from pathlib import Path

import numpy as np

from . import camera, capi

REL_EPS = 0.01


def relmse(x, ref, albedo=None):
    """mean over pixels and rgb of (x - ref)^2 / (ref^2 + 0.01); with the G-buffer
    albedo (uint8), of the remodulated radiance (albedo + 1e-6) * x, as the final image
    is formed (bmfrPost.comp)."""
    x = np.asarray(x, np.float64)[..., :3]
    r = np.asarray(ref, np.float64)[..., :3]
    if albedo is not None:
        a = np.asarray(albedo)[..., :3].astype(np.float64) / 255.0 + 1e-6
        x, r = a * x, a * r
    return float(np.mean((x - r) ** 2 / (r ** 2 + REL_EPS)))


WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"
DEFAULT_WEIGHTS = WEIGHTS_DIR / "kpn6.bin"  # 6 layers (the learned-denoiser configuration)
TWO_LAYER_WEIGHTS = WEIGHTS_DIR / "kpn2.bin"  # 2 layers, fused
COMPACT_WEIGHTS = WEIGHTS_DIR / "kpnc.bin"  # 2 layers, the second pointwise, fused


def weights_for(kind):
    """Committed parameters of a network kind: 6, 2 or "c" (compact)."""
    return {6: DEFAULT_WEIGHTS, 2: TWO_LAYER_WEIGHTS, "c": COMPACT_WEIGHTS}[kind]


class Network:
    """A pbrt_cnn handle (2..6 layers, or "c": the compact two-layer network whose second
    convolution is pointwise) with host parameter I/O (float32, canonical order)."""

    def __init__(self, seed: int = 1, layers=6):
        self.handle = C.c_void_p()
        self.kind = layers
        if layers == "c":
            self.layers = 2
            capi.check(capi.lib().pbrt_cnn_create_compact(seed, C.byref(self.handle)))
        else:
            self.layers = int(layers)
            capi.check(capi.lib().pbrt_cnn_create_layers(self.layers, seed, C.byref(self.handle)))

    def parameter_count(self) -> int:
        n = C.c_uint32()
        capi.check(capi.lib().pbrt_cnn_layers_parameter_count(self.layers, C.byref(n)))
        return n.value

    def get(self) -> np.ndarray:
        out = np.zeros(self.parameter_count(), np.float32)
        capi.check(capi.lib().pbrt_cnn_get_parameters(self.handle, out.ctypes.data))
        return out

    def set(self, params) -> None:
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.parameter_count():
            raise ValueError(f"expected {self.parameter_count()} parameters, got {p.size}")
        capi.check(capi.lib().pbrt_cnn_set_parameters(self.handle, p.ctypes.data))

    def save(self, path) -> None:
        """Raw little-endian float32 in canonical order (.bin), or .npy."""
        path = Path(path)
        if path.suffix == ".npy":
            np.save(path, self.get())
        else:
            self.get().astype("<f4").tofile(path)

    def load(self, path) -> None:
        path = Path(path)
        if path.suffix == ".npy":
            self.set(np.load(path, allow_pickle=False))
        else:
            self.set(np.fromfile(path, dtype="<f4"))

    def close(self) -> None:
        if self.handle:
            capi.lib().pbrt_cnn_destroy(self.handle)
            self.handle = C.c_void_p()


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class PairSet:
    """N images of w x h on the device: accumulator / G-buffer inputs and references."""

    KEYS = ("noisy", "depth", "normal", "albedo", "spp", "motion", "target")

    def __init__(self, torch, n, w, h):
        d = "cuda"
        self.n, self.w, self.h = n, w, h
        self.noisy = torch.zeros((n, h, w, 4), dtype=torch.int16, device=d)
        self.depth = torch.zeros((n, h, w), dtype=torch.float32, device=d)
        self.normal = torch.zeros((n, h, w, 2), dtype=torch.float32, device=d)
        self.albedo = torch.zeros((n, h, w, 4), dtype=torch.uint8, device=d)
        self.spp = torch.zeros((n, h, w), dtype=torch.uint8, device=d)
        self.motion = torch.zeros((n, h, w, 2), dtype=torch.int16, device=d)
        self.target = torch.zeros((n, h, w, 4), dtype=torch.float32, device=d)

    def subset(self, torch, idx):
        out = PairSet.__new__(PairSet)
        out.n, out.w, out.h = len(idx), self.w, self.h
        ix = torch.as_tensor(np.asarray(idx), device="cuda")
        for k in self.KEYS:
            setattr(out, k, getattr(self, k).index_select(0, ix).contiguous())
        return out

    def images(self, denoised=None, final=None, workspace=None):
        return capi.CnnImages(self.w, self.h, self.n, *(capi.ptr(getattr(self, k)) for k in
                                                          ("noisy", "depth", "normal", "albedo", "spp", "motion")),
                              capi.ptr(denoised), capi.ptr(final), capi.ptr(self.target), capi.ptr(workspace),
                              workspace.numel() if workspace is not None else 0)


def render_pairs(torch, seeds, w, h, target_spp, max_depth=2, progress=None) -> PairSet:
    lib = capi.lib()
    ps = PairSet(torch, len(seeds), w, h)
    st = _stream(torch)
    tws = C.c_size_t()
    capi.check(lib.pbrt_trace_workspace_size(w, h, C.byref(tws)))
    trace_ws = torch.zeros(tws.value, dtype=torch.uint8, device="cuda")
    img = dict(depth=torch.zeros((h, w), dtype=torch.float32, device="cuda"),
               normal=torch.zeros((h, w, 2), dtype=torch.float32, device="cuda"),
               material=torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda"),
               albedo=torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda"),
               illumination=torch.zeros((h, w, 4), dtype=torch.float32, device="cuda"))
    zeros = dict(depth=torch.zeros((h, w), dtype=torch.float32, device="cuda"),
                 spp=torch.zeros((h, w), dtype=torch.uint8, device="cuda"),
                 acc=torch.zeros((h, w, 4), dtype=torch.int16, device="cuda"))

    def trace(scene, pc, gbuffer):
        p = capi.TraceParams(w, h, max_depth, 0, capi.LIGHT_SAMPLE_AUTO, capi.OUTPUT_DEMOD_ILLUMINATION_FLOAT,
                             1 if gbuffer else 0, capi.TRACER_WAVEFRONT)
        ti = capi.TraceImages(*(capi.ptr(img[k]) for k in ("depth", "normal", "material", "albedo", "illumination")),
                              capi.ptr(trace_ws), tws.value)
        capi.check(lib.pbrt_trace(scene.handle, C.byref(p), C.byref(pc), C.byref(ti), None, st))

    for i, seed in enumerate(seeds):
        gen = capi.GeneratedScene(1, int(seed))
        scene = capi.Scene(gen.desc)
        view, proj = camera.pose_matrices(gen.camera(0), w, h)
        pc = camera.push_constants(view, proj, view, 7)
        trace(scene, pc, True)
        a = capi.AccumulateParams(w, h, 0, 0)
        a.inv_proj[:] = camera.colmajor(np.linalg.inv(proj)).tolist()
        a.inv_view[:] = camera.colmajor(np.linalg.inv(view)).tolist()
        a.prev_view[:] = camera.colmajor(view).tolist()
        o = np.linalg.inv(view)[:, 3].astype(np.float32)
        o[3] = 1
        a.prev_origin[:] = o.tolist()
        ai = capi.AccumulateImages(capi.ptr(img["illumination"]), capi.ptr(img["depth"]), capi.ptr(zeros["depth"]),
                                   capi.ptr(zeros["spp"]), capi.ptr(zeros["acc"]), capi.ptr(ps.motion[i]),
                                   capi.ptr(ps.spp[i]), capi.ptr(ps.noisy[i]))
        capi.check(lib.pbrt_accumulate(C.byref(a), C.byref(ai), st))
        ps.depth[i].copy_(img["depth"])
        ps.normal[i].copy_(img["normal"])
        ps.albedo[i].copy_(img["albedo"])
        for k in range(target_spp):
            pc.frame_number = 1000 + k
            trace(scene, pc, False)
            capi.check(lib.pbrt_running_mean(w, h, capi.ptr(img["illumination"]), capi.ptr(ps.target[i]), k, st))
        torch.cuda.synchronize()
        scene.close()
        if progress and (i + 1) % progress == 0:
            print(f"  rendered {i + 1}/{len(seeds)} pairs", flush=True)
    return ps


def concat_pairs(torch, sets):
    """One PairSet holding the images of `sets` (all of one size)."""
    out = PairSet.__new__(PairSet)
    out.w, out.h = sets[0].w, sets[0].h
    out.n = sum(s.n for s in sets)
    for k in PairSet.KEYS:
        setattr(out, k, torch.cat([getattr(s, k) for s in sets], 0).contiguous())
    return out


def render_sequence_pairs(torch, config, seeds, poses, w, h, crop, crops, ref_spp, max_history=8, seed=0,
                          progress=False) -> PairSet:
    """Crops of the frame pipeline's own denoiser input on generated scenes of `config`:
    for each scene seed and each of `poses` poses, a random start frame of the scene's
    camera path and a random history length n in 1..max_history are drawn, the n
    frames are traced at 1 spp (w x h, the workload's bounce count) and accumulated with
    reprojection as the frame pipeline runs them, and the last frame's accumulated
    illumination, G-buffer, sample counts and motion are the input; the running mean of
    `ref_spp` independent 1-spp frames at the last camera is the target.  `crops`
    random crop x crop windows of every such frame make the pairs."""
    from .frame import Frame
    rng = np.random.default_rng(seed)
    depth = {1: 2, 2: 4, 5: 8}.get(config, 4)
    out = PairSet(torch, len(seeds) * poses * crops, crop, crop)
    ref = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
    i = 0
    for s in seeds:
        gen = capi.GeneratedScene(config, int(s))
        gen.desc.bvh_builder = capi.PBRT_BVH_SAH_HOST
        scene = capi.Scene(gen.desc, keepalive=(gen,))
        fr = Frame(torch, capi, camera, gen, scene, w, h, depth, "megakernel", "gram")
        for p in range(poses):
            start, n = int(rng.integers(0, 64)), int(rng.integers(1, max_history + 1))
            fr.reset_history()
            for f in range(start, start + n):
                pc, view, proj = fr.push_constants(f)
                fr.trace(pc)
                fr.accumulate(f - start, view, proj)
                fr.prev_view = view
            last = start + n - 1
            cur = (n - 1) & 1
            src = dict(noisy=fr.acc[cur], depth=fr.img["depth"], normal=fr.img["normal"],
                       albedo=fr.img["albedo"], spp=fr.spp[cur], motion=fr.motion)
            windows = [(int(rng.integers(0, h - crop + 1)), int(rng.integers(0, w - crop + 1)))
                       for _ in range(crops)]
            for c, (y, x) in enumerate(windows):
                for k, t in src.items():
                    getattr(out, k)[i + c].copy_(t[y:y + crop, x:x + crop])
            render_reference(torch, fr, last, ref_spp, ref)
            for c, (y, x) in enumerate(windows):
                out.target[i + c].copy_(ref[y:y + crop, x:x + crop])
            torch.cuda.synchronize()
            i += crops
            if progress:
                print(f"  scene {s} pose {p}: frames {start}..{last} ({n} accumulated), {i}/{out.n} crops",
                      flush=True)
        for net in fr.nets.values():
            net.close()
        scene.close()
    return out


def train(torch, net: Network, data: PairSet, steps, batch, lr, seed=0, log_every=250):
    lib = capi.lib()
    rng = np.random.default_rng(seed)
    ws_bytes = C.c_size_t()
    capi.check(lib.pbrt_cnn_workspace_size(data.w, data.h, batch, 1, C.byref(ws_bytes)))
    ws = torch.zeros(ws_bytes.value, dtype=torch.uint8, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    history = []
    t0 = time.perf_counter()
    for step in range(steps):
        # cosine decay of the learning rate
        lr_t = lr * 0.5 * (1.0 + np.cos(np.pi * step / steps))
        sub = data.subset(torch, rng.choice(data.n, batch, replace=False))
        im = sub.images(workspace=ws)
        capi.check(lib.pbrt_cnn_train_step(net.handle, C.byref(im), C.c_float(max(lr_t, lr * 1e-3)),
                                           capi.ptr(loss), _stream(torch)))
        if step % log_every == 0 or step == steps - 1:
            torch.cuda.synchronize()
            history.append((step, float(loss.item())))
            print(f"  step {step:5d} loss {loss.item():.5f} ({time.perf_counter() - t0:.1f} s)", flush=True)
    torch.cuda.synchronize()
    return history, time.perf_counter() - t0


def denoise_cnn(torch, net: Network, data: PairSet):
    lib = capi.lib()
    ws_bytes = C.c_size_t()
    capi.check(lib.pbrt_cnn_workspace_size(data.w, data.h, data.n, 0, C.byref(ws_bytes)))
    ws = torch.zeros(ws_bytes.value, dtype=torch.uint8, device="cuda")
    out = torch.zeros((data.n, data.h, data.w, 4), dtype=torch.float32, device="cuda")
    im = data.images(denoised=out, workspace=ws)
    capi.check(lib.pbrt_cnn_denoise(net.handle, C.byref(im), _stream(torch)))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def denoise_bmfr(torch, data: PairSet, block=32):
    """BMFR (Gram FP64 fit) of every image as the first frame; returns its demodulated output."""
    lib = capi.lib()
    out = []
    p = capi.BmfrParams(data.w, data.h, block, 0, capi.BMFR_FIT_GRAM_F64)
    n = C.c_size_t()
    capi.check(lib.pbrt_bmfr_workspace_size(C.byref(p), C.byref(n)))
    ws = torch.zeros(n.value, dtype=torch.uint8, device="cuda")
    den = torch.zeros((2, data.h, data.w, 4), dtype=torch.int16, device="cuda")
    final = torch.zeros((data.h, data.w, 4), dtype=torch.uint8, device="cuda")
    for i in range(data.n):
        im = capi.BmfrImages(capi.ptr(data.depth[i]), capi.ptr(data.normal[i]), capi.ptr(data.albedo[i]),
                             capi.ptr(data.motion[i]), capi.ptr(data.spp[i]), capi.ptr(data.noisy[i]),
                             capi.ptr(den), capi.ptr(final), capi.ptr(ws), n.value)
        capi.check(lib.pbrt_bmfr(C.byref(p), C.byref(im), _stream(torch)))
        torch.cuda.synchronize()
        out.append(den[1].cpu().numpy().view(np.float16).astype(np.float32))
    return np.stack(out)


def evaluate(torch, net: Network, data: PairSet):
    ref = data.target.cpu().numpy()
    noisy = data.noisy.cpu().numpy().view(np.float16).astype(np.float32)
    cnn_out = denoise_cnn(torch, net, data)
    bmfr_out = denoise_bmfr(torch, data)
    alb = data.albedo.cpu().numpy()
    out = {}
    for name, img in (("noisy", noisy), ("bmfr", bmfr_out), ("cnn", cnn_out)):
        out[f"relmse_{name}"] = relmse(img, ref, alb)
        out[f"relmse_{name}_demodulated"] = relmse(img, ref)
    return out


def box5(img):
    """The 5x5 box filter of an (n, h, w, c) stack with taps clamped to the image: the
    kernel-prediction network's output when every logit is equal (a network without
    learned kernels)."""
    x = np.asarray(img, np.float64)
    n, h, w = x.shape[:3]
    ys = np.clip(np.arange(h)[:, None] + np.arange(-2, 3)[None, :], 0, h - 1)
    xs = np.clip(np.arange(w)[:, None] + np.arange(-2, 3)[None, :], 0, w - 1)
    rows = x[:, ys].mean(axis=2)             # (n, h, w, c)
    return rows[:, :, xs].mean(axis=3)


def guided5(img, depth, normal, albedo, sigma_albedo, sigma_normal, sigma_depth):
    """A 5x5 cross-bilateral filter of an (n, h, w, c) stack guided by the G-buffer, taps
    clamped to the image: tap q of pixel p weighs
    exp(-|a_q - a_p|^2 / (2 sa^2) - (1 - n_q . n_p) / sn - ((z_q - z_p) / (sz z_p))^2 / 2)
    (albedo a in [0, 1], unit normal n from the G-buffer's (theta, phi), depth z) - the
    hand-designed feature-guided kernel of the network's footprint."""
    x = np.asarray(img, np.float64)[..., :3]
    n, h, w = x.shape[:3]
    a = np.asarray(albedo)[..., :3].astype(np.float64) / 255.0
    th, ph = np.asarray(normal, np.float64)[..., 0], np.asarray(normal, np.float64)[..., 1]
    nv = np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)], axis=-1)
    z = np.asarray(depth, np.float64)
    z = np.where(np.isfinite(z), z, 1e30)
    zs = np.maximum(np.abs(z), 1e-3)
    acc = np.zeros_like(x)
    wsum = np.zeros(x.shape[:3])
    yy, xx = np.arange(h), np.arange(w)
    for dy in range(-2, 3):
        ys = np.clip(yy + dy, 0, h - 1)
        for dx in range(-2, 3):
            xs = np.clip(xx + dx, 0, w - 1)
            sel = (slice(None), ys[:, None], xs[None, :])
            wt = np.exp(-((a[sel] - a) ** 2).sum(-1) / (2 * sigma_albedo ** 2)
                        - (1.0 - (nv[sel] * nv).sum(-1)) / sigma_normal
                        - ((z[sel] - z) / (sigma_depth * zs)) ** 2 / 2)
            acc += wt[..., None] * x[sel]
            wsum += wt
    return acc / wsum[..., None]


def render_reference(torch, fr, frame, spp, out, first_number=1_000_000):
    """Running mean into `out` (h x w x 4 float32) of `spp` 1-spp demodulated frames at
    camera pose `frame` of the sequence `fr` (a vulkanpbrt_amd.frame.Frame), each with
    its own frame number (independent random streams).  Overwrites fr's trace outputs."""
    lib = capi.lib()
    pc, _, _ = fr.push_constants(frame)
    p = fr.trace_params()
    p.write_gbuffer = 0
    ti = fr.trace_images()
    st = fr.stream()
    for k in range(spp):
        pc.frame_number = first_number + k
        capi.check(lib.pbrt_trace(fr.scene.handle, C.byref(p), C.byref(pc), C.byref(ti), None, st))
        capi.check(lib.pbrt_running_mean(fr.w, fr.h, capi.ptr(fr.img["illumination"]), capi.ptr(out), k, st))
        if k % 256 == 255:
            torch.cuda.synchronize()


def sequence_quality(torch, fr, frames, score_frames, ref_spp, layers=(6, 2), progress=False):
    """relMSE of the frame pipeline's denoisers on a camera sequence, against `ref_spp`
    references.  Frames 0 .. frames-1 of `fr`'s camera path are traced at 1 spp and
    accumulated (motion-vector reprojection), then denoised by BMFR with its temporal
    history (bmfrPost's blend, as the frame pipeline runs it) and by each learned
    network in `layers` on the same accumulated input.  Beside that pipeline every
    frame's own 1-spp trace is denoised alone, as a first frame (no history: the
    `*_single` outputs), so the effect of the temporal reuse shows on its own.  At every
    frame in `score_frames` each output, the accumulated input and the 1-spp trace
    ("traced") are scored by relMSE of the remodulated radiance (and of the demodulated
    one) against the running mean of `ref_spp` independent 1-spp frames at that camera.
    Returns a dict per score frame plus their means."""
    lib = capi.lib()
    h, w = fr.h, fr.w
    d = "cuda"
    fr.reset_history()
    for n in layers:
        fr.ensure_cnn(n)
    cnn_out = {n: torch.zeros((h, w, 4), dtype=torch.float32, device=d) for n in layers}
    ref = torch.zeros((h, w, 4), dtype=torch.float32, device=d)
    # single-frame inputs: the accumulator's first frame of the current trace
    zero_depth = torch.zeros((h, w), dtype=torch.float32, device=d)
    zero_spp = torch.zeros((h, w), dtype=torch.uint8, device=d)
    zero_acc = torch.zeros((h, w, 4), dtype=torch.int16, device=d)
    one = dict(motion=torch.zeros((h, w, 2), dtype=torch.int16, device=d),
               spp=torch.zeros((h, w), dtype=torch.uint8, device=d),
               acc=torch.zeros((h, w, 4), dtype=torch.int16, device=d),
               den=torch.zeros((2, h, w, 4), dtype=torch.int16, device=d),
               final=torch.zeros((h, w, 4), dtype=torch.uint8, device=d))
    bws = torch.zeros(fr.ws_bytes, dtype=torch.uint8, device=d)
    single_out = {n: torch.zeros((h, w, 4), dtype=torch.float32, device=d) for n in layers}
    rows = {}
    for f in range(frames):
        pc, view, proj = fr.push_constants(f)
        fr.trace(pc)
        fr.accumulate(f, view, proj)
        fr.bmfr(f)
        for n in layers:
            fr.cnn(f, n, denoised=cnn_out[n])
        fr.prev_view = view
        if f not in score_frames:
            continue
        st = fr.stream()
        a = capi.AccumulateParams(w, h, 0, 0)
        a.inv_proj[:] = camera.colmajor(np.linalg.inv(proj)).tolist()
        a.inv_view[:] = camera.colmajor(np.linalg.inv(view)).tolist()
        a.prev_view[:] = camera.colmajor(view).tolist()
        o = np.linalg.inv(view)[:, 3].astype(np.float32)
        o[3] = 1
        a.prev_origin[:] = o.tolist()
        ai = capi.AccumulateImages(capi.ptr(fr.img["illumination"]), capi.ptr(fr.img["depth"]),
                                   capi.ptr(zero_depth), capi.ptr(zero_spp), capi.ptr(zero_acc),
                                   capi.ptr(one["motion"]), capi.ptr(one["spp"]), capi.ptr(one["acc"]))
        capi.check(lib.pbrt_accumulate(C.byref(a), C.byref(ai), st))
        bp = capi.BmfrParams(w, h, 32, 0, fr.fit, fr.bmfr_threads)
        bi = capi.BmfrImages(capi.ptr(fr.img["depth"]), capi.ptr(fr.img["normal"]), capi.ptr(fr.img["albedo"]),
                             capi.ptr(one["motion"]), capi.ptr(one["spp"]), capi.ptr(one["acc"]),
                             capi.ptr(one["den"]), capi.ptr(one["final"]), capi.ptr(bws), fr.ws_bytes)
        capi.check(lib.pbrt_bmfr(C.byref(bp), C.byref(bi), st))
        for n in layers:
            ci = capi.CnnImages(w, h, 1, capi.ptr(one["acc"]), capi.ptr(fr.img["depth"]),
                                capi.ptr(fr.img["normal"]), capi.ptr(fr.img["albedo"]), capi.ptr(one["spp"]),
                                capi.ptr(one["motion"]), capi.ptr(single_out[n]), None, None,
                                capi.ptr(fr.cnn_ws), fr.cnn_ws.numel())
            capi.check(lib.pbrt_cnn_denoise(fr.nets[n].handle, C.byref(ci), st))
        torch.cuda.synchronize()
        half = lambda t: t.cpu().numpy().view(np.float16).astype(np.float32)
        outs = {"traced": fr.img["illumination"].cpu().numpy(), "accumulated": half(fr.acc[f & 1]),
                "bmfr": half(fr.denoised[(f & 1) ^ 1])}
        for n in layers:
            outs[f"cnn{n}"] = cnn_out[n].cpu().numpy()
        outs["bmfr_single"] = half(one["den"][1])
        for n in layers:
            outs[f"cnn{n}_single"] = single_out[n].cpu().numpy()
        alb = fr.img["albedo"].cpu().numpy()
        spp = fr.spp[f & 1].cpu().numpy()
        render_reference(torch, fr, f, ref_spp, ref)
        torch.cuda.synchronize()
        r = ref.cpu().numpy()
        row = {"mean_accumulated_spp": round(float(spp.mean()), 2),
               "reused_fraction": round(float((spp > 1).mean()), 4)}
        for name, img in outs.items():
            row[name] = relmse(img, r, alb)
            row[f"{name}_demodulated"] = relmse(img, r)
        rows[f] = row
        if progress:
            print(f"  frame {f}: " + " ".join(f"{k} {v:.4f}" for k, v in row.items()), flush=True)
        # the reference traces overwrote this frame's trace outputs; the next frame
        # retraces them, and the accumulator's depth history was saved before
    names = [k for k in next(iter(rows.values())) if k not in ("mean_accumulated_spp", "reused_fraction")]
    mean = {k: float(np.mean([rows[f][k] for f in rows])) for k in names}
    return {"frames": rows, "mean": mean}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-scenes", type=int, default=192, help="Cornell-box training scenes")
    ap.add_argument("--eval-scenes", type=int, default=16)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--target-spp", type=int, default=256, help="samples per pixel of the Cornell-box targets")
    ap.add_argument("--atrium-seeds", type=int, default=0,
                    help="atrium scenes (seeds 100, 101, ...; the timed workload is seed 2) whose camera-path "
                         "frames, accumulated as the frame pipeline does, add crops to the training set")
    ap.add_argument("--atrium-poses", type=int, default=4, help="poses per atrium scene")
    ap.add_argument("--atrium-crops", type=int, default=32, help="crops per atrium pose")
    ap.add_argument("--atrium-spp", type=int, default=1024, help="samples per pixel of the atrium targets")
    ap.add_argument("--atrium-width", type=int, default=1920)
    ap.add_argument("--atrium-height", type=int, default=1080)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--layers", nargs="+", default=["6"],
                    help="convolutions (2..6) or c (the compact network: 3x3 then pointwise), one network each")
    ap.add_argument("--out", type=Path, nargs="+", required=True,
                    help="trained parameters (.bin raw float32 or .npy), one per --layers entry")
    args = ap.parse_args()
    if len(args.out) != len(args.layers):
        ap.error("give one --out per --layers entry")
    import torch
    t0 = time.perf_counter()
    train_seeds = np.arange(10_000, 10_000 + args.train_scenes)
    eval_seeds = np.arange(90_000, 90_000 + args.eval_scenes)
    print(f"rendering {args.train_scenes} training and {args.eval_scenes} evaluation pairs "
          f"({args.size}x{args.size}, {args.target_spp} spp references)", flush=True)
    sets = [render_pairs(torch, train_seeds, args.size, args.size, args.target_spp, progress=32)]
    ev = render_pairs(torch, eval_seeds, args.size, args.size, args.target_spp)
    if args.atrium_seeds:
        print(f"rendering {args.atrium_seeds} atrium scenes x {args.atrium_poses} poses x {args.atrium_crops} "
              f"crops ({args.atrium_spp} spp references)", flush=True)
        sets.append(render_sequence_pairs(torch, 2, np.arange(100, 100 + args.atrium_seeds), args.atrium_poses,
                                          args.atrium_width, args.atrium_height, args.size, args.atrium_crops,
                                          args.atrium_spp, seed=args.seed, progress=True))
    tr = concat_pairs(torch, sets) if len(sets) > 1 else sets[0]
    t_render = time.perf_counter() - t0
    for layers, out in zip(args.layers, args.out):
        layers = layers if layers == "c" else int(layers)
        net = Network(args.seed, layers)
        before = evaluate(torch, net, ev)
        print(f"{layers} layers, before training", before, flush=True)
        history, t_train = train(torch, net, tr, args.steps, args.batch, args.lr, seed=args.seed)
        after = evaluate(torch, net, ev)
        print(f"{layers} layers, after training", after, flush=True)
        net.save(out)
        report = dict(train_scenes=args.train_scenes, eval_scenes=args.eval_scenes, size=args.size,
                      target_spp=args.target_spp, atrium_seeds=args.atrium_seeds, atrium_poses=args.atrium_poses,
                      atrium_crops=args.atrium_crops, atrium_spp=args.atrium_spp,
                      atrium_size=[args.atrium_width, args.atrium_height], training_pairs=tr.n,
                      steps=args.steps, batch=args.batch, lr=args.lr, seed=args.seed, layers=layers,
                      render_s=round(t_render, 1), train_s=round(t_train, 1), loss_history=history,
                      eval_untrained=before, eval_trained=after)
        out.with_suffix(".json").write_text(json.dumps(report, indent=1))
        net.close()


if __name__ == "__main__":
    main()
