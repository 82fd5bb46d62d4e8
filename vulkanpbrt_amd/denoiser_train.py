"""Training pairs from the path tracer, Adam training of the learned denoiser, and the
relMSE comparison with BMFR, all through the C ABI of libpbrt.so (torch only holds
device memory).

Pairs: for every seed, the Cornell-box scene of that seed (config 1) is traced at
1 sample per pixel with its G-buffer and passed through the accumulator's first
frame (the input BMFR and the network read), and traced again `target_spp` times
with other frame numbers (independent random streams); the running mean of those
demodulated frames is the reference.

    python -m vulkanpbrt_amd.denoiser_train --train-scenes 192 --eval-scenes 16 \\
        --size 128 --target-spp 256 --steps 3000 --out weights.npy
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
DEFAULT_WEIGHTS = WEIGHTS_DIR / "kpn_cornell.bin"        # 6 layers (the learned-denoiser configuration)
TWO_LAYER_WEIGHTS = WEIGHTS_DIR / "kpn2_cornell.bin"     # 2 layers, fused (equal quality to BMFR, faster)


class Network:
    """A pbrt_cnn handle (2..6 layers) with host parameter I/O (float32, canonical order)."""

    def __init__(self, seed: int = 1, layers: int = 6):
        self.handle = C.c_void_p()
        self.layers = layers
        capi.check(capi.lib().pbrt_cnn_create_layers(layers, seed, C.byref(self.handle)))

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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-scenes", type=int, default=192)
    ap.add_argument("--eval-scenes", type=int, default=16)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--target-spp", type=int, default=256)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--layers", type=int, default=6, help="convolutions (2..6)")
    ap.add_argument("--out", type=Path, required=True, help="trained parameters (.bin raw float32 or .npy)")
    args = ap.parse_args()
    import torch
    t0 = time.perf_counter()
    train_seeds = np.arange(10_000, 10_000 + args.train_scenes)
    eval_seeds = np.arange(90_000, 90_000 + args.eval_scenes)
    print(f"rendering {args.train_scenes} training and {args.eval_scenes} evaluation pairs "
          f"({args.size}x{args.size}, {args.target_spp} spp references)", flush=True)
    tr = render_pairs(torch, train_seeds, args.size, args.size, args.target_spp, progress=32)
    ev = render_pairs(torch, eval_seeds, args.size, args.size, args.target_spp)
    t_render = time.perf_counter() - t0
    net = Network(args.seed, args.layers)
    before = evaluate(torch, net, ev)
    print("before training", before, flush=True)
    history, t_train = train(torch, net, tr, args.steps, args.batch, args.lr, seed=args.seed)
    after = evaluate(torch, net, ev)
    print("after training", after, flush=True)
    net.save(args.out)
    report = dict(train_scenes=args.train_scenes, eval_scenes=args.eval_scenes, size=args.size,
                  target_spp=args.target_spp, steps=args.steps, batch=args.batch, lr=args.lr, seed=args.seed,
                  layers=args.layers,
                  render_s=round(t_render, 1), train_s=round(t_train, 1), loss_history=history,
                  eval_untrained=before, eval_trained=after)
    args.out.with_suffix(".json").write_text(json.dumps(report, indent=1))
    net.close()


if __name__ == "__main__":
    main()
