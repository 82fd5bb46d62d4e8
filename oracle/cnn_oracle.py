"""CPU oracle of the learned kernel-predicting denoiser (test infrastructure only).

The reference has no learned denoiser; this network is the project's own stand-in
for BMFR's per-block fit (bmfrFit.comp), so the oracle restates the project's
definition, in numpy, independently of the HIP kernels:

  features (12 channels per pixel, from the accumulator / G-buffer images that BMFR
  also reads): log1p(max(noisy rgb, 0)), albedo rgb, decoded normal xyz,
  depth / (1 + depth), min(spp, 16) / 16, motion valid (1 if motion.x >= 0)
  6 layers of 3x3 convolution (cross-correlation, zero padding):
      12 -> 32 -> 32 -> 32 -> 32 -> 32 (ReLU after each) -> 25 logits
  softmax over the 25 logits = a 5x5 kernel (tap k = ky * 5 + kx, offset
  (kx - 2, ky - 2), taps clamped to the image edge) applied to the noisy
  demodulated illumination
  loss = mean over pixels and rgb of (a (out - target))^2 / ((a target)^2 + 0.01),
  a = albedo + 1e-6 (the remodulation of bmfrPost.comp): the relative squared error
  of the remodulated radiance

A network of L layers (2..6, shape_of(L)) is the same definition with L
convolutions, 12 -> 32 -> ... -> 32 -> 25; the two-layer one is the fused
equal-quality alternative to BMFR.

Parameters in canonical order, layer by layer: W[cout][cin][ky][kx], then b[cout].
`bf16=True` mirrors the device storage: weights, features, activations and
logits rounded to bfloat16 (round to nearest even); arithmetic in float64.
"""
from __future__ import annotations

import numpy as np

CIN = [12, 32, 32, 32, 32, 32]
COUT = [32, 32, 32, 32, 32, 25]
LAYERS = 6
EPS_REL = 0.01
DEFAULT = (CIN, COUT)


def shape_of(layers):
    """Layer shapes of a network of `layers` convolutions: 12 -> 32 -> ... -> 32 -> 25."""
    return ([12] + [32] * (layers - 1), [32] * (layers - 1) + [25])


def layer_shapes(shape=DEFAULT):
    cin, cout = shape
    return [((co, ci, 3, 3), (co,)) for ci, co in zip(cin, cout)]


def parameter_count(shape=DEFAULT):
    return sum(int(np.prod(w)) + b[0] for w, b in layer_shapes(shape))


def unpack(params, shape=DEFAULT):
    """Canonical flat vector -> list of (W, b)."""
    out, o = [], 0
    for ws, bs in layer_shapes(shape):
        n = int(np.prod(ws))
        out.append((params[o:o + n].reshape(ws), params[o + n:o + n + bs[0]]))
        o += n + bs[0]
    return out


def pack(layers):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in layers])


def bf16(x):
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def half_to_float(h):
    return np.asarray(h, np.uint16).view(np.float16).astype(np.float64)


def features(noisy, depth, normal, albedo, spp, motion):
    """noisy (N,H,W,4) RGBA16F bits, depth (N,H,W) f32, normal (N,H,W,2) f32 spherical,
    albedo (N,H,W,4) u8, spp (N,H,W) u8, motion (N,H,W,2) RG16F bits -> (N,H,W,12) float64."""
    nz = half_to_float(noisy)[..., :3]
    f = np.zeros(depth.shape + (12,))
    f[..., 0:3] = np.log1p(np.maximum(nz, 0.0))
    f[..., 3:6] = albedo[..., :3].astype(np.float64) / 255.0
    th, ph = normal[..., 0].astype(np.float64), normal[..., 1].astype(np.float64)
    f[..., 6] = np.cos(ph) * np.sin(th)
    f[..., 7] = np.sin(ph) * np.sin(th)
    f[..., 8] = np.cos(th)
    d = depth.astype(np.float64)
    f[..., 9] = d / (1.0 + d)
    cnt = spp.astype(np.float64) / 255.0 * 256.0
    f[..., 10] = np.minimum(cnt, 16.0) / 16.0
    f[..., 11] = (half_to_float(motion)[..., 0] >= 0).astype(np.float64)
    return f


def conv3x3(x, w, b):
    """x (N,H,W,Ci), w (Co,Ci,3,3), b (Co,) -> (N,H,W,Co); zero padding."""
    n, h, wd, ci = x.shape
    xp = np.zeros((n, h + 2, wd + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    out = np.zeros((n, h, wd, w.shape[0])) + b
    for ky in range(3):
        for kx in range(3):
            out += np.einsum("nhwc,oc->nhwo", xp[:, ky:ky + h, kx:kx + wd], w[:, :, ky, kx], optimize=True)
    return out


def conv3x3_backward(x, w, dy):
    """Gradients of conv3x3 wrt x, w, b for upstream dy (N,H,W,Co)."""
    n, h, wd, ci = x.shape
    xp = np.zeros((n, h + 2, wd + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for ky in range(3):
        for kx in range(3):
            win = xp[:, ky:ky + h, kx:kx + wd]
            dw[:, :, ky, kx] = np.einsum("nhwo,nhwc->oc", dy, win, optimize=True)
            dxp[:, ky:ky + h, kx:kx + wd] += np.einsum("nhwo,oc->nhwc", dy, w[:, :, ky, kx], optimize=True)
    return dxp[:, 1:-1, 1:-1], dw, dy.sum(axis=(0, 1, 2))


def gather_taps(img):
    """img (N,H,W,C) -> (N,H,W,25,C): 5x5 neighbourhood, clamped to the edge."""
    n, h, w, c = img.shape
    ys = np.arange(h)
    xs = np.arange(w)
    out = np.zeros((n, h, w, 25, c))
    for ky in range(5):
        for kx in range(5):
            yy = np.clip(ys + ky - 2, 0, h - 1)
            xx = np.clip(xs + kx - 2, 0, w - 1)
            out[:, :, :, ky * 5 + kx] = img[:, yy][:, :, xx]
    return out


def forward(params, feats, noisy_rgb, bf16_mirror=True, shape=DEFAULT):
    """Returns (output rgb (N,H,W,3), cache for backward)."""
    layers = unpack(np.asarray(params, np.float64), shape)
    rnd = bf16 if bf16_mirror else (lambda v: np.asarray(v, np.float64))
    a = rnd(feats)
    acts = [a]
    for li, (w, b) in enumerate(layers):
        z = conv3x3(a, rnd(w), np.asarray(b, np.float64))
        a = rnd(np.maximum(z, 0.0)) if li < len(layers) - 1 else rnd(z)
        acts.append(a)
    logits = acts[-1]
    m = logits.max(-1, keepdims=True)
    e = np.exp(logits - m)
    wk = e / e.sum(-1, keepdims=True)
    taps = gather_taps(noisy_rgb)
    out = np.einsum("nhwk,nhwkc->nhwc", wk, taps)
    return out, dict(acts=acts, wk=wk, taps=taps, layers=layers, rnd=rnd)


def remod(albedo_u8):
    return albedo_u8[..., :3].astype(np.float64) / 255.0 + 1e-6


def loss(out, target, albedo):
    """albedo: (N,H,W,>=3) uint8 G-buffer albedo."""
    a = remod(albedo)
    return float(np.mean((a * (out - target)) ** 2 / ((a * target) ** 2 + EPS_REL)))


def backward(cache, out, target, albedo):
    """Gradient of loss() wrt the canonical parameters."""
    acts, wk, taps, layers, rnd = cache["acts"], cache["wk"], cache["taps"], cache["layers"], cache["rnd"]
    a = remod(albedo)
    g = 2.0 * a * a * (out - target) / ((a * target) ** 2 + EPS_REL) / out.size
    dwk = np.einsum("nhwc,nhwkc->nhwk", g, taps)
    dz = wk * (dwk - (wk * dwk).sum(-1, keepdims=True))
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        w, _ = layers[li]
        dx, dw, db = conv3x3_backward(acts[li], rnd(w), dz)
        grads[li] = (dw, db)
        if li > 0:
            dz = dx * (acts[li] > 0)
    return pack(grads)


def init_parameters(seed, shape=DEFAULT):
    """He-normal weights (last layer scaled by 0.1), zero biases."""
    rng = np.random.default_rng(seed)
    layers = []
    shapes = layer_shapes(shape)
    for li, ((co, ci, _, _), _) in enumerate(shapes):
        std = np.sqrt(2.0 / (9 * ci)) * (0.1 if li == len(shapes) - 1 else 1.0)
        layers.append((rng.normal(0.0, std, (co, ci, 3, 3)), np.zeros(co)))
    return pack(layers)
