"""The learned denoiser's HIP kernels against the numpy oracle (oracle/cnn_oracle.py):
forward output, loss and the gradient of every parameter, plus a training run.

Tolerances: the device stores weights, activations and gradients in bfloat16 and
accumulates in fp32 on the matrix cores; the oracle mirrors the bf16 storage of the
forward pass (weights, features, activations, logits) in float64 and keeps the
backward pass in float64, so gradients differ by bf16 rounding of the backward
intermediates (relative ~1e-2 per element, cosine similarity > 0.999 per layer)."""
import ctypes as C

import numpy as np
import pytest

from oracle import cnn_oracle as cnn
from vulkanpbrt_amd import capi
from tests import helpers

pytestmark = pytest.mark.gpu

W, H, N = 37, 29, 2  # partial 16x16 tiles on both axes, two images


def half(a):
    return np.asarray(a, np.float32).astype(np.float16).view(np.uint16)


def synthetic_batch(seed):
    rng = np.random.default_rng(seed)
    noisy = np.ones((N, H, W, 4), np.float32)
    noisy[..., :3] = rng.gamma(1.0, 0.5, (N, H, W, 3))
    inp = dict(
        noisy=half(noisy),
        depth=rng.uniform(0.5, 20.0, (N, H, W)).astype(np.float32),
        normal=np.stack([rng.uniform(0, np.pi, (N, H, W)), rng.uniform(-np.pi, np.pi, (N, H, W))], -1).astype(
            np.float32),
        albedo=rng.integers(0, 256, (N, H, W, 4), dtype=np.uint8),
        spp=rng.integers(0, 40, (N, H, W), dtype=np.uint8),
        motion=half(rng.uniform(-1, 1, (N, H, W, 2))),
    )
    target = np.ones((N, H, W, 4), np.float32)
    target[..., :3] = rng.uniform(0.05, 1.5, (N, H, W, 3))
    inp["target"] = target
    return inp


class Net:
    def __init__(self, seed=1, layers=6):
        """layers: 2..6, or "c" for the compact network (two layers, the second pointwise)."""
        self.h = C.c_void_p()
        if layers == "c":
            self.layers = 2
            capi.check(capi.lib().pbrt_cnn_create_compact(seed, C.byref(self.h)))
        else:
            self.layers = layers
            capi.check(capi.lib().pbrt_cnn_create_layers(layers, seed, C.byref(self.h)))

    def set(self, params):
        p = np.ascontiguousarray(params, np.float32)
        capi.check(capi.lib().pbrt_cnn_set_parameters(self.h, p.ctypes.data))

    def get(self):
        n = C.c_uint32()
        capi.check(capi.lib().pbrt_cnn_layers_parameter_count(self.layers, C.byref(n)))
        out = np.zeros(n.value, np.float32)
        capi.check(capi.lib().pbrt_cnn_get_parameters(self.h, out.ctypes.data))
        return out

    def close(self):
        capi.lib().pbrt_cnn_destroy(self.h)


def device_images(torch, inp, training, w=W, h=H, n=N):
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in inp.items()}
    size = C.c_size_t()
    capi.check(capi.lib().pbrt_cnn_workspace_size(w, h, n, int(training), C.byref(size)))
    t["ws"] = torch.zeros(size.value, dtype=torch.uint8, device="cuda")
    t["denoised"] = torch.zeros((n, h, w, 4), dtype=torch.float32, device="cuda")
    t["final"] = torch.zeros((n, h, w, 4), dtype=torch.uint8, device="cuda")
    im = capi.CnnImages(w, h, n, *(capi.ptr(t[k]) for k in ("noisy", "depth", "normal", "albedo", "spp", "motion",
                                                             "denoised", "final")),
                        capi.ptr(t.get("target")), capi.ptr(t["ws"]), size.value)
    return im, t


def oracle_forward(params, inp, layers=6):
    feats = cnn.features(inp["noisy"], inp["depth"], inp["normal"], inp["albedo"], inp["spp"], inp["motion"])
    noisy = cnn.half_to_float(inp["noisy"])[..., :3]
    out, cache = cnn.forward(params, feats, noisy, bf16_mirror=True, shape=cnn.shape_of(layers))
    return out, cache


def outer_taps(layers):
    """Canonical indices of the last layer's eight outer taps (W[cout][cin][ky][kx])."""
    n = cnn.parameter_count(cnn.shape_of(layers))
    o = n - 25 - 9 * 32 * 25
    k = np.arange(9 * 32 * 25)
    return o + k[k % 9 != 4]


@pytest.mark.parametrize("layers", [6, 2, "c"])
def test_forward_matches_oracle(cuda, layers):
    """Six layers: the layer-by-layer path with kernel prediction fused into the last
    layer; two layers: the single fused kernel (features to filtered radiance in LDS);
    the compact network ("c": the second layer pointwise, its outer taps zero): its own
    fused kernel (both layers in registers per 16-pixel row).  The oracle runs the
    two-layer shape with the same parameters."""
    inp = synthetic_batch(3)
    net = Net(7, layers)
    nl = net.layers
    params = net.get()
    if layers == "c":
        assert not params[outer_taps(2)].any()
    # a non-trivial kernel: larger last-layer weights than the init
    o = cnn.parameter_count(cnn.shape_of(nl)) - 25 - 9 * 32 * 25
    params[o:o + 9 * 32 * 25] *= 10.0
    net.set(params)
    np.testing.assert_array_equal(net.get(), params)
    im, t = device_images(cuda, inp, training=False)
    capi.check(capi.lib().pbrt_cnn_denoise(net.h, C.byref(im), helpers.stream(cuda)))
    cuda.cuda.synchronize()
    g = t["denoised"].cpu().numpy()[..., :3]
    ref, _ = oracle_forward(params.astype(np.float64), inp, nl)
    rel = np.abs(g - ref) / np.maximum(np.abs(ref), 1e-3)
    print(f"cnn forward ({layers} layers): rel err median {np.median(rel):.2e} p99 {np.quantile(rel, 0.99):.2e} max {rel.max():.2e}")
    assert np.quantile(rel, 0.99) < 2e-2 and np.median(rel) < 2e-3
    # remodulated final image: BGRA8 of pow(albedo * out, 1/2.2)
    f = t["final"].cpu().numpy()
    alb = inp["albedo"][..., :3].astype(np.float64) / 255 + 1e-6
    expect = np.clip(np.maximum(alb * g, 0) ** 0.454545, 0, 1) * 255
    assert np.abs(f[..., 2::-1].astype(float) - expect).max() <= 1.0
    assert np.all(f[..., 3] == 255)
    net.close()


@pytest.mark.parametrize("layers", [6, 2, "c"])
def test_gradients_match_oracle(cuda, layers):
    """Every layer's gradient against the oracle's; the compact network reports zero
    gradient for its pointwise layer's outer taps and the oracle's for every other
    parameter."""
    inp = synthetic_batch(4)
    net = Net(9, layers)
    compact = layers == "c"
    layers = net.layers
    params = net.get().astype(np.float64)
    im, t = device_images(cuda, inp, training=True)
    n = cnn.parameter_count(cnn.shape_of(layers))
    grad = cuda.zeros(n, dtype=cuda.float32, device="cuda")
    loss = cuda.zeros(1, dtype=cuda.float32, device="cuda")
    capi.check(capi.lib().pbrt_cnn_gradients(net.h, C.byref(im), capi.ptr(loss), capi.ptr(grad),
                                             helpers.stream(cuda)))
    cuda.cuda.synchronize()
    g = grad.cpu().numpy().astype(np.float64)
    out, cache = oracle_forward(params, inp, layers)
    tgt = inp["target"][..., :3].astype(np.float64)
    ref_loss = cnn.loss(out, tgt, inp["albedo"])
    ref = cnn.backward(cache, out, tgt, inp["albedo"])
    if compact:
        outer = outer_taps(2)
        assert not g[outer].any()
        ref[outer] = 0.0
    print(f"cnn loss gpu {loss.item():.6f} oracle {ref_loss:.6f}")
    assert abs(loss.item() - ref_loss) <= 2e-3 * ref_loss
    o = 0
    for li, (ws, bs) in enumerate(cnn.layer_shapes(cnn.shape_of(layers))):
        sz = int(np.prod(ws)) + bs[0]
        a, b = g[o:o + sz], ref[o:o + sz]
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30)
        relnorm = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
        print(f"layer {li}: cosine {cos:.6f} relative L2 error {relnorm:.2e}")
        assert cos > 0.999 and relnorm < 0.03
        o += sz
    net.close()


def test_training_reduces_loss(cuda):
    """Adam steps on one batch: the loss falls well below the initial (near box-filter) network's."""
    inp = synthetic_batch(5)
    # target: a smooth function of the inputs the network can learn (the noisy image's
    # 5x5 box mean), so the optimum is reachable
    noisy = cnn.half_to_float(inp["noisy"])[..., :3]
    tgt = np.ones((N, H, W, 4), np.float32)
    tgt[..., :3] = cnn.gather_taps(noisy)[..., 12, :] * 0.5 + cnn.gather_taps(noisy).mean(3) * 0.5
    inp["target"] = tgt
    net = Net(11)
    im, t = device_images(cuda, inp, training=True)
    loss = cuda.zeros(1, dtype=cuda.float32, device="cuda")
    losses = []
    for step in range(300):
        capi.check(capi.lib().pbrt_cnn_train_step(net.h, C.byref(im), C.c_float(2e-3), capi.ptr(loss),
                                                  helpers.stream(cuda)))
        if step % 50 == 0 or step == 299:
            cuda.cuda.synchronize()
            losses.append(loss.item())
    print("cnn training losses", [f"{x:.4f}" for x in losses])
    assert losses[-1] < 0.5 * losses[0]
    net.close()


def test_invalid_arguments(cuda):
    net = Net(1)
    im = capi.CnnImages()
    assert capi.lib().pbrt_cnn_denoise(net.h, C.byref(im), None) != 0
    inp = synthetic_batch(6)
    im, t = device_images(cuda, inp, training=False)
    im.workspace_bytes = 16
    assert capi.lib().pbrt_cnn_denoise(net.h, C.byref(im), None) != 0
    assert capi.lib().pbrt_cnn_train_step(net.h, C.byref(im), C.c_float(1e-3), None, None) != 0  # no target / ws
    net.close()
    h = C.c_void_p()
    for bad in (0, 1, 7):
        assert capi.lib().pbrt_cnn_create_layers(bad, 1, C.byref(h)) != 0
    two = Net(1, 2)
    im, t = device_images(cuda, inp, training=False)
    im.denoised = None
    im.final_image = None
    assert capi.lib().pbrt_cnn_denoise(two.h, C.byref(im), None) != 0  # nothing to write
    n = C.c_uint32()
    assert capi.lib().pbrt_cnn_layers(two.h, C.byref(n)) == 0 and n.value == 2
    two.close()


def test_compact_training_keeps_pointwise_layer(cuda):
    """Adam steps on the compact network: the loss falls and the second layer stays
    pointwise (its outer taps exactly zero, also after set_parameters is handed
    non-zero ones)."""
    inp = synthetic_batch(6)
    net = Net(5, "c")
    params = net.get()
    params[outer_taps(2)] = 0.25
    net.set(params)
    assert not net.get()[outer_taps(2)].any()
    im, t = device_images(cuda, inp, training=True)
    loss = cuda.zeros(1, dtype=cuda.float32, device="cuda")
    losses = []
    for _ in range(60):
        capi.check(capi.lib().pbrt_cnn_train_step(net.h, C.byref(im), C.c_float(2e-3), capi.ptr(loss),
                                                  helpers.stream(cuda)))
        cuda.cuda.synchronize()
        losses.append(loss.item())
    print(f"compact training: loss {losses[0]:.5f} -> {losses[-1]:.5f}")
    assert losses[-1] < 0.8 * losses[0]
    p = net.get()
    assert not p[outer_taps(2)].any() and p[outer_taps(2)[0] + 4::9][:10].any()
    flag = C.c_uint32()
    capi.check(capi.lib().pbrt_cnn_is_compact(net.h, C.byref(flag)))
    assert flag.value == 1
    net.close()
