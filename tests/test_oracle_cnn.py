"""Pins of the CPU oracle of the learned denoiser (oracle/cnn_oracle.py).

  * the convolution equals scipy.signal.correlate2d summed over input channels
    (an independent implementation), with zero padding;
  * the 5x5 kernel gather equals explicit clamped indexing;
  * the analytic gradient equals central finite differences of the loss in
    float64 for every layer's weights and biases (bf16 rounding off);
  * a uniform softmax (all logits equal) makes the output the 5x5 box filter of
    the noisy input.
"""
import numpy as np
import pytest
from scipy.signal import correlate2d

from oracle import cnn_oracle as cnn


def test_conv_matches_scipy():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 7, 9, 3))
    w = rng.normal(size=(4, 3, 3, 3))
    b = rng.normal(size=4)
    y = cnn.conv3x3(x, w, b)
    for n in range(2):
        for o in range(4):
            ref = b[o] + sum(correlate2d(x[n, :, :, c], w[o, c], mode="same", boundary="fill") for c in range(3))
            np.testing.assert_allclose(y[n, :, :, o], ref, rtol=1e-12, atol=1e-12)


def test_gather_taps_clamps_to_edge():
    img = np.arange(2 * 4 * 5 * 1, dtype=np.float64).reshape(2, 4, 5, 1)
    t = cnn.gather_taps(img)
    for y in range(4):
        for x in range(5):
            for ky in range(5):
                for kx in range(5):
                    yy = min(max(y + ky - 2, 0), 3)
                    xx = min(max(x + kx - 2, 0), 4)
                    assert t[1, y, x, ky * 5 + kx, 0] == img[1, yy, xx, 0]


def test_uniform_kernel_is_box_filter():
    rng = np.random.default_rng(1)
    params = cnn.init_parameters(3) * 0.0  # zero weights and biases: all logits 0
    feats = rng.normal(size=(1, 6, 7, 12))
    noisy = rng.uniform(0, 2, size=(1, 6, 7, 3))
    out, _ = cnn.forward(params, feats, noisy, bf16_mirror=False)
    box = cnn.gather_taps(noisy).mean(axis=3)
    np.testing.assert_allclose(out, box, rtol=1e-12)


@pytest.mark.parametrize("layers", [6, 2])
def test_gradient_matches_finite_differences(layers):
    rng = np.random.default_rng(2)
    shape = cnn.shape_of(layers)
    params = cnn.init_parameters(5, shape)
    params[-25:] = rng.normal(0, 0.3, 25)  # non-trivial last-layer bias
    feats = rng.normal(size=(1, 5, 6, 12))
    noisy = rng.uniform(0, 2, size=(1, 5, 6, 3))
    target = rng.uniform(0, 2, size=(1, 5, 6, 3))
    albedo = rng.integers(0, 256, size=(1, 5, 6, 4), dtype=np.uint8)
    out, cache = cnn.forward(params, feats, noisy, bf16_mirror=False, shape=shape)
    g = cnn.backward(cache, out, target, albedo)
    # probe a handful of entries in every layer's weights and bias
    idx, o = [], 0
    for ws, bs in cnn.layer_shapes(shape):
        n = int(np.prod(ws))
        idx += list(rng.choice(n, 4, replace=False) + o) + [o + n, o + n + bs[0] - 1]
        o += n + bs[0]
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        l1 = cnn.loss(cnn.forward(p1, feats, noisy, bf16_mirror=False, shape=shape)[0], target, albedo)
        l2 = cnn.loss(cnn.forward(p2, feats, noisy, bf16_mirror=False, shape=shape)[0], target, albedo)
        fd = (l1 - l2) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_parameter_count():
    assert cnn.parameter_count() == 9 * 12 * 32 + 32 + 4 * (9 * 32 * 32 + 32) + 9 * 32 * 25 + 25
    assert cnn.shape_of(6) == cnn.DEFAULT
    assert cnn.parameter_count(cnn.shape_of(2)) == 9 * 12 * 32 + 32 + 9 * 32 * 25 + 25


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38], np.float32)
    r = cnn.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie -> even mantissa
    assert r[2] == 1.0 + 2 ** -6  # tie -> even (rounds up)
    assert r[3] == -2.5
    assert np.isfinite(r[4])
