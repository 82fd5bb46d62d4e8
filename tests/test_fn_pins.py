"""Pins of the project's transcendentals (vulkanpbrt_amd/csrc/pbrt_fn.hpp), which the
device's shading code and the oracle's FP32 mirror both evaluate: each against the
float64 libm value over the arguments the path tracer passes, to a stated number of
fp32 ulps, plus its special values."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.usefixtures("oracle_lib")


def ulps(got, want):
    want = np.asarray(want, np.float64)
    spacing = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / np.maximum(spacing, np.float64(np.spacing(np.float32(0))))


def test_log2():
    rng = np.random.default_rng(1)
    x = np.concatenate([np.exp2(rng.uniform(-126, 127, 200_000)), rng.uniform(0, 4, 200_000),
                        1 + rng.uniform(-1e-3, 1e-3, 10_000)]).astype(np.float32)
    x = x[x > 0]
    got = oracle.fn_probe("log2", x)
    want = np.log2(x.astype(np.float64))
    # absolute error in log units: what exp2(y log2 x) amplifies
    assert np.abs(got - want).max() <= 3e-7 * np.maximum(1.0, np.abs(want)).max()
    assert np.all(np.abs(got - want) <= 2.5e-7 * np.maximum(1.0, np.abs(want)))
    sub = np.float32(1e-40)  # subnormal
    assert abs(oracle.fn_probe("log2", [sub])[0] - np.log2(1e-40)) < 1e-4
    assert oracle.fn_probe("log2", [1.0])[0] == 0.0 and oracle.fn_probe("log2", [2.0])[0] == 1.0
    assert oracle.fn_probe("log2", [0.0])[0] == -np.inf
    assert np.isnan(oracle.fn_probe("log2", [-1.0])[0]) and oracle.fn_probe("log2", [np.inf])[0] == np.inf


def test_exp2():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-125, 127, 200_000), rng.uniform(-1, 1, 200_000)]).astype(np.float32)
    got = oracle.fn_probe("exp2", x)
    assert ulps(got, np.exp2(x.astype(np.float64))).max() <= 2.0
    assert np.array_equal(oracle.fn_probe("exp2", [0.0, 1.0, -3.0, 10.0]), [1.0, 2.0, 0.125, 1024.0])
    assert oracle.fn_probe("exp2", [200.0])[0] == np.inf and oracle.fn_probe("exp2", [-200.0])[0] == 0.0


@pytest.mark.parametrize("e", [2.2, 1 / 2.2])
def test_pow_gamma(e):
    """sRGB <-> linear of color.glsl and the final image, x in (0, 16] (radiance is
    clamped to 10).  exp2(y log2 x) in fp32 carries the log's absolute error into the
    exponent, so the bound scales with |y log2 x|: relative error <= 1.2e-7 (1 +
    |y log2 x|) (measured 8e-8 (1 + |y log2 x|)); pow(0, e) = 0 and pow(1, e) = 1."""
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(0, 1, 300_000), rng.uniform(1, 16, 50_000),
                        np.arange(1, 256) / 255.0]).astype(np.float32)
    x = x[x > 0]
    ef = np.float32(e)
    got = oracle.fn_probe("pow", x, np.full_like(x, ef)).astype(np.float64)
    want = np.power(x.astype(np.float64), np.float64(ef))
    scale = 1.0 + np.abs(np.float64(ef) * np.log2(x.astype(np.float64)))
    assert np.all(np.abs(got - want) / want <= 1.2e-7 * scale)
    assert oracle.fn_probe("pow", [0.0, 1.0], [ef, ef]).tolist() == [0.0, 1.0]


def test_pow5():
    """(x^2)^2 x: three roundings, relative error <= 4 * 2^-24."""
    x = np.random.default_rng(4).uniform(1e-3, 1, 300_000).astype(np.float32)
    got, want = oracle.fn_probe("pow5", x).astype(np.float64), x.astype(np.float64) ** 5
    assert (np.abs(got - want) / want).max() <= 4 * 2.0 ** -24


def test_sincos():
    """The angles of the BRDF samplers (2 pi u, u in [0, 1)) and of the G-buffer normal
    decode ([-pi, pi]): within 1.5 ulp, or 1e-7 absolute near the zeros."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-np.pi, 2 * np.pi, 400_000), np.linspace(-4, 4, 100_001)]).astype(np.float32)
    for kind, fn in (("sin", np.sin), ("cos", np.cos)):
        got = oracle.fn_probe(kind, x)
        want = fn(x.astype(np.float64))
        err = np.abs(got - want)
        assert np.all((ulps(got, want) <= 1.5) | (err <= 1e-7)), kind
    assert oracle.fn_probe("sin", [0.0])[0] == 0.0 and oracle.fn_probe("cos", [0.0])[0] == 1.0


def test_planted_mistake_is_caught():
    """The pins fail on a wrong polynomial: a log2 that drops its highest term
    (emulated here on the probe's own output) misses the bound."""
    x = np.linspace(0.7, 1.4, 10_001).astype(np.float32)
    f = x.astype(np.float64) - 1
    wrong = oracle.fn_probe("log2", x) - np.float32(1.228052825e-01) * f ** 9
    assert np.abs(wrong - np.log2(x.astype(np.float64))).max() > 3e-7
