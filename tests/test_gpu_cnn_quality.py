"""What the trained network gains: on held-out Cornell-box scenes (seeds never used in
training), the committed parameters (vulkanpbrt_amd/weights/kpn_cornell.bin) against the
method without them - the 1-spp input itself, the untrained network and BMFR - by relMSE
of the remodulated radiance against 256-spp references rendered by the tracer."""
import numpy as np
import pytest

from vulkanpbrt_amd import denoiser_train as dt

pytestmark = pytest.mark.gpu


def test_trained_network_beats_untrained_noisy_and_bmfr(cuda, pbrt_lib):
    pairs = dt.render_pairs(cuda, np.arange(95_000, 95_008), 128, 128, 256)
    trained = dt.Network(1)
    trained.load(dt.DEFAULT_WEIGHTS)
    untrained = dt.Network(1)
    t = dt.evaluate(cuda, trained, pairs)
    u = dt.evaluate(cuda, untrained, pairs)
    print(f"relMSE noisy {t['relmse_noisy']:.4f} bmfr {t['relmse_bmfr']:.4f} "
          f"untrained {u['relmse_cnn']:.4f} trained {t['relmse_cnn']:.4f}")
    assert t["relmse_cnn"] < 0.25 * t["relmse_noisy"]
    assert t["relmse_cnn"] < 0.5 * t["relmse_bmfr"]
    assert t["relmse_cnn"] < 0.01 * u["relmse_cnn"]
    trained.close()
    untrained.close()


def test_two_layer_network_beats_bmfr(cuda, pbrt_lib):
    """The fused two-layer network (vulkanpbrt_amd/weights/kpn2_cornell.bin) reaches
    BMFR's relMSE with room to spare on held-out scenes, so it is the network the
    equal-quality timing comparison runs."""
    pairs = dt.render_pairs(cuda, np.arange(95_000, 95_008), 128, 128, 256)
    net = dt.Network(1, layers=2)
    net.load(dt.TWO_LAYER_WEIGHTS)
    r = dt.evaluate(cuda, net, pairs)
    print(f"two layers: relMSE noisy {r['relmse_noisy']:.4f} bmfr {r['relmse_bmfr']:.4f} cnn {r['relmse_cnn']:.4f}")
    assert r["relmse_cnn"] < 0.5 * r["relmse_bmfr"]
    assert r["relmse_cnn"] < 0.25 * r["relmse_noisy"]
    net.close()
