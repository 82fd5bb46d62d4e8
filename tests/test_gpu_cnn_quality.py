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
