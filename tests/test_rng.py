"""random.glsl: the oracle's C++ restatement against the independent Python
known-answer vectors in tests/golden/rng_kat.json."""
import json
from pathlib import Path

import numpy as np

KAT = json.loads((Path(__file__).parent / "golden" / "rng_kat.json").read_text())


def test_scale_literal_is_two_to_minus_32():
    # random.glsl:46: the literal rounds to exactly 2^-32 in binary32
    assert np.float32(2.3283064365387e-10) == np.float32(2.0 ** -32)


def test_u32_to_float_edge_gives_one():
    # float(u32) rounds to nearest: values >= 2^32 - 128 become 2^32, so randomFloat == 1.0
    assert np.float32(np.uint32(0xFFFFFF80)) * np.float32(2.0 ** -32) == np.float32(1.0)
    assert np.float32(np.uint32(0xFFFFFF7F)) * np.float32(2.0 ** -32) < np.float32(1.0)


def test_wang_hash_kat(oracle_lib):
    for s, h in KAT["wang_hash"].items():
        assert oracle_lib.wang_hash(int(s)) == h


def test_streams_kat(oracle_lib):
    for case in KAT["streams"]:
        x, y, f = case["x"], case["y"], case["frame"]
        seq = oracle_lib.rng_sequence(x + 1, y + 1, f, 64)[y, x]
        assert seq.tolist() == case["bits"], case


def test_streams_in_unit_interval(oracle_lib):
    seq = oracle_lib.rng_sequence(64, 64, 7, 32).view(np.float32)
    assert (seq >= 0).all() and (seq <= 1).all()
    # mean of uniforms: 64*64*32 samples, 6 sigma band
    n = seq.size
    assert abs(seq.mean() - 0.5) < 6 * (1 / np.sqrt(12 * n))
