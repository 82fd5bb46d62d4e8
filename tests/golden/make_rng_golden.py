"""Write tests/golden/rng_kat.json: known-answer vectors for shaders/random.glsl.

An independent restatement in Python integers (mod 2^32 arithmetic) and numpy
float32, written from the published definitions: Thomas Wang's 32-bit integer
hash (random.glsl:74-81, as popularised by N. Reed's "Quick and easy GPU random
numbers"), L'Ecuyer's three-component Tausworthe combined with an LCG
(1664525, 1013904223) as in GPU Gems 3 ch. 37 (random.glsl:13-49).
Run: python tests/golden/make_rng_golden.py
"""
import json
from pathlib import Path

import numpy as np

M = 0xFFFFFFFF


def wang(s):
    s = ((s ^ 61) ^ (s >> 16)) & M
    s = (s * 9) & M
    s = s ^ (s >> 4)
    s = (s * 0x27D4EB2D) & M
    return s ^ (s >> 15)


def taus(z, s1, s2, s3, m):
    b = (((z << s1) & M) ^ z) >> s2
    return (((z & m) << s3) & M) ^ b


def stream(x, y, frame, n):
    st = [wang(x), wang(y), wang(frame), wang(((x + y) & M) * frame & M)]
    out = []
    scale = np.float32(2.0 ** -32)
    for _ in range(n):
        st[0] = taus(st[0], 13, 19, 12, 4294967294)
        st[1] = taus(st[1], 2, 25, 4, 4294967288)
        st[2] = taus(st[2], 3, 11, 17, 4294967280)
        st[3] = (1664525 * st[3] + 1013904223) & M
        v = np.float32(np.uint32(st[0] ^ st[1] ^ st[2] ^ st[3])) * scale
        out.append(int(np.float32(v).view(np.uint32)))
    return out


def main():
    hashes = {str(s): wang(s) for s in [0, 1, 2, 61, 255, 256, 1023, 65535, 65536, 123456789, 0xDEADBEEF, M]}
    cases = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (17, 42, 3), (255, 255, 1), (1919, 1079, 63), (4095, 4095, 12345),
             (3, 5, 0xFFFFFFFF)]
    streams = [{"x": x, "y": y, "frame": f, "bits": stream(x, y, f, 64)} for (x, y, f) in cases]
    out = {"wang_hash": hashes, "streams": streams}
    Path(__file__).with_name("rng_kat.json").write_text(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
