"""Write tests/golden/bmfr_noise_kat.json: known answers of the BMFR feature noise.

Independent Python-integer restatement of random(uint) and addRandom
(bmfrGeneral.comp:83-97): Bob Jenkins' 6-shift 32-bit integer hash, scaled by
float(0xffffffff) (= 2^32 in binary32), noise = 2e-4 * (hash - 0.5).
Run: python tests/golden/make_bmfr_noise_golden.py
"""
import json
from pathlib import Path

import numpy as np

M = 0xFFFFFFFF


def jenkins(a):
    a = ((a + 0x7ED55D16) + (a << 12)) & M
    a = ((a ^ 0xC761C23C) ^ (a >> 19)) & M
    a = ((a + 0x165667B1) + (a << 5)) & M
    a = ((a + 0xD3A2646C) ^ (a << 9)) & M
    a = ((a + 0xFD7046C5) + (a << 3)) & M
    a = ((a ^ 0xB55A4F09) ^ (a >> 16)) & M
    return a


def noised(value, key):
    r = np.float32(np.uint32(jenkins(key))) / np.float32(4294967296.0)
    return np.float32(np.float32(value) + np.float32(np.float32(2e-4) * np.float32(r - np.float32(0.5))))


def key(i, col, frame, pb):
    # id + sub * BLOCK_WIDTH == row index i; int arithmetic wraps
    return (i + col * pb * pb + frame * 13 * pb * pb) & M


def main():
    cases = []
    for pb in (64, 256, 1024):
        for (i, col, frame) in [(0, 0, 0), (5, 3, 1), (1023 % pb, 9, 15), (17, 4, 157), (17, 4, 158),
                                (pb - 1, 7, 100000)]:
            k = key(i, col, frame, pb)
            cases.append({"pb": pb, "i": i, "col": col, "frame": frame, "key": k, "hash": jenkins(k),
                          "noised_one_bits": int(noised(1.0, k).view(np.uint32))})
    Path(__file__).with_name("bmfr_noise_kat.json").write_text(json.dumps(cases, indent=1))


if __name__ == "__main__":
    main()
