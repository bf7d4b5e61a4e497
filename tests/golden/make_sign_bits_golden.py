"""Writes tests/golden/sign_bits_golden.npz: the first 4096 interleaved sign
bits (bitHD1, bitHD2) of src/FalconnPP.h:52-75 for seeds 0, 1, 42, 12345.

The bits are computed from the closed form of libstdc++'s
std::default_random_engine (minstd_rand0: x <- 16807 x mod 2^31-1, x0 = seed
mod m, 0 -> 1) and uniform_int_distribution<uint32_t>(0, 1) (down-scaling:
urange 2147483645, scaling = 2147483645 // 2, reject x-1 >= 2 * scaling,
bit = (x-1) // scaling).  The oracle test then checks libstdc++ itself
against these vectors.  Run from the repository root.
"""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "sign_bits_golden.npz")
M = 2147483647
A = 16807


def closed_form_bits(seed: int, nbits: int):
    x = seed % M
    if x == 0:
        x = 1
    urng_range = (M - 1) - 1  # max() - min() of minstd_rand0
    scaling = urng_range // 2
    past = 2 * scaling

    def draw():
        nonlocal x
        while True:
            x = (A * x) % M
            r = x - 1
            if r < past:
                return r // scaling

    b1 = np.zeros(nbits, np.uint8)
    b2 = np.zeros(nbits, np.uint8)
    for d in range(nbits):
        b1[d] = draw()
        b2[d] = draw()
    return b1, b2


def main():
    data = {}
    for s in (0, 1, 42, 12345):
        b1, b2 = closed_form_bits(s, 4096)
        data[f"b1_{s}"] = b1
        data[f"b2_{s}"] = b2
    np.savez(OUT, **data)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
