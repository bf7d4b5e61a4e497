"""Writes tests/golden/fwht_golden.npz from the reference's own fht_float.

The reference FFHT (src/fht.c + src/fast_copy.c, pure C) is compiled from the
reference tree by oracle/Makefile into oracle/_ref/ (SSE and AVX builds).  This
script feeds seeded random float32 vectors through both builds and stores
inputs and outputs for log_n = 1..10.  Run from the repository root:

    make -C oracle && python tests/golden/make_fwht_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "fwht_golden.npz")


def main():
    libs = {}
    for v in ("sse", "avx"):
        p = os.path.join(ROOT, "oracle", "_ref", f"libreffht_{v}.so")
        if not os.path.exists(p):
            sys.exit(f"missing {p}: run make -C oracle with the reference tree present")
        L = C.CDLL(p)
        L.fht_float.argtypes = [np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS"), C.c_int]
        libs[v] = L
    rng = np.random.default_rng(20240601)
    data = {}
    for log_n in range(1, 11):
        n = 1 << log_n
        x = rng.standard_normal((6, n)).astype(np.float32)
        # one row of small integers (exact arithmetic) and one with zeros/ties
        x[4] = rng.integers(-3, 4, n).astype(np.float32)
        x[5, : n // 2] = 0.0
        outs = []
        for v in ("sse", "avx"):
            y = x.copy()
            for row in y:
                r = np.ascontiguousarray(row)
                libs[v].fht_float(r, log_n)
                row[:] = r
            outs.append(y)
        assert np.array_equal(outs[0], outs[1]), f"SSE/AVX builds disagree at log_n={log_n}"
        data[f"x{log_n}"] = x
        data[f"y{log_n}"] = outs[0]
    np.savez(OUT, **data)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
