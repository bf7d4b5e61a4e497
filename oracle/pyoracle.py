"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use
this module, to check results or as the CPU baseline; the product never imports
it.  The C++ behind it restates the reference's 1D path
(src/FalconnPP.cpp:21-707, src/FalconnPP.h:52-100) with libstdc++ heaps; see
the header of falconn_oracle.cpp for how each function is pinned.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def build_oracle() -> None:
    """Compile oracle/liboracle.so (and oracle/_ref when the reference exists)."""
    subprocess.check_call(["make", "-s", "-C", _HERE], stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build_oracle()
        L = C.CDLL(_LIB)
        L.orc_sign_bits.argtypes = [C.c_uint, C.c_int, u8p, u8p]
        L.orc_fwht.argtypes = [f32p, C.c_int]
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_index_param.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_int]
        L.orc_set_qprobes.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.orc_fht_dim.argtypes = [C.c_void_p]
        L.orc_build.argtypes = [C.c_void_p, f32p, C.c_int64, C.c_void_p]
        L.orc_build_seconds.restype = C.c_double
        L.orc_build_seconds.argtypes = [C.c_void_p]
        L.orc_limited_buckets.restype = C.c_int64
        L.orc_limited_buckets.argtypes = [C.c_void_p]
        L.orc_num_ids.restype = C.c_int64
        L.orc_num_ids.argtypes = [C.c_void_p]
        L.orc_num_buckets.restype = C.c_int64
        L.orc_num_buckets.argtypes = [C.c_void_p]
        L.orc_get_center.argtypes = [C.c_void_p, f32p]
        L.orc_get_sign_bits.argtypes = [C.c_void_p, u8p, u8p]
        L.orc_get_buckets.argtypes = [C.c_void_p, u64p, u32p, i32p]
        L.orc_hash_build.argtypes = [C.c_void_p, f32p, C.c_int, i32p, f32p]
        L.orc_set_sign_bits.argtypes = [C.c_void_p, C.c_uint]
        L.orc_hash_query_layers.argtypes = [C.c_void_p, f32p, C.c_int, C.c_int, i32p, f32p, i32p, f32p]
        L.orc_per_table.argtypes = [C.c_void_p]
        L.orc_per_layer.argtypes = [C.c_void_p]
        L.orc_query.argtypes = [C.c_void_p, f32p, C.c_int, C.c_int, i32p, f32p, i64p]
        L.orc_query_debug.restype = C.c_int64
        L.orc_query_debug.argtypes = [C.c_void_p, f32p, C.c_int, i32p, f32p, i32p, f32p, i32p, C.c_int64, i64p]
        L.orc_load_index.argtypes = [C.c_void_p, f32p, C.c_int64, f32p, u64p, i32p]
        L.orc_bruteforce_f64.argtypes = [f32p, C.c_int, C.c_int, f32p, C.c_int, C.c_int, i32p, f64p, C.c_int]
        _lib = L
    return _lib


def sign_bits(seed: int, nbits: int):
    b1 = np.zeros(nbits, np.uint8)
    b2 = np.zeros(nbits, np.uint8)
    lib().orc_sign_bits(seed, nbits, b1, b2)
    return b1, b2


def fwht(x: np.ndarray) -> np.ndarray:
    y = np.ascontiguousarray(x, dtype=np.float32).copy()
    n = y.shape[-1]
    log_n = n.bit_length() - 1
    for row in y.reshape(-1, n):
        r = np.ascontiguousarray(row)
        lib().orc_fwht(r, log_n)
        row[:] = r
    return y


def _cm(a: np.ndarray) -> np.ndarray:
    """N x D (row per point) float32, C-contiguous == col-major D x N."""
    return np.ascontiguousarray(a, dtype=np.float32)


class OracleIndex:
    """CPU restatement of the reference FalconnPP class (1D index)."""

    def __init__(self, n_points: int, n_features: int):
        self.n = n_points
        self.d = n_features
        self.h = lib().orc_create(n_points, n_features)
        self.qprobes = 1

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def setIndexParam(self, n_tables, n_proj, bucket_minSize, bucket_scale, iProbes, n_threads=8, random_seed=-1):
        self.L, self.n_proj, self.iProbes = n_tables, n_proj, iProbes
        lib().orc_set_index_param(self.h, n_tables, n_proj, bucket_minSize, bucket_scale, iProbes, n_threads,
                                  random_seed)
        self.seed = random_seed

    @property
    def fht_dim(self) -> int:
        return lib().orc_fht_dim(self.h)

    def set_qProbes(self, p: int):
        self.qprobes = p
        lib().orc_set_qprobes(self.h, p)

    def set_threads(self, t: int):
        lib().orc_set_threads(self.h, t)

    def build(self, X_nd: np.ndarray, center: np.ndarray | None = None):
        """X_nd: N x D points (rows); the reference's D x N col-major layout."""
        X = _cm(X_nd)
        c = None if center is None else np.ascontiguousarray(center, dtype=np.float32)
        lib().orc_build(self.h, X, X.shape[1], None if c is None else c.ctypes.data)
        self._center_keep = c

    @property
    def build_seconds(self) -> float:
        return lib().orc_build_seconds(self.h)

    def center(self) -> np.ndarray:
        c = np.zeros(self.d, np.float32)
        lib().orc_get_center(self.h, c)
        return c

    def buckets(self):
        nb = lib().orc_num_buckets(self.h)
        ni = lib().orc_num_ids(self.h)
        pos = np.zeros(nb, np.uint64)
        size = np.zeros(nb, np.uint32)
        ids = np.zeros(max(ni, 1), np.int32)
        lib().orc_get_buckets(self.h, pos, size, ids)
        return pos, size, ids[:ni]

    def load_index(self, X_nd: np.ndarray, center: np.ndarray, off: np.ndarray, ids: np.ndarray):
        """Adopt CSR tables built elsewhere (uses the seed given to setIndexParam)."""
        X = _cm(X_nd)
        lib().orc_load_index(self.h, X, X.shape[1], np.ascontiguousarray(center, np.float32),
                             np.ascontiguousarray(off, np.uint64), np.ascontiguousarray(ids, np.int32))

    def set_sign_bits(self, seed: int):
        lib().orc_set_sign_bits(self.h, seed)

    def hash_build(self, V_md: np.ndarray):
        V = _cm(V_md)
        m = V.shape[0]
        codes = np.zeros(self.L * m * self.iProbes, np.int32)
        vals = np.zeros(self.L * m * self.iProbes, np.float32)
        lib().orc_hash_build(self.h, V, m, codes, vals)
        return codes.reshape(self.L, m, self.iProbes), vals.reshape(self.L, m, self.iProbes)

    def hash_query_layers(self, q: np.ndarray, table: int, per_layer: int):
        c1 = np.zeros(per_layer, np.int32)
        c2 = np.zeros(per_layer, np.int32)
        v1 = np.zeros(per_layer, np.float32)
        v2 = np.zeros(per_layer, np.float32)
        lib().orc_hash_query_layers(self.h, np.ascontiguousarray(q, np.float32), table, per_layer, c1, v1, c2, v2)
        return c1, v1, c2, v2

    def per_table(self) -> int:
        return lib().orc_per_table(self.h)

    def per_layer(self) -> int:
        return lib().orc_per_layer(self.h)

    def query(self, Q_nd: np.ndarray, k: int):
        Q = _cm(Q_nd)
        nq = Q.shape[0]
        topk = np.zeros(nq * k, np.int32)
        scores = np.zeros(nq * k, np.float32)
        stats = np.zeros(3, np.int64)
        lib().orc_query(self.h, Q, nq, k, topk, scores, stats)
        return topk.reshape(nq, k), scores.reshape(nq, k), stats

    def query_debug(self, q: np.ndarray, k: int, cap: int | None = None):
        cap = self.n if cap is None else cap
        topk = np.zeros(k, np.int32)
        scores = np.zeros(k, np.float32)
        probes = np.zeros(self.qprobes, np.int32)
        pvals = np.zeros(self.qprobes, np.float32)
        cands = np.zeros(max(cap, 1), np.int32)
        ctot = np.zeros(1, np.int64)
        nc = lib().orc_query_debug(self.h, np.ascontiguousarray(q, np.float32), k, topk, scores, probes, pvals,
                                   cands, cap, ctot)
        return dict(topk=topk, scores=scores, probes=probes, probe_vals=pvals, cands=cands[:min(nc, cap)],
                    n_cands=nc, ctot=int(ctot[0]))


def bruteforce_f64(X_nd: np.ndarray, Q_nd: np.ndarray, k: int, threads: int = 0):
    X = _cm(X_nd)
    Q = _cm(Q_nd)
    nq = Q.shape[0]
    ids = np.zeros(nq * k, np.int32)
    sc = np.zeros(nq * k, np.float64)
    lib().orc_bruteforce_f64(X, X.shape[0], X.shape[1], Q, nq, k, ids, sc, threads)
    return ids.reshape(nq, k), sc.reshape(nq, k)
