"""ctypes binding of the engine's C ABI (include/falconnpp_amd.h).

This is how tests and the benchmark drive the HIP library directly, including
the device-pointer entry points.  The drop-in Python surface of the reference
(``FalconnPP.FalconnPP``) is the pybind11 module built beside this file.
There is no fallback: if the native library is missing, loading fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfalconnpp_amd.so")

# Every symbol include/falconnpp_amd.h declares.
EXPORTS = (
    "fpp_last_error", "fpp_create", "fpp_destroy", "fpp_set_index_param", "fpp_set_qprobes", "fpp_set_threads",
    "fpp_clear", "fpp_build", "fpp_build_device", "fpp_set_insertable", "fpp_insert", "fpp_insert_device", "fpp_query", "fpp_query_2d", "fpp_query_device", "fpp_query_device_noalloc", "fpp_get_shape",
    "fpp_get_center", "fpp_get_sign_bits", "fpp_get_buckets", "fpp_get_bucket_counts", "fpp_hash_points", "fpp_get_stats",
    "fpp_get_last_probes", "fpp_get_last_candidates", "fpp_bruteforce_f64", "fpp_set_profiling",
    "fpp_get_timings", "fpp_set_tuning", "fpp_set_precision", "fpp_get_fp8_scale", "fpp_autotune", "fpp_autotune_batch",
    "fpp_query_batch_device",
    "fpp_get_query_times", "fpp_load_txt", "fpp_write_topk_txt", "fpp_topk_file_name", "fpp_topk_file_name_2d",
    "fpp_save", "fpp_load",
)
PREC_FP32, PREC_BF16, PREC_FP8 = 0, 1, 2
TUNE_QUERY_ORDER, TUNE_GATHER_THREADS, TUNE_PERSISTENT, TUNE_GRAPH, TUNE_VERBOSE, TUNE_FP8_PER_ROW = 1, 2, 3, 4, 5, 6
TUNE_KEEP_CANDIDATES = 7
TUNE_REPLAY = 8  # 0 per call, 1 separate kernels, 2 inside the gather
TUNE_WORKSPACE_MB = 9  # bound of the per-workgroup candidate scratch
TUNE_SMALL_QUERIES = 10  # queries of at most 512 bucket ids served one per wave (default on)

FPP_EINVAL, FPP_ESTATE, FPP_EHIP, FPP_ENOMEM, FPP_EINTERNAL, FPP_EIO = 1, 2, 3, 4, 5, 6

_lib = None


class FppError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def share_hip_runtime():
    """Import torch, when installed, before the library loads.  torch carries
    its own HIP runtime (loaded as libamdhip64.so); the library asks for
    libamdhip64.so.7.  Loaded first, torch's runtime serves both.  Loaded
    after the library, torch would load a second runtime into the process,
    and that one finds no GPU."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load():
    """Load libfalconnpp_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C falconnpp_amd/csrc` "
                          "(or __graft_entry__.build())")
    share_hip_runtime()
    L = C.CDLL(LIB_PATH)
    vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int
    L.fpp_last_error.restype = C.c_char_p
    L.fpp_create.argtypes = [i32, i32, C.POINTER(vp)]
    L.fpp_destroy.argtypes = [vp]
    L.fpp_destroy.restype = None
    L.fpp_set_index_param.argtypes = [vp, i32, i32, i32, C.c_float, i32, i32, i32]
    L.fpp_set_qprobes.argtypes = [vp, i32]
    L.fpp_set_threads.argtypes = [vp, i32]
    L.fpp_clear.argtypes = [vp]
    L.fpp_build.argtypes = [vp, vp, i64]
    L.fpp_build_device.argtypes = [vp, vp, i64, vp]
    L.fpp_set_insertable.argtypes = [vp, i32]
    L.fpp_insert.argtypes = [vp, vp, i64, i32]
    L.fpp_insert_device.argtypes = [vp, vp, i64, i32, vp]
    L.fpp_query.argtypes = [vp, vp, i64, i32, i32, vp, i32]
    L.fpp_query_2d.argtypes = [vp, vp, i64, i32, i32, vp, i32]
    L.fpp_query_device.argtypes = [vp, vp, i64, i32, i32, vp, vp, vp]
    L.fpp_query_device_noalloc.argtypes = [vp, vp, i64, i32, i32, vp, vp, vp]
    L.fpp_get_shape.argtypes = [vp, vp]
    L.fpp_get_center.argtypes = [vp, vp]
    L.fpp_get_sign_bits.argtypes = [vp, vp, vp]
    L.fpp_get_buckets.argtypes = [vp, vp, vp]
    L.fpp_get_bucket_counts.argtypes = [vp, vp]
    L.fpp_hash_points.argtypes = [vp, vp, i64, i32, vp, vp]
    L.fpp_get_stats.argtypes = [vp, vp]
    L.fpp_get_last_probes.argtypes = [vp, vp]
    L.fpp_set_profiling.argtypes = [vp, i32]
    L.fpp_get_timings.argtypes = [vp, vp]
    L.fpp_set_tuning.argtypes = [vp, i32, i32]
    L.fpp_set_precision.argtypes = [vp, i32, i32]
    L.fpp_get_fp8_scale.argtypes = [vp, vp]
    L.fpp_autotune.argtypes = [vp, vp, i64, i32, i32, vp, vp]
    L.fpp_autotune_batch.argtypes = [vp, vp, i64, i32, vp, vp, i32, vp, vp]
    L.fpp_get_query_times.argtypes = [vp, vp, vp]
    L.fpp_query_batch_device.argtypes = [vp, vp, i64, i32, vp, vp, i32, vp, vp, vp]
    L.fpp_get_last_candidates.restype = i64
    L.fpp_get_last_candidates.argtypes = [vp, i32, vp, i64]
    L.fpp_bruteforce_f64.argtypes = [vp, i64, i64, i32, vp, i64, i32, i32, vp, vp, vp]
    L.fpp_load_txt.argtypes = [C.c_char_p, i32, i32, vp, vp]
    L.fpp_save.argtypes = [vp, C.c_char_p]
    L.fpp_load.argtypes = [vp, C.c_char_p]
    L.fpp_write_topk_txt.argtypes = [C.c_char_p, vp, i32, i32]
    L.fpp_topk_file_name.argtypes = [i32, i32, i32, i32, i32, C.c_float, i32, C.c_char_p, i32]
    L.fpp_topk_file_name_2d.argtypes = [i32, i32, i32, i32, i32, C.c_float, i32, C.c_char_p, i32]
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise FppError(rc, load().fpp_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _points(a) -> np.ndarray:
    """N x D row-per-point float32 (== the reference's D x N column-major)."""
    return np.ascontiguousarray(a, dtype=np.float32)


def load_txt(path: str, n_points: int, n_features: int):
    """The reference's loadtxtData (src/Utilities.cpp:41-60): n_points x
    n_features float32 (row per point) and the number of values read; values
    the file lacks stay 0.  No device use."""
    out = np.zeros((n_points, n_features), np.float32)
    n = np.zeros(1, np.int64)
    _check(load().fpp_load_txt(os.fsencode(path), n_points, n_features, _ptr(out), _ptr(n)))
    return out, int(n[0])


def write_topk_txt(path: str, ids: np.ndarray):
    """The reference's outputFile (src/Utilities.cpp:13-31): one line per
    query, each id followed by a space.  No device use."""
    a = np.ascontiguousarray(ids, np.int32)
    _check(load().fpp_write_topk_txt(os.fsencode(path), _ptr(a), a.shape[0], a.shape[1]))


def topk_file_name(k: int, n_proj: int, n_tables: int, iProbes: int, bucket_minSize: int, bucket_scale: float,
                   qProbes: int, layout: str = "1D") -> str:
    """File name a verbose query writes: 1D_Top_... for query
    (src/FalconnPP.cpp:694-700), 2D_Top_... for query2D (:1400-1406)."""
    buf = C.create_string_buffer(512)
    fn = load().fpp_topk_file_name if layout == "1D" else load().fpp_topk_file_name_2d
    _check(fn(k, n_proj, n_tables, iProbes, bucket_minSize, bucket_scale, qProbes, buf, 512))
    return buf.value.decode()


class Index:
    """One engine index, driven through the C ABI."""

    def __init__(self, n_points: int, n_features: int):
        L = load()
        h = C.c_void_p()
        _check(L.fpp_create(n_points, n_features, C.byref(h)))
        self.h = h
        self.n, self.d = n_points, n_features
        self.qprobes = 1

    def close(self):
        if getattr(self, "h", None):
            load().fpp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def setIndexParam(self, n_tables, n_proj, bucket_minSize, bucket_scale, iProbes, n_threads=8, random_seed=-1):
        self.L, self.n_proj, self.iProbes = n_tables, n_proj, iProbes
        _check(load().fpp_set_index_param(self.h, n_tables, n_proj, bucket_minSize, bucket_scale, iProbes,
                                          n_threads, random_seed))

    def save(self, path: str):
        """Write the built index to path (no reference counterpart)."""
        _check(load().fpp_save(self.h, os.fsencode(path)))

    def load(self, path: str):
        """Restore an index written by save (same n_points and n_features)."""
        _check(load().fpp_load(self.h, os.fsencode(path)))

    def set_qProbes(self, p: int):
        self.qprobes = p
        _check(load().fpp_set_qprobes(self.h, p))

    def set_threads(self, t: int):
        _check(load().fpp_set_threads(self.h, t))

    def clear(self):
        _check(load().fpp_clear(self.h))

    def build(self, X_nd):
        X = _points(X_nd)
        assert X.shape == (self.n, self.d)
        _check(load().fpp_build(self.h, _ptr(X), self.d))

    def build_device(self, dX_ptr: int, ld: int, stream: int = 0):
        _check(load().fpp_build_device(self.h, C.c_void_p(dX_ptr), ld, C.c_void_p(stream or None)))

    def set_insertable(self, on: bool = True):
        """Keep the pre-filter bucket entries at the next build, for insert()."""
        _check(load().fpp_set_insertable(self.h, 1 if on else 0))

    def insert(self, X_md):
        """Append m points (rows) as ids n .. n+m-1; buckets are filtered again."""
        X = _points(X_md)
        assert X.shape[1] == self.d
        _check(load().fpp_insert(self.h, _ptr(X), self.d, X.shape[0]))
        self.n += X.shape[0]

    def query(self, Q_nd, k: int, verbose: bool = False) -> np.ndarray:
        Q = _points(Q_nd)
        nq = Q.shape[0]
        out = np.zeros((nq, k), np.int32)
        _check(load().fpp_query(self.h, _ptr(Q), self.d, nq, k, _ptr(out), 1 if verbose else 0))
        return out

    def query2D(self, Q_nd, k: int, verbose: bool = False) -> np.ndarray:
        """query2Layers (src/FalconnPP.cpp:1013-1413): same ids as query()."""
        Q = _points(Q_nd)
        nq = Q.shape[0]
        out = np.zeros((nq, k), np.int32)
        _check(load().fpp_query_2d(self.h, _ptr(Q), self.d, nq, k, _ptr(out), 1 if verbose else 0))
        return out

    def query_batch_device(self, dQ_ptr: int, ldq: int, nq: int, qprobes, ks, kmax: int, dout_ptr: int,
                           dscores_ptr: int = 0, stream: int = 0):
        """Per-query qProbes and k (host arrays); outputs nq x kmax on the device."""
        qp = np.ascontiguousarray(qprobes, dtype=np.int32)
        kk = np.ascontiguousarray(ks, dtype=np.int32)
        assert len(qp) == nq and len(kk) == nq
        _check(load().fpp_query_batch_device(self.h, C.c_void_p(dQ_ptr), ldq, nq, _ptr(qp), _ptr(kk), kmax,
                                             C.c_void_p(dout_ptr), C.c_void_p(dscores_ptr or None),
                                             C.c_void_p(stream or None)))

    def autotune(self, dQ_ptr: int, ldq: int, nq: int, k: int, stream: int = 0) -> dict:
        """Time the launch shapes on these queries and keep the fastest."""
        ch = np.zeros(4, np.int32)
        _check(load().fpp_autotune(self.h, C.c_void_p(dQ_ptr), ldq, nq, k, C.c_void_p(stream or None), _ptr(ch)))
        return {"gather_threads": int(ch[0]), "query_order": int(ch[1]), "persistent": int(ch[2]),
                "us_per_batch": int(ch[3])}

    def autotune_batch(self, dQ_ptr: int, ldq: int, nq: int, qprobes, ks, kmax: int, stream: int = 0) -> dict:
        """As autotune, timing batched calls with these per-query qProbes and k."""
        qp = np.ascontiguousarray(qprobes, dtype=np.int32)
        kk = np.ascontiguousarray(ks, dtype=np.int32)
        assert len(qp) == nq and len(kk) == nq
        ch = np.zeros(4, np.int32)
        _check(load().fpp_autotune_batch(self.h, C.c_void_p(dQ_ptr), ldq, nq, _ptr(qp), _ptr(kk), kmax,
                                         C.c_void_p(stream or None), _ptr(ch)))
        return {"gather_threads": int(ch[0]), "query_order": int(ch[1]), "persistent": int(ch[2]),
                "us_per_batch": int(ch[3])}

    def query_device(self, dQ_ptr: int, ldq: int, nq: int, k: int, dout_ptr: int, dscores_ptr: int = 0,
                     stream: int = 0):
        _check(load().fpp_query_device(self.h, C.c_void_p(dQ_ptr), ldq, nq, k, C.c_void_p(dout_ptr),
                                       C.c_void_p(dscores_ptr or None), C.c_void_p(stream or None)))

    def query_device_noalloc(self, dQ_ptr: int, ldq: int, nq: int, k: int, dout_ptr: int, dscores_ptr: int = 0,
                             stream: int = 0):
        """query_device with every scratch buffer in the library's fixed
        device array: the call allocates nothing."""
        _check(load().fpp_query_device_noalloc(self.h, C.c_void_p(dQ_ptr), ldq, nq, k, C.c_void_p(dout_ptr),
                                               C.c_void_p(dscores_ptr or None), C.c_void_p(stream or None)))

    # ---- inspection ----
    def shape(self):
        s = np.zeros(6, np.int64)
        _check(load().fpp_get_shape(self.h, _ptr(s)))
        return dict(fht_dim=int(s[0]), buckets_per_table=int(s[1]), n_tables=int(s[2]), n_ids=int(s[3]),
                    ldx=int(s[4]), n_points=int(s[5]))

    def center(self) -> np.ndarray:
        c = np.zeros(self.d, np.float32)
        _check(load().fpp_get_center(self.h, _ptr(c)))
        return c

    def sign_bits(self):
        sh = self.shape()
        n = sh["fht_dim"] * sh["n_tables"] * 3
        b1 = np.zeros(n, np.uint8)
        b2 = np.zeros(n, np.uint8)
        _check(load().fpp_get_sign_bits(self.h, _ptr(b1), _ptr(b2)))
        return b1, b2

    def buckets(self):
        sh = self.shape()
        nb = sh["buckets_per_table"] * sh["n_tables"]
        off = np.zeros(nb + 1, np.uint64)
        ids = np.zeros(max(sh["n_ids"], 1), np.int32)
        _check(load().fpp_get_buckets(self.h, _ptr(off), _ptr(ids)))
        return off, ids[: sh["n_ids"]]

    def bucket_counts(self) -> np.ndarray:
        """Bucket sizes before the alpha / bucketLimit filter."""
        sh = self.shape()
        cnt = np.zeros(sh["buckets_per_table"] * sh["n_tables"], np.uint32)
        _check(load().fpp_get_bucket_counts(self.h, _ptr(cnt)))
        return cnt

    def hash_points(self, V_md):
        V = _points(V_md)
        m = V.shape[0]
        codes = np.zeros(self.L * m * self.iProbes, np.uint32)
        vals = np.zeros(self.L * m * self.iProbes, np.float32)
        _check(load().fpp_hash_points(self.h, _ptr(V), self.d, m, _ptr(codes), _ptr(vals)))
        return codes.reshape(self.L, m, self.iProbes), vals.reshape(self.L, m, self.iProbes)

    def stats(self):
        s = np.zeros(8, np.int64)
        _check(load().fpp_get_stats(self.h, _ptr(s)))
        return dict(hash_ties=int(s[0]), bucket_ties=int(s[1]), ctot=int(s[2]), cuniq=int(s[3]),
                    query_ties=int(s[4]), nq=int(s[5]), probe_path=int(s[6]), small_queries=int(s[7]))

    def set_profiling(self, level: int):
        """0 off, 1 HIP events per query stage, 2 also per-query phase stamps."""
        _check(load().fpp_set_profiling(self.h, int(level)))

    def timings(self):
        """Device ms of the last query (hash, probe select, tie replay, gather+rerank),
        mean per-query us of the gather kernel's phases, the longest query's us
        in that kernel and the longest first phase (replay + bucket gather) of
        a tie-flagged query."""
        t = np.zeros(10, np.float64)
        _check(load().fpp_get_timings(self.h, _ptr(t)))
        return dict(hash=t[0], probe=t[1], replay=t[2], gather=t[3], ph_gather_us=t[4], ph_compact_us=t[5],
                    ph_score_us=t[6], ph_topk_us=t[7], longest_query_us=t[8], longest_flagged_first_us=t[9])

    def set_precision(self, mode: int, refine: int = 0):
        """0 FP32 (reference), 1 BF16 + FP32 refine, 2 FP8 + FP32 refine."""
        _check(load().fpp_set_precision(self.h, mode, refine))

    def fp8_shared_scale(self) -> float:
        """Scale every FP8 row shares (0.0 when rows keep their own)."""
        s = C.c_float(0.0)
        _check(load().fpp_get_fp8_scale(self.h, C.byref(s)))
        return float(s.value)

    def set_tuning(self, which: int, value: int):
        _check(load().fpp_set_tuning(self.h, which, value))

    def last_probes(self, nq: int) -> np.ndarray:
        out = np.zeros((nq, self.qprobes), np.uint32)
        _check(load().fpp_get_last_probes(self.h, _ptr(out)))
        return out

    def query_times(self, nq: int):
        """(service_us, done_us) per query of the last level-2 profiled call."""
        a = np.zeros(nq, np.float64)
        b = np.zeros(nq, np.float64)
        _check(load().fpp_get_query_times(self.h, _ptr(a), _ptr(b)))
        return a, b

    def last_probes_batch(self, qprobes) -> list:
        """Probes of the last batched call, one array per query."""
        qp = np.asarray(qprobes, np.int64)
        flat = np.zeros(max(int(qp.sum()), 1), np.uint32)
        _check(load().fpp_get_last_probes(self.h, _ptr(flat)))
        return np.split(flat[:int(qp.sum())], np.cumsum(qp)[:-1])

    def last_candidates(self, j: int) -> np.ndarray:
        n = load().fpp_get_last_candidates(self.h, j, None, 0)
        if n < 0:
            _check(int(-n))
        out = np.zeros(max(n, 1), np.int32)
        load().fpp_get_last_candidates(self.h, j, _ptr(out), n)
        return out[:n]


def bruteforce_f64_device(dX_ptr: int, ldx: int, N: int, D: int, dQ_ptr: int, ldq: int, nq: int, k: int,
                          seed_ids: np.ndarray | None = None):
    """Exact FP64 top-k on the GPU (verification only)."""
    ids = np.zeros((nq, k), np.int32)
    sc = np.zeros((nq, k), np.float64)
    seeds = None if seed_ids is None else np.ascontiguousarray(seed_ids, np.int32)
    _check(load().fpp_bruteforce_f64(C.c_void_p(dX_ptr), ldx, N, D, C.c_void_p(dQ_ptr), ldq, nq, k,
                                     None if seeds is None else _ptr(seeds), _ptr(ids), _ptr(sc)))
    return ids, sc


def recall(truth: np.ndarray, answer: np.ndarray, k: int) -> float:
    """Mean over queries of |truth[:k] & answer| / k (test/falconnpp.py:181-186)."""
    s = 0.0
    for t, a in zip(truth, answer):
        s += len(set(t[:k].tolist()) & set(a.tolist())) / k
    return s / max(len(truth), 1)
