"""Seeded synthetic workloads (numpy only; identical bytes on every host).

Shapes follow the benchmark configurations of this engine:
  * ``lognormal_sphere``: x = rho * z / |z|, z ~ N(0, I_d), rho ~ LogNormal(0, s)
  * ``gaussian_mixture``: y = c_j + sigma * eps with centres c ~ N(0, I_d),
    x = rho * y / |y|, rho ~ LogNormal(0, s)
Rows are points (N x d, C-contiguous float32), i.e. the reference's D x N
column-major layout (src/FalconnPP.h:30) viewed from numpy.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

DATA_SEED = 20240601
QUERY_SEED = 20240602
INDEX_SEED = 42


def _normalise_rows(y: np.ndarray, rho: np.ndarray | None) -> np.ndarray:
    nrm = np.sqrt(np.einsum("ij,ij->i", y, y, dtype=np.float64)).astype(np.float32)
    nrm[nrm == 0] = 1.0
    y /= nrm[:, None]
    if rho is not None:
        y *= rho[:, None]
    return y


# Rows are generated in chunks of CHUNK; chunk c draws from its own stream
# default_rng([seed, c]), so the bytes do not depend on how many threads run
# the chunks.
CHUNK = 1 << 18


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def _chunked(n: int, d: int, fill) -> np.ndarray:
    out = np.empty((n, d), np.float32)
    starts = list(range(0, n, CHUNK))

    def run(ci):
        s = starts[ci]
        fill(ci, out[s:min(n, s + CHUNK)])

    if len(starts) <= 1:
        for ci in range(len(starts)):
            run(ci)
    else:
        with ThreadPoolExecutor(_threads()) as ex:
            list(ex.map(run, range(len(starts))))
    return out


def lognormal_sphere(n: int, d: int, sigma: float = 0.3, seed: int = DATA_SEED, unit: bool = False) -> np.ndarray:
    def fill(ci, o):
        rng = np.random.default_rng([seed, ci])
        z = rng.standard_normal(o.shape, dtype=np.float32)
        rho = None if unit else rng.lognormal(0.0, sigma, len(o)).astype(np.float32)
        o[:] = _normalise_rows(z, rho)

    return _chunked(n, d, fill)


def mixture_centres(n_centres: int, d: int, seed: int = DATA_SEED) -> np.ndarray:
    rng = np.random.default_rng(seed + 7919)
    return rng.standard_normal((n_centres, d), dtype=np.float32)


def gaussian_mixture(n: int, d: int, n_centres: int, within: float = 0.3, norm_sigma: float = 0.25,
                     seed: int = DATA_SEED, centres: np.ndarray | None = None) -> np.ndarray:
    if centres is None:
        centres = mixture_centres(n_centres, d)

    def fill(ci, o):
        rng = np.random.default_rng([seed, ci])
        j = rng.integers(0, centres.shape[0], len(o))
        rng.standard_normal(o.shape, dtype=np.float32, out=o)
        o *= np.float32(within)
        o += centres[j]
        rho = rng.lognormal(0.0, norm_sigma, len(o)).astype(np.float32)
        _normalise_rows(o, rho)

    return _chunked(n, d, fill)


# name -> (generator kwargs, index params, query params)
CONFIGS = {
    # Plain Falconn++ on the CPU reference: N=1e5 x d=100, unit Gaussian queries.
    "c1": dict(n=100_000, d=100, kind="sphere", n_centres=0, nq=1000, L=50, n_proj=128, iProbes=1,
               alpha=0.01, bucketLimit=20, qProbes=200, k=20),
    # Plain FP32 GPU drop-in: N=1e6 x d=200 mixture of 1000 centres.
    "c2": dict(n=1_000_000, d=200, kind="mixture", n_centres=1000, nq=10_000, L=100, n_proj=256, iProbes=3,
               alpha=0.005, bucketLimit=20, qProbes=4000, k=20),
    # Batched serving: the N=1e7 x d=128 index, 50k queries, each with its own
    # qProbes (uniform in [100, 8000]) and k (one of 1, 10, 100).
    "c4": dict(n=10_000_000, d=128, kind="mixture", n_centres=5000, nq=50_000, L=64, n_proj=128, iProbes=2,
               alpha=0.01, bucketLimit=20, qProbes=2000, k=100, qp_range=(100, 8000), k_choices=(1, 10, 100)),
    # Precision variant: N=1e7 x d=128, 5000 centres.
    "c3": dict(n=10_000_000, d=128, kind="mixture", n_centres=5000, nq=10_000, L=64, n_proj=128, iProbes=2,
               alpha=0.01, bucketLimit=20, qProbes=2000, k=20),
    # Largest single-GPU case: N=5e7 x d=96 (fhtDim 128), 20000 centres,
    # 6.4e9 (id, table) entries before the bucket filter; FP8 scoring keeps
    # 200 survivors for the FP32 refine.
    "c5": dict(n=50_000_000, d=96, kind="mixture", n_centres=20_000, nq=100_000, L=64, n_proj=128, iProbes=2,
               alpha=0.01, bucketLimit=20, qProbes=4000, k=20, refine=200, fp8_per_row=True),
}


def make_config(name: str, scale: float = 1.0):
    """Data and queries of a named configuration; ``scale`` shrinks N and nq."""
    cfg = dict(CONFIGS[name])
    n = max(1000, int(cfg["n"] * scale))
    nq = max(10, int(cfg["nq"] * scale)) if scale < 1 else cfg["nq"]
    d = cfg["d"]
    if cfg["kind"] == "sphere":
        X = lognormal_sphere(n, d, 0.3, DATA_SEED)
        Q = lognormal_sphere(nq, d, 0.3, QUERY_SEED, unit=True)
    else:
        C = mixture_centres(cfg["n_centres"], d)
        X = gaussian_mixture(n, d, cfg["n_centres"], 0.3, 0.25, DATA_SEED, C)
        Q = gaussian_mixture(nq, d, cfg["n_centres"], 0.3, 0.25, QUERY_SEED, C)
    cfg["n"], cfg["nq"] = n, nq
    return X, Q, cfg


def per_query_params(cfg: dict, nq: int, seed: int = QUERY_SEED + 1):
    """Per-query qProbes (uniform integers in cfg["qp_range"]) and k (uniform
    over cfg["k_choices"]) of a batched-serving configuration."""
    rng = np.random.default_rng(seed)
    lo, hi = cfg["qp_range"]
    qps = rng.integers(lo, hi + 1, nq).astype(np.int32)
    ks = rng.choice(np.asarray(cfg["k_choices"], np.int32), nq).astype(np.int32)
    return qps, ks
