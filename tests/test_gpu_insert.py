"""Streaming insert against the oracle (needs an MI355X).

An index built on the first points and then grown by fpp_insert must hold,
bucket for bucket, what the oracle's build2Layers_1D (src/FalconnPP.cpp:21-303)
of all points holds with the same sign bits and the centre of the first
build: the alpha / bucketLimit rule is applied again over old and new
entries, in point order, so ties resolve as in one build.  Queries on the
grown index then match the oracle's (probes and candidates bit-exact, top-k
within 1e-5)."""
import os

import numpy as np
import pytest

from falconnpp_amd import engine as E
from falconnpp_amd import synth
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu


def _int_grid(n_sym, n_more, d, seed):
    """n_sym integer points with zero mean (each with its negative), then
    n_more other integer points: the first build's centre is exactly 0, so
    every projection stays an integer and exact ties are everywhere."""
    rng = np.random.default_rng(seed)
    A = rng.integers(-2, 3, (n_sym // 2, d)).astype(np.float32)
    S = np.concatenate([A, -A])[rng.permutation(n_sym // 2 * 2)]
    return np.concatenate([S, rng.integers(-2, 3, (n_more, d)).astype(np.float32)])


CASES = {
    # name: (data, first build, insert batches, L, n_proj, bucketLimit, alpha, iProbes, qProbes, k)
    "mixture": (lambda: synth.gaussian_mixture(20000, 64, 50, seed=51), 12000, (5000, 3000), 10, 64, 10, 0.1, 2,
                100, 10),
    # integer data: exact ties in hashing and in the bucket filter
    "int_ties": (lambda: _int_grid(2500, 1001, 16, 52), 2500, (1000, 1), 6, 16, 5, 0.3, 2, 40, 10),
}


def _flat(off, ids):
    size = np.diff(off.astype(np.int64))
    return size, np.repeat(np.arange(len(size)), size), ids


def _oracle_flat(o):
    pos, size, oids = o.buckets()
    size = size.astype(np.int64)
    starts = np.cumsum(size) - size
    flat = np.repeat(pos.astype(np.int64), size) + (np.arange(size.sum()) - np.repeat(starts, size))
    ob = np.repeat(np.arange(len(size)), size)
    ov = oids[flat]
    order = np.lexsort((ov, ob))
    return size, ob[order], ov[order]


@pytest.mark.parametrize("name", list(CASES))
def test_insert_equals_build_of_all_points(name):
    data, n_a, batches, L, n_proj, blim, alpha, p, qp, k = CASES[name]
    X = data()
    n, d = X.shape
    assert n_a + sum(batches) == n
    print(f"[insert] case={name} n={n} first={n_a} batches={batches}", flush=True)
    g = E.Index(n_a, d)
    g.setIndexParam(L, n_proj, blim, alpha, p, 8, 42)
    g.set_insertable(True)
    g.build(X[:n_a])
    center = g.center()
    s = n_a
    for m in batches:
        g.insert(X[s:s + m])
        s += m
    assert g.shape()["n_points"] == n and g.n == n
    assert np.array_equal(g.center(), center)
    if name == "int_ties":
        assert np.all(center == 0)
    cnt = g.bucket_counts()
    assert int(cnt.astype(np.int64).sum()) == n * L * p
    o = P.OracleIndex(n, d)
    o.setIndexParam(L, n_proj, blim, alpha, p, 8, 42)
    o.build(X, center)
    off, gids = g.buckets()
    gsize, gb, gv = _flat(off, gids)
    osize, ob, ov = _oracle_flat(o)
    assert np.array_equal(gsize, osize)
    assert np.array_equal(gb, ob) and np.array_equal(gv, ov)
    # queries on the grown index
    rng = np.random.default_rng(3)
    Q = X[rng.choice(n, 30, replace=False)] + np.float32(0.05) * rng.standard_normal((30, d)).astype(np.float32)
    if name == "int_ties":
        Q = np.round(Q)
    Q = Q.astype(np.float32)
    g.set_tuning(E.TUNE_KEEP_CANDIDATES, 1)
    g.set_qProbes(qp)
    o.set_qProbes(qp)
    out = g.query(Q, k)
    probes = g.last_probes(len(Q))
    for j, q in enumerate(Q):
        dbg = o.query_debug(q, k)
        assert np.array_equal(np.sort(probes[j]), np.sort(dbg["probes"].astype(np.uint32))), j
        assert np.array_equal(g.last_candidates(j), dbg["cands"]), j
        m = min(k, len(dbg["cands"]))
        sg = (X[out[j, :m]] - center).astype(np.float64) @ q.astype(np.float64)
        so = dbg["scores"][:m].astype(np.float64)
        assert np.all(np.abs(sg - so) <= 1e-5 * np.maximum(np.abs(so), 1e-6) + 1e-6), j
    # the reduced-precision copy is made again over all points: with every
    # candidate refined it answers like FP32
    g.set_qProbes(6)
    ref = g.query(Q, k)
    if max(len(g.last_candidates(j)) for j in range(len(Q))) <= 256:
        g.set_precision(E.PREC_BF16, 256)
        try:
            assert np.array_equal(g.query(Q, k), ref)
        finally:
            g.set_precision(E.PREC_FP32, 0)


def test_insert_needs_an_insertable_build_and_pybind_surface():
    import sys
    X = synth.gaussian_mixture(3000, 32, 20, seed=53)
    g = E.Index(2000, 32)
    g.setIndexParam(4, 32, 5, 0.1, 1, 8, 7)
    with pytest.raises(E.FppError) as e:
        g.insert(X[2000:])
    assert e.value.code == E.FPP_ESTATE
    g.build(X[:2000])
    with pytest.raises(E.FppError) as e:
        g.insert(X[2000:])
    assert e.value.code == E.FPP_ESTATE
    g.set_insertable(True)
    g.build(X[:2000])
    g.insert(X[2000:])
    g.set_qProbes(40)
    ref = g.query(X[:10], 5)
    sys.path.insert(0, os.path.dirname(E.LIB_PATH))
    import FalconnPP
    f = FalconnPP.FalconnPP(2000, 32)
    f.setIndexParam(4, 32, 5, 0.1, 1, 8, 7)
    f.set_insertable(True)
    f.build(np.asfortranarray(X[:2000].T))
    f.insert(np.asfortranarray(X[2000:].T))
    f.set_qProbes(40)
    assert np.array_equal(f.query(np.asfortranarray(X[:10].T), 5), ref)
    with pytest.raises(ValueError):
        f.insert(np.zeros((31, 4), np.float32))
