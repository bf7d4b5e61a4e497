"""Oracle parity at the sizes of the benchmark workloads (needs an MI355X).

* N=1e5 x d=100, 50 tables, qProbes 200: every table's buckets bit-exact;
  probes and candidates of all 1000 queries bit-exact; top-k within 1e-5.
* N=1e6 x d=200 (100 tables) and N=1e7 x d=128 (64 tables): a table's sign
  bits are the same prefix of one bit stream whatever n_tables is
  (src/FalconnPP.h:52-75), so the oracle builds only the first tables and
  those must equal the engine's first tables bit for bit.  The oracle then
  adopts the engine's whole index and answers a sample of the queries:
  probes and candidates bit-exact, top-k within 1e-5.
* N=1e7, 50,000 queries each with its own qProbes (100..8000) and k (1, 10,
  100) in one batched call: a sample of them, each against the oracle at its
  own (qProbes, k).
* N=1e7: BF16 / FP8 + FP32 refine of the top 80, recall@20 against the FP64
  brute force (src/BF.cpp:17-87 in FP64) and exact scores of the answers.
* N=5e7 x d=96, 64 tables: full-size properties of the index (every point in
  iProbes buckets per table before the filter, 6.4e9 entries in all; the
  iLimit rule per bucket; ids unique per bucket and in range) and a sample of
  queries against the oracle; FP8 with one scale per row, recall against
  FP64.
"""

import numpy as np
import pytest

from falconnpp_amd import engine as E
from falconnpp_amd import synth
from oracle import pyoracle as P

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

SCORE_RTOL = 1e-5
THREADS = 16


def _log(msg):
    print(f"[configs] {msg}", flush=True)


def _oracle(n, d, c, L=None):
    o = P.OracleIndex(n, d)
    o.setIndexParam(L or c["L"], c["n_proj"], c["bucketLimit"], c["alpha"], c["iProbes"], THREADS, synth.INDEX_SEED)
    return o


def _engine(X, c):
    n, d = X.shape
    g = E.Index(n, d)
    g.setIndexParam(c["L"], c["n_proj"], c["bucketLimit"], c["alpha"], c["iProbes"], THREADS, synth.INDEX_SEED)
    g.build(X)
    g.set_tuning(E.TUNE_KEEP_CANDIDATES, 1)
    return g


def _oracle_flat(o):
    """Oracle buckets as (bucket of each id, ids ascending inside a bucket)."""
    pos, size, oids = o.buckets()
    size = size.astype(np.int64)
    starts = np.cumsum(size) - size
    flat = np.repeat(pos.astype(np.int64), size) + (np.arange(size.sum()) - np.repeat(starts, size))
    ob = np.repeat(np.arange(len(size)), size)
    ov = oids[flat]
    order = np.lexsort((ov, ob))
    return size, ob[order], ov[order]


def _assert_prefix_tables_equal(g, o, tables):
    off, gids = g.buckets()
    nb = g.shape()["buckets_per_table"]
    m = tables * nb
    gsize = np.diff(off[:m + 1].astype(np.int64))
    osize, ob, ov = _oracle_flat(o)
    assert len(osize) == m
    assert np.array_equal(gsize, osize)
    gb = np.repeat(np.arange(m), gsize)
    assert np.array_equal(gb, ob)
    assert np.array_equal(gids[:int(off[m])], ov)


def _check_query(g, o, X, center, q, j, probes_j, k, out_j):
    """Probes and candidates bit-exact; at every rank the engine's id scores
    within 1e-5 of the oracle's score there, so ids differ only among scores
    tied within that tolerance."""
    d = o.query_debug(q, k)
    assert np.array_equal(np.sort(probes_j), np.sort(d["probes"].astype(np.uint32))), j
    assert np.array_equal(g.last_candidates(j), d["cands"]), j
    m = min(k, len(d["cands"]))
    sg = (X[out_j[:m]] - center).astype(np.float64) @ q.astype(np.float64)
    so = d["scores"][:m].astype(np.float64)
    tol = SCORE_RTOL * np.maximum(np.abs(so), 1e-6) + 1e-6
    assert np.all(np.abs(sg - so) <= tol), (j, sg, so)
    assert len(set(out_j[:m].tolist())) == m
    assert np.all(out_j[m:] == 0)


def test_c1_full_size_parity():
    X, Q, c = synth.make_config("c1")
    n, d = X.shape
    _log(f"c1 n={n} d={d} L={c['L']} qProbes={c['qProbes']} nq={len(Q)}")
    g = _engine(X, c)
    o = _oracle(n, d, c)
    o.build(X, g.center())
    _assert_prefix_tables_equal(g, o, c["L"])
    k = c["k"]
    g.set_qProbes(c["qProbes"])
    o.set_qProbes(c["qProbes"])
    out = g.query(Q, k)
    probes = g.last_probes(len(Q))
    center = g.center()
    for j in range(len(Q)):
        _check_query(g, o, X, center, Q[j], j, probes[j], k, out[j])


@pytest.mark.parametrize("name,prefix,sample", [("c2", 4, 200)])
def test_prefix_tables_and_query_sample(name, prefix, sample):
    X, Q, c = synth.make_config(name)
    n, d = X.shape
    _log(f"{name} n={n} d={d} L={c['L']}: engine build, oracle build of {prefix} tables")
    g = _engine(X, c)
    center = g.center()
    o = _oracle(n, d, c, L=prefix)
    o.build(X, center)
    _assert_prefix_tables_equal(g, o, prefix)
    del o
    off, ids = g.buckets()
    o = _oracle(n, d, c)
    o.load_index(X, center, off, ids)
    del off, ids
    k = c["k"]
    Qs = Q[:sample]
    g.set_qProbes(c["qProbes"])
    o.set_qProbes(c["qProbes"])
    out = g.query(Qs, k)
    probes = g.last_probes(len(Qs))
    _log(f"{name}: {sample} queries against the oracle")
    for j in range(len(Qs)):
        _check_query(g, o, X, center, Qs[j], j, probes[j], k, out[j])


@pytest.fixture(scope="module")
def c3():
    X, Q, c = synth.make_config("c3")
    n, d = X.shape
    _log(f"c3 n={n} d={d} L={c['L']}: engine build")
    g = _engine(X, c)
    center = g.center()
    off, ids = g.buckets()
    o = _oracle(n, d, c)
    o.load_index(X, center, off, ids)
    del off, ids
    yield X, Q, c, g, o, center
    g.close()


def test_c3_prefix_tables(c3):
    X, Q, c, g, o, center = c3
    n, d = X.shape
    _log("c3: oracle build of 2 tables")
    ob = _oracle(n, d, c, L=2)
    ob.build(X, center)
    _assert_prefix_tables_equal(g, ob, 2)


def test_c3_query_sample(c3):
    X, Q, c, g, o, center = c3
    k, sample = c["k"], 200
    Qs = Q[:sample]
    g.set_qProbes(c["qProbes"])
    o.set_qProbes(c["qProbes"])
    out = g.query(Qs, k)
    probes = g.last_probes(len(Qs))
    _log(f"c3: {sample} queries against the oracle")
    for j in range(len(Qs)):
        _check_query(g, o, X, center, Qs[j], j, probes[j], k, out[j])


def _truth(X, Q, k, seeds):
    import torch
    dX = torch.from_numpy(X).cuda()
    dQ = torch.from_numpy(np.ascontiguousarray(Q)).cuda()
    ids, sc = E.bruteforce_f64_device(dX.data_ptr(), X.shape[1], X.shape[0], X.shape[1], dQ.data_ptr(),
                                      Q.shape[1], len(Q), k, seeds)
    del dX, dQ
    torch.cuda.empty_cache()
    return ids, sc


def _lowprec_answer(g, Q, k, mode, refine):
    import torch
    dQ = torch.from_numpy(np.ascontiguousarray(Q)).cuda()
    out = torch.zeros((len(Q), k), dtype=torch.int32, device="cuda")
    sc = torch.zeros((len(Q), k), dtype=torch.float32, device="cuda")
    try:
        g.set_precision(mode, refine)
        g.query_device(dQ.data_ptr(), Q.shape[1], len(Q), k, out.data_ptr(), sc.data_ptr())
        torch.cuda.synchronize()
    finally:
        g.set_precision(E.PREC_FP32, 0)
    return out.cpu().numpy(), sc.cpu().numpy()


def _check_exact_scores(X, center, Q, ids, scores, k):
    Xc_rows = lambda r: (X[r] - center).astype(np.float64)  # noqa: E731
    for j in range(len(Q)):
        exact = Xc_rows(ids[j]) @ Q[j].astype(np.float64)
        np.testing.assert_allclose(scores[j], exact, rtol=1e-5, atol=1e-5)
        assert np.all(np.diff(scores[j]) <= 0)


def test_c3_precision_recall_against_fp64(c3):
    X, Q, c, g, o, center = c3
    k, sample, refine = c["k"], 1000, 80
    Qs = Q[:sample]
    g.set_qProbes(c["qProbes"])
    ref = g.query(Qs, k)
    truth, _ = _truth(X, Qs, k, ref)
    r32 = E.recall(truth, ref, k)
    _log(f"c3: fp32 recall@{k} = {r32:.4f}")
    assert r32 >= 0.95
    for mode in (E.PREC_BF16, E.PREC_FP8):
        ids, sc = _lowprec_answer(g, Qs, k, mode, refine)
        r = E.recall(truth, ids, k)
        _log(f"c3: precision {mode} refine {refine}: recall@{k} = {r:.4f}")
        assert r >= 0.95 and r >= r32 - 0.02
        _check_exact_scores(X, center, Qs, ids, sc, k)


def test_c4_mixed_batch_sample(c3):
    import torch
    X, _, c3cfg, g, o, center = c3
    _, Q, c = synth.make_config("c4")
    nq = len(Q)
    qps, ks = synth.per_query_params(c, nq)
    kmax = int(max(c["k_choices"]))
    dQ = torch.from_numpy(np.ascontiguousarray(Q)).cuda()
    out = torch.zeros((nq, kmax), dtype=torch.int32, device="cuda")
    sc = torch.zeros((nq, kmax), dtype=torch.float32, device="cuda")
    _log(f"c4: batch of {nq} queries, qProbes {qps.min()}..{qps.max()}, k in {sorted(set(ks.tolist()))}")
    g.query_batch_device(dQ.data_ptr(), Q.shape[1], nq, qps, ks, kmax, out.data_ptr(), sc.data_ptr())
    torch.cuda.synchronize()
    out = out.cpu().numpy()
    probes = g.last_probes_batch(qps)
    rng = np.random.default_rng(11)
    sel = np.sort(rng.choice(nq, 200, replace=False))
    _log("c4: 200 queries against the oracle at their own (qProbes, k)")
    for j in sel:
        o.set_qProbes(int(qps[j]))
        _check_query(g, o, X, center, Q[j], int(j), probes[j], int(ks[j]), out[j, :ks[j]])
        assert np.all(out[j, ks[j]:] == 0)


def _ilimit(B, alpha, p, blim):
    """src/FalconnPP.cpp:233-242 in float arithmetic: ceil(alpha * B / p), raised
    to min(bucketLimit, B) below bucketLimit, at most B."""
    Bf = B.astype(np.float32)
    lim = np.ceil((np.float32(alpha) * Bf) / np.float32(p)).astype(np.int64)
    B64 = B.astype(np.int64)
    lim = np.where(lim < blim, np.minimum(blim, B64), lim)
    return np.minimum(lim, B64)


def test_c5_full_size_properties_and_sample():
    X, Q, c = synth.make_config("c5")
    n, d = X.shape
    L, p = c["L"], c["iProbes"]
    _log(f"c5 n={n} d={d} L={L}: engine build")
    g = _engine(X, c)
    _log("c5: built; reading the tables")
    cnt = g.bucket_counts()
    off, ids = g.buckets()
    nb = g.shape()["buckets_per_table"]
    _log(f"c5: {len(ids)} ids; checking sizes and the iLimit rule")
    # every point lands in iProbes buckets per table before the filter
    per_table = cnt.reshape(L, nb).astype(np.int64).sum(1)
    assert np.all(per_table == n * p)
    assert int(cnt.astype(np.int64).sum()) == n * L * p and n * L * p > 2 ** 32
    size = np.diff(off.astype(np.int64))
    assert np.array_equal(size, _ilimit(cnt, c["alpha"], p, c["bucketLimit"]))
    assert int(off[-1]) == len(ids)
    assert ids.min() >= 0 and ids.max() < n
    # ascending, hence unique, inside every bucket
    inside = np.ones(len(ids), bool)
    inside[off[:-1][size > 0].astype(np.int64)] = False
    assert np.all(np.diff(ids.astype(np.int64))[inside[1:]] > 0)
    # a point is kept at most iProbes times per table
    _log("c5: ids per table")
    for t in range(L):
        a, b = int(off[t * nb]), int(off[(t + 1) * nb])
        srt = np.sort(ids[a:b])
        assert len(srt) <= p or np.all(srt[p:] != srt[:-p]), t
    _log(f"c5: {len(ids)} ids kept of {n * L * p}; oracle adopts the index")
    center = g.center()
    o = _oracle(n, d, c)
    o.load_index(X, center, off, ids)
    del off, ids, cnt
    _log("c5: sample queries against the oracle")
    k, sample = c["k"], 40
    Qs = Q[:sample]
    g.set_qProbes(c["qProbes"])
    o.set_qProbes(c["qProbes"])
    out = g.query(Qs, k)
    probes = g.last_probes(len(Qs))
    for j in range(len(Qs)):
        _check_query(g, o, X, center, Qs[j], j, probes[j], k, out[j])
    del o
    # FP8 with one scale per row (the forced per-vector scale), refine 200
    _log("c5: FP64 truth, FP8 with per-row scales")
    truth, _ = _truth(X, Qs, k, out)
    g.set_tuning(E.TUNE_FP8_PER_ROW, 1)
    try:
        ids8, sc8 = _lowprec_answer(g, Qs, k, E.PREC_FP8, c["refine"])
        assert g.fp8_shared_scale() == 0.0
    finally:
        g.set_tuning(E.TUNE_FP8_PER_ROW, 0)
    r32, r8 = E.recall(truth, out, k), E.recall(truth, ids8, k)
    _log(f"c5: recall@{k} fp32 {r32:.4f}, fp8 per-row scales + refine {c['refine']} {r8:.4f}")
    assert r8 >= 0.95 and r8 >= r32 - 0.02
    _check_exact_scores(X, center, Qs, ids8, sc8, k)
    g.close()
