"""GPU parity of the HIP engine against the CPU oracle (needs an MI355X).

Bit-exact: sign bits, hash codes and values, bucket contents (after the
alpha / bucketLimit filter), probe sets and candidate sets.  Rerank: top-k ids
equal except where scores tie within 1e-5 relative; scores within 1e-5
relative (the reference's Eigen dot order is not reproducible).  The centre
is an engine output (fixed-order FP64 sum) fed to the oracle; it is checked
against the float64 mean separately.
"""
import math

import numpy as np
import pytest

from falconnpp_amd import engine as E
from falconnpp_amd import synth
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu

SCORE_RTOL = 1e-5


def _sym_int_grid(n, d, lo, hi, seed):
    """Zero-mean integer data: every projection is an integer, so exact float
    ties are everywhere and the heap-replay paths run."""
    rng = np.random.default_rng(seed)
    A = rng.integers(lo, hi + 1, (n // 2, d)).astype(np.float32)
    return np.concatenate([A, -A])[rng.permutation(n // 2 * 2)]


def _far_block(X, lo, hi, seed):
    """Rows lo..hi replaced by a tight cluster around one direction: queries
    drawn from the other rows find no candidate in that id range, queries
    drawn from it find tens of thousands there."""
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(X.shape[1]).astype(np.float32)
    u /= np.linalg.norm(u)
    X[lo:hi] = u + np.float32(0.05) * rng.standard_normal((hi - lo, X.shape[1])).astype(np.float32)
    return X


CASES = {
    # name: (data, L, n_proj, bucketLimit, alpha, iProbes, qProbes, k, nq)
    "sphere_d100_p1": (lambda: synth.lognormal_sphere(20000, 100, 0.3, seed=11), 10, 128, 20, 0.01, 1, 100, 20, 60),
    "mix_d200_p3": (lambda: synth.gaussian_mixture(6000, 200, 40, seed=12), 6, 256, 20, 0.005, 3, 300, 20, 40),
    "mix_d128_p2": (lambda: synth.gaussian_mixture(12000, 128, 60, seed=13), 8, 128, 10, 0.2, 2, 160, 10, 50),
    "nproj16_d40": (lambda: synth.lognormal_sphere(5000, 40, 0.3, seed=14), 5, 16, 8, 0.05, 2, 60, 15, 40),
    "int_grid_ties": (lambda: _sym_int_grid(3000, 24, -2, 2, 15), 4, 32, 5, 0.3, 2, 40, 10, 40),
    # 200 tables: the exact probe merge keeps its heap in 4 register slots per lane
    "int_ties_l200": (lambda: _sym_int_grid(2000, 16, -2, 2, 17), 200, 16, 5, 0.3, 2, 300, 10, 20),
    # N above the LDS bitset window (~1.2M ids): the dedup runs over two windows
    "multiwindow_n1.3m": (lambda: synth.gaussian_mixture(1_300_000, 16, 300, seed=16), 2, 16, 30, 0.05, 1, 8, 20, 12),
    # more than 31 probes per layer: every query's probes come from the exact
    # replay (32 per layer; the merge stages its lists in LDS) ...
    "sphere_pl32_exact": (lambda: synth.lognormal_sphere(8000, 50, 0.3, seed=19), 10, 128, 20, 0.01, 1, 5000, 20, 30),
    # ... and with ties everywhere (40 per layer; 20 x 1600 lists of 256 KB,
    # read by the merge in place)
    "int_grid_pl40_exact": (lambda: _sym_int_grid(2000, 24, -2, 2, 18), 20, 64, 5, 0.3, 2, 16000, 10, 20),
    # 64 tables x 18^2 pairs: the pool of every pair (97.5 KB of LDS with the
    # tables' code lists) still fits and is taken
    "full_pool_l64_pl18": (lambda: synth.lognormal_sphere(20000, 64, 0.3, seed=21), 64, 64, 20, 0.01, 1, 9280, 20, 20),
    # the last step of the reference CLI's sweep (qProbes 10000) at 100 and 350
    # tables: the pool of every pair fits, the per-table lists would not
    "l100_qp10000": (lambda: synth.lognormal_sphere(4000, 32, 0.3, seed=22), 100, 32, 10, 0.1, 1, 10000, 20, 20),
    "l350_qp10000": (lambda: synth.lognormal_sphere(3000, 16, 0.3, seed=23), 350, 16, 10, 0.1, 1, 10000, 20, 20),
    # 4 probes per layer over 1800 tables: the pool of every pair (173 KB)
    # does not fit, the top-10 lists per table (144 KB) do: two-stage selection
    "two_stage_l1800_pl4": (lambda: synth.lognormal_sphere(2000, 16, 0.3, seed=24), 1800, 16, 5, 0.1, 1, 9000, 10, 10),
    # neither pool fits (192 KB and 240 KB): every query's probes come from the
    # exact replay, which reuses the fast per-layer lists
    "exact_pool_l600": (lambda: synth.lognormal_sphere(2000, 16, 0.3, seed=25), 600, 16, 5, 0.1, 1, 15000, 10, 10),
    # two dedup windows with C_tot above N (alpha = 1 keeps every id; 2000
    # probes over 16 tables of buckets holding ~1270 ids each): the scatter by
    # window holds every bucket id
    "ctot_over_n_windows": (lambda: synth.gaussian_mixture(1_300_000, 16, 300, seed=26), 16, 16, 30, 1.0, 1, 2000, 20, 6),
    # a tight cluster at ids [2^19, 2^20) of 1.6M: an empty window between a
    # query's non-empty ones, windows of one round and of tens, and a query
    # small enough for the open-addressing set, in one batch
    "windows_far_block": (lambda: _far_block(synth.gaussian_mixture(1_600_000, 16, 300, seed=27), 1 << 19, 1 << 20, 28),
                          4, 16, 30, 0.05, 1, 40, 20, 12),
}

# probe selection each case must take (engine stats: 0 pool of every pair,
# 1 per-table lists, 2 exact replay of every query)
PROBE_PATH = {"full_pool_l64_pl18": 0, "l100_qp10000": 0, "l350_qp10000": 0, "two_stage_l1800_pl4": 1,
              "exact_pool_l600": 2, "sphere_d100_p1": 0}

_cache = {}


def _built(name):
    if name in _cache:
        return _cache[name]
    data, L, n_proj, blim, alpha, p, qp, k, nq = CASES[name]
    X = data()
    n, d = X.shape
    print(f"[parity] case={name} n={n} d={d} L={L} n_proj={n_proj} p={p}", flush=True)
    g = E.Index(n, d)
    g.setIndexParam(L, n_proj, blim, alpha, p, 8, 42)
    g.build(X)
    g.set_tuning(E.TUNE_KEEP_CANDIDATES, 1)  # candidate lists read by the tests
    o = P.OracleIndex(n, d)
    o.setIndexParam(L, n_proj, blim, alpha, p, 8, 42)
    o.build(X, g.center())
    rng = np.random.default_rng(7)
    Q = X[rng.choice(n, nq, replace=False)] + np.float32(0.05) * rng.standard_normal((nq, d)).astype(np.float32)
    if name.startswith("int_grid"):
        Q = np.round(Q)
    _cache[name] = (X, Q.astype(np.float32), g, o)
    return _cache[name]


@pytest.mark.parametrize("name", list(CASES))
def test_sign_bits_and_center(name):
    X, Q, g, o = _built(name)
    b1, b2 = g.sign_bits()
    o1, o2 = P.sign_bits(42, len(b1))
    assert np.array_equal(b1, o1) and np.array_equal(b2, o2)
    c = g.center()
    m64 = X.astype(np.float64).mean(0)
    assert np.all(np.abs(c.astype(np.float64) - m64) <= np.spacing(np.abs(m64).astype(np.float32)) + 1e-30)


@pytest.mark.parametrize("name", list(CASES))
def test_hash_codes_bit_exact(name):
    X, Q, g, o = _built(name)
    Xc = X - g.center()
    m = min(len(X), 1500)
    gc, gv = g.hash_points(Xc[:m])
    oc, ov = o.hash_build(Xc[:m])
    # canonical order inside each (table, point): by value, then code
    for arr_c, arr_v in ((gc, gv), (oc, ov)):
        idx = np.lexsort((arr_c.astype(np.int64), arr_v), axis=-1)
        arr_c[:] = np.take_along_axis(arr_c, idx, -1)
        arr_v[:] = np.take_along_axis(arr_v, idx, -1)
    assert np.array_equal(gc.astype(np.int64), oc.astype(np.int64))
    assert np.array_equal(gv, ov)


def _bucket_pairs_gpu(g):
    off, ids = g.buckets()
    sizes = np.diff(off.astype(np.int64))
    b = np.repeat(np.arange(len(sizes)), sizes)
    return sizes, b, ids


@pytest.mark.parametrize("name", list(CASES))
def test_bucket_contents_bit_exact(name):
    X, Q, g, o = _built(name)
    sizes, gb, gids = _bucket_pairs_gpu(g)
    pos, osize, oids = o.buckets()
    assert np.array_equal(sizes, osize.astype(np.int64))
    # GPU stores every bucket in ascending id order
    same = gb[1:] == gb[:-1]
    assert np.all(gids[1:][same] > gids[:-1][same])
    ob = np.repeat(np.arange(len(osize)), osize.astype(np.int64))
    oflat = np.concatenate([oids[p:p + s] for p, s in zip(pos.astype(np.int64), osize.astype(np.int64)) if s]) \
        if osize.sum() else np.zeros(0, np.int32)
    ob_sorted = np.lexsort((oflat, ob))
    assert np.array_equal(gb, ob[ob_sorted])
    assert np.array_equal(gids, oflat[ob_sorted])


def _check_topk(Xc, q, gids, oids, oscores, k):
    qd = q.astype(np.float64)
    sg = Xc[gids].astype(np.float64) @ qd
    so = oscores.astype(np.float64)
    nvalid = int(np.count_nonzero(so)) if np.all(oids == 0) else len(so)
    tol = SCORE_RTOL * np.maximum(np.abs(so), 1e-6)
    assert np.all(np.abs(sg - so) <= tol + 1e-6), (sg, so)
    # ids differ only inside groups of scores tied within tolerance
    for j in range(len(oids)):
        if gids[j] != oids[j]:
            assert abs(sg[j] - so[j]) <= tol[j] + 1e-6


@pytest.mark.parametrize("name", list(CASES))
def test_query_probes_candidates_topk(name):
    X, Q, g, o = _built(name)
    qp, k = CASES[name][6], CASES[name][7]
    g.set_qProbes(qp)
    o.set_qProbes(qp)
    out = g.query(Q, k)
    probes = g.last_probes(len(Q))
    st = g.stats()
    Xc = X - g.center()
    ctot = 0
    for j, q in enumerate(Q):
        d = o.query_debug(q, k)
        assert np.array_equal(np.sort(probes[j]), np.sort(d["probes"].astype(np.uint32))), j
        cg = g.last_candidates(j)
        assert np.array_equal(cg, d["cands"]), j
        ctot += d["ctot"]
        m = min(k, len(d["cands"]))
        _check_topk(Xc, q, out[j, :m], d["topk"][:m], d["scores"][:m], k)
        assert np.all(out[j, m:] == 0)
    assert st["ctot"] == ctot
    if name in PROBE_PATH:
        assert st["probe_path"] == PROBE_PATH[name]
    if name == "ctot_over_n_windows":
        assert all(o.query_debug(q, k)["ctot"] > len(X) for q in Q[:2])
    if name == "windows_far_block":
        # some query skips a window of 2^19 ids (the FP32 gather's window at
        # two workgroups per CU) between its first and last candidate
        def gap(c):
            w = np.unique(c.astype(np.int64) >> 19)
            return len(w) > 0 and w[-1] - w[0] + 1 > len(w)
        assert any(gap(g.last_candidates(j)) for j in range(len(Q)))


@pytest.mark.parametrize("k", [10, 64, 65, 200])
def test_topk_ties_break_to_smaller_id(k):
    """Integer data and integer queries: every score is an exact integer, so
    equal scores are common; both top-k paths (wave lists for k <= 64, radix
    select beyond) must keep the smaller ids among equal scores."""
    X, _, g, _ = _built("int_grid_ties")
    assert np.all(g.center() == 0)
    Q = X[:16].copy()
    g.set_qProbes(40)
    out = g.query(Q, k)
    for j, q in enumerate(Q):
        c = g.last_candidates(j)
        s = X[c].astype(np.float64) @ q.astype(np.float64)
        order = np.lexsort((c, -s))
        m = min(k, len(c))
        assert np.array_equal(out[j, :m], c[order[:m]]), j
        assert np.all(out[j, m:] == 0)


def test_tie_replay_paths_exercised():
    X, Q, g, o = _built("int_grid_ties")
    st = g.stats()
    assert st["bucket_ties"] > 0
    g.set_qProbes(40)
    g.query(Q, 10)
    assert g.stats()["query_ties"] > 0
    Xc = X - g.center()
    g.hash_points(Xc[:500])
    assert g.stats()["hash_ties"] > 0


def test_zero_fill_and_empty_batch():
    X, Q, g, o = _built("nproj16_d40")
    g.set_qProbes(1)
    o.set_qProbes(1)
    out = g.query(Q[:5], 300)
    for j in range(5):
        d = o.query_debug(Q[j], 300)
        m = len(d["cands"])
        assert m < 300
        assert np.all(out[j, m:] == 0)
        assert set(out[j, :m].tolist()) == set(d["cands"].tolist())
    assert g.query(Q[:0], 5).shape == (0, 5)


def test_error_paths():
    g = E.Index(100, 8)
    with pytest.raises(E.FppError) as e:
        g.query(np.zeros((2, 8), np.float32), 5)
    assert e.value.code == E.FPP_ESTATE
    g.setIndexParam(2, 6, 5, 0.1, 1, 1, 1)  # n_proj not a power of two
    with pytest.raises(E.FppError) as e:
        g.build(np.zeros((100, 8), np.float32))
    assert e.value.code == E.FPP_EINVAL
    g.setIndexParam(2, 4, 5, 0.1, 1, 1, 1)
    g.build(np.random.default_rng(0).standard_normal((100, 8)).astype(np.float32))
    g.set_qProbes(1000)  # perLayer > n_proj
    with pytest.raises(E.FppError) as e:
        g.query(np.zeros((2, 8), np.float32), 5)
    assert e.value.code == E.FPP_EINVAL
    g.set_qProbes(2)
    g.clear()
    with pytest.raises(E.FppError) as e:
        g.query(np.zeros((2, 8), np.float32), 5)
    assert e.value.code == E.FPP_ESTATE


def test_pybind_surface_matches_c_abi():
    import falconnpp_amd
    FPP = falconnpp_amd.FalconnPP
    X, Q, g, o = _built("sphere_d100_p1")
    n, d = X.shape
    idx = FPP.FalconnPP(n, d)
    idx.setIndexParam(10, 128, 20, 0.01, 1, 8, 42)
    idx.build(np.transpose(X))  # D x N view, F-contiguous: used in place
    idx.set_qProbes(100)
    ans = idx.query(np.transpose(Q), 20)
    assert ans.shape == (len(Q), 20) and ans.dtype == np.int32 and ans.flags.f_contiguous
    g.set_qProbes(100)
    assert np.array_equal(ans, g.query(Q, 20))
    idx.build2D(np.transpose(X))
    assert np.array_equal(idx.query2D(np.transpose(Q), 20), ans)
    with pytest.raises(ValueError):
        idx.query(np.zeros((d + 1, 3), np.float32), 5)


def test_bruteforce_f64_matches_oracle():
    import torch
    X, Q, g, o = _built("mix_d128_p2")
    k = 10
    dX = torch.from_numpy(X).cuda()
    dQ = torch.from_numpy(Q).cuda()
    g.set_qProbes(160)
    seeds = g.query(Q, k)
    ids, sc = E.bruteforce_f64_device(dX.data_ptr(), X.shape[1], X.shape[0], X.shape[1], dQ.data_ptr(),
                                      Q.shape[1], len(Q), k, seeds)
    oi, os_ = P.bruteforce_f64(X, Q, k)
    np.testing.assert_allclose(sc, os_, rtol=1e-12, atol=1e-12)
    assert np.array_equal(ids, oi)
    r = E.recall(ids, seeds, k)
    assert 0.0 <= r <= 1.0


def test_query_order_and_launch_shape_do_not_change_results():
    X, Q, g, o = _built("mix_d200_p3")
    g.set_qProbes(300)
    ref = g.query(Q, 20)
    g.set_tuning(E.TUNE_QUERY_ORDER, 0)
    try:
        assert np.array_equal(g.query(Q, 20), ref)
    finally:
        g.set_tuning(E.TUNE_QUERY_ORDER, 1)
    for t in (256, 1024):
        g.set_tuning(E.TUNE_GATHER_THREADS, t)
        assert np.array_equal(g.query(Q, 20), ref), t
    g.set_tuning(E.TUNE_GATHER_THREADS, 512)
    g.set_profiling(2)
    assert np.array_equal(g.query(Q, 20), ref)
    t = g.timings()
    assert t["gather"] > 0 and t["ph_score_us"] > 0
    g.set_profiling(0)


@pytest.mark.parametrize("mode", [E.PREC_BF16, E.PREC_FP8])
def test_low_precision_refine_equals_fp32_when_every_candidate_is_refined(mode):
    # few probes: every query has <= 256 unique candidates, so the FP32 refine
    # rescores all of them and the result must equal the FP32 path exactly
    X, Q, g, o = _built("mix_d128_p2")
    g.set_qProbes(6)
    ref = g.query(Q, 10)
    nc = [len(g.last_candidates(j)) for j in range(len(Q))]
    assert max(nc) <= 256
    try:
        g.set_precision(mode, 256)
        got = g.query(Q, 10)
    finally:
        g.set_precision(E.PREC_FP32, 0)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("case,mode,min_recall", [("mix_d128_p2", E.PREC_BF16, 0.98), ("mix_d128_p2", E.PREC_FP8, 0.9),
                                                  ("mix_d200_p3", E.PREC_BF16, 0.98), ("mix_d200_p3", E.PREC_FP8, 0.9)])
def test_low_precision_recall_and_exact_scores(case, mode, min_recall):
    import torch
    X, Q, g, o = _built(case)
    qp, k = CASES[case][6], CASES[case][7]
    g.set_qProbes(qp)
    ref = g.query(Q, k)
    d = X.shape[1]
    dQ = torch.from_numpy(np.ascontiguousarray(Q)).cuda()
    out = torch.zeros((len(Q), k), dtype=torch.int32, device="cuda")
    sc = torch.zeros((len(Q), k), dtype=torch.float32, device="cuda")
    try:
        g.set_precision(mode, 4 * k)
        g.query_device(dQ.data_ptr(), d, len(Q), k, out.data_ptr(), sc.data_ptr())
        torch.cuda.synchronize()
    finally:
        g.set_precision(E.PREC_FP32, 0)
    got, scores = out.cpu().numpy(), sc.cpu().numpy()
    assert E.recall(ref, got, k) >= min_recall
    # returned scores are the exact FP32 scores of the returned ids
    Xc = (X - g.center()).astype(np.float64)
    for j in range(len(Q)):
        m = min(k, len(g.last_candidates(j)))
        exact = Xc[got[j, :m]] @ Q[j].astype(np.float64)
        np.testing.assert_allclose(scores[j, :m], exact, rtol=1e-5, atol=1e-5)
        assert np.all(np.diff(scores[j, :m]) <= 0)
        assert np.all(got[j, m:] == 0) and np.all(scores[j, m:] == 0)


@pytest.mark.parametrize("mode", [E.PREC_BF16, E.PREC_FP8])
def test_low_precision_candidates_match_oracle_over_windows(mode):
    """The BF16/FP8 gathers collect with their own launch shape (512-thread
    workgroups, rounds of 4096 ids); over dedup windows that are empty,
    of one round and of tens, their candidate lists equal the oracle's."""
    import torch
    X, Q, g, o = _built("windows_far_block")
    qp, k = CASES["windows_far_block"][6], CASES["windows_far_block"][7]
    g.set_qProbes(qp)
    o.set_qProbes(qp)
    d = X.shape[1]
    dQ = torch.from_numpy(np.ascontiguousarray(Q)).cuda()
    out = torch.zeros((len(Q), k), dtype=torch.int32, device="cuda")
    sc = torch.zeros((len(Q), k), dtype=torch.float32, device="cuda")
    try:
        g.set_precision(mode, 4 * k)
        g.query_device(dQ.data_ptr(), d, len(Q), k, out.data_ptr(), sc.data_ptr())
        torch.cuda.synchronize()
        got = [g.last_candidates(j) for j in range(len(Q))]
    finally:
        g.set_precision(E.PREC_FP32, 0)
    for j, q in enumerate(Q):
        assert np.array_equal(got[j], o.query_debug(q, k)["cands"]), j


# ---- small queries: one wave per query ----------------------------------

SMALL_CASES = [
    # (case, qProbes, k): few probes, so (most) queries hold at most 512 ids
    ("mix_d128_p2", 6, 10),
    ("sphere_d100_p1", 20, 20),
    ("mix_d200_p3", 12, 20),
    ("l100_qp10000", 60, 20),        # 100 tables, about one id per bucket
    ("int_grid_ties", 10, 10),       # tie-flagged queries stay on the workgroup path
    ("nproj16_d40", 30, 15),
]


@pytest.mark.parametrize("keep", [0, 1])
@pytest.mark.parametrize("case,qp,k", SMALL_CASES)
def test_small_queries_equal_workgroup_path(case, qp, k, keep):
    """Queries of at most 512 bucket ids served one per wave return the ids,
    scores and candidate counts of the workgroup gather."""
    X, Q, g, o = _built(case)
    g.set_tuning(E.TUNE_KEEP_CANDIDATES, keep)
    try:
        res = {}
        for on in (0, 1):
            g.set_tuning(E.TUNE_SMALL_QUERIES, on)
            ids, sc = _run_single(g, Q, qp, k)
            st = g.stats()
            cands = [g.last_candidates(j) for j in range(len(Q))] if keep else None
            res[on] = (ids, sc, st, cands)
    finally:
        g.set_tuning(E.TUNE_SMALL_QUERIES, 1)
        g.set_tuning(E.TUNE_KEEP_CANDIDATES, 1)
    assert res[0][2]["small_queries"] == 0
    assert res[1][2]["small_queries"] > 0, case
    assert res[0][2]["cuniq"] == res[1][2]["cuniq"] and res[0][2]["ctot"] == res[1][2]["ctot"]
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
    if keep:
        for j in range(len(Q)):
            assert np.array_equal(res[0][3][j], res[1][3][j]), j


@pytest.mark.parametrize("case,qp,k", [("mix_d128_p2", 6, 10), ("l100_qp10000", 60, 20),
                                       ("sphere_d100_p1", 20, 20)])
def test_small_queries_match_oracle(case, qp, k):
    """The one-wave path against the oracle: probes, candidate sets bit-exact,
    top-k within the score tolerance."""
    X, Q, g, o = _built(case)
    g.set_qProbes(qp)
    o.set_qProbes(qp)
    out = g.query(Q, k)
    probes = g.last_probes(len(Q))
    assert g.stats()["small_queries"] > 0
    Xc = X - g.center()
    for j, q in enumerate(Q):
        d = o.query_debug(q, k)
        assert np.array_equal(np.sort(probes[j]), np.sort(d["probes"].astype(np.uint32))), j
        assert np.array_equal(g.last_candidates(j), d["cands"]), j
        m = min(k, len(d["cands"]))
        _check_topk(Xc, q, out[j, :m], d["topk"][:m], d["scores"][:m], k)
        assert np.all(out[j, m:] == 0)


@pytest.mark.parametrize("case,qp,k", [("l100_qp10000", 40, 20), ("int_ties_l200", 60, 10),
                                       ("int_grid_ties", 2, 10), ("full_pool_l64_pl18", 20, 20),
                                       ("sphere_d100_p1", 5, 20)])
def test_one_probe_per_layer_matches_oracle(case, qp, k):
    """qProbes <= L / 2: one probe per layer, every table's best pair; the
    one-wave selection (and the tie flag on integer data, whose queries the
    exact replay then serves) against the oracle."""
    X, Q, g, o = _built(case)
    L = CASES[case][1]
    assert math.ceil(2 * qp / L) == 1
    g.set_qProbes(qp)
    o.set_qProbes(qp)
    out = g.query(Q, k)
    probes = g.last_probes(len(Q))
    Xc = X - g.center()
    ctot = 0
    for j, q in enumerate(Q):
        d = o.query_debug(q, k)
        assert np.array_equal(np.sort(probes[j]), np.sort(d["probes"].astype(np.uint32))), j
        assert np.array_equal(g.last_candidates(j), d["cands"]), j
        ctot += d["ctot"]
        m = min(k, len(d["cands"]))
        _check_topk(Xc, q, out[j, :m], d["topk"][:m], d["scores"][:m], k)
    st = g.stats()
    assert st["ctot"] == ctot
    print(f"[parity] {case} qProbes={qp}: {st['query_ties']} of {len(Q)} queries replayed for ties", flush=True)


@pytest.mark.parametrize("mode", [E.PREC_BF16, E.PREC_FP8])
def test_small_queries_are_scored_exactly_in_every_precision(mode):
    """A small query's candidates are all scored in FP32, whatever the
    precision mode: with every query small the answer is the FP32 answer."""
    X, Q, g, o = _built("mix_d128_p2")
    ref, rsc = _run_single(g, Q, 6, 10)
    assert g.stats()["small_queries"] == len(Q)
    try:
        g.set_precision(mode, 40)
        got, gsc = _run_single(g, Q, 6, 10)
    finally:
        g.set_precision(E.PREC_FP32, 0)
    assert g.stats()["small_queries"] == len(Q)
    assert np.array_equal(got, ref) and np.array_equal(gsc, rsc)


def _bf16(a):
    """float32 -> bfloat16 (round to nearest even) -> float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def test_bf16_refined_set_is_the_best_by_approximate_score():
    """With k equal to the refine count R, the answer is the set the
    approximate pass kept: the R best candidates by BF16 score (smaller id on
    ties).  Thousands of candidates per query, so the selection goes through
    the short list of its best score bins; the host recomputes the BF16
    scores in float64, so the sets may differ only between candidates whose
    scores lie within rounding of the boundary."""
    import torch
    X = synth.gaussian_mixture(400_000, 64, 100, seed=31)
    n, d = X.shape
    R, qp, nq = 80, 400, 16
    g = E.Index(n, d)
    g.setIndexParam(8, 64, 20, 0.05, 1, 8, 42)
    g.build(X)
    rng = np.random.default_rng(5)
    Q = (X[rng.choice(n, nq, replace=False)] + np.float32(0.05) * rng.standard_normal((nq, d))).astype(np.float32)
    g.set_qProbes(qp)
    g.set_tuning(E.TUNE_KEEP_CANDIDATES, 1)
    g.query(Q, 20)
    cands = [g.last_candidates(j) for j in range(nq)]
    g.set_tuning(E.TUNE_KEEP_CANDIDATES, 0)
    print(f"[parity] candidates per query {min(map(len, cands))} .. {max(map(len, cands))}", flush=True)
    assert min(map(len, cands)) > 20 * R
    dQ = torch.from_numpy(Q).cuda()
    out = torch.zeros((nq, R), dtype=torch.int32, device="cuda")
    try:
        g.set_precision(E.PREC_BF16, R)
        g.query_device(dQ.data_ptr(), d, nq, R, out.data_ptr(), 0)
        torch.cuda.synchronize()
    finally:
        g.set_precision(E.PREC_FP32, 0)
    got = out.cpu().numpy()
    Xb = _bf16(X - g.center())
    for j in range(nq):
        c = cands[j]
        approx = Xb[c] @ _bf16(Q[j])
        order = np.lexsort((c, -approx))
        want = set(c[order[:R]].tolist())
        bound = approx[order[R - 1]]
        score = dict(zip(c.tolist(), approx.tolist()))
        diff = set(got[j].tolist()) ^ want
        tol = 1e-5 * (abs(bound) + np.abs(approx).max())
        assert all(abs(score[i] - bound) <= tol for i in diff), (j, sorted(diff))
        assert len(set(got[j].tolist())) == R


@pytest.mark.parametrize("d", [128, 100])
def test_fp8_rows_keep_their_own_scales_when_magnitudes_spread(d):
    # rows scaled by 2^-6 .. 2^6, with their negatives so that the centre is
    # about zero: the centred rows' largest magnitudes span more than 2^8, so
    # each row keeps its own FP8 scale (the cases above share one); at d=100
    # an FP8 row has 28 bytes of padding and keeps its scale there
    import torch
    rng = np.random.default_rng(21)
    A = synth.gaussian_mixture(6000, d, 60, seed=13)
    A = A * np.exp2(rng.integers(-6, 7, len(A))).astype(np.float32)[:, None]
    X = np.concatenate([A, -A])[rng.permutation(2 * len(A))].astype(np.float32)
    n, d = X.shape
    g = E.Index(n, d)
    g.setIndexParam(8, 128, 10, 0.2, 2, 8, 42)
    g.build(X)
    g.set_tuning(E.TUNE_KEEP_CANDIDATES, 1)
    am = np.abs(X - g.center()).max(1)
    am = am[am > 0]
    assert am.max() > 256 * am.min()
    Q = X[rng.choice(n, 50, replace=False)] + np.float32(0.05) * rng.standard_normal((50, d)).astype(np.float32)
    Q = Q.astype(np.float32)
    # every candidate refined: equal to the FP32 path
    g.set_qProbes(6)
    ref = g.query(Q, 10)
    assert max(len(g.last_candidates(j)) for j in range(len(Q))) <= 256
    try:
        g.set_precision(E.PREC_FP8, 256)
        got = g.query(Q, 10)
    finally:
        g.set_precision(E.PREC_FP32, 0)
    assert np.array_equal(got, ref)
    # R = 4k survivors: recall against the FP32 path, exact scores
    g.set_qProbes(160)
    ref = g.query(Q, 10)
    dQ = torch.from_numpy(np.ascontiguousarray(Q)).cuda()
    out = torch.zeros((len(Q), 10), dtype=torch.int32, device="cuda")
    sc = torch.zeros((len(Q), 10), dtype=torch.float32, device="cuda")
    try:
        g.set_precision(E.PREC_FP8, 40)
        g.query_device(dQ.data_ptr(), d, len(Q), 10, out.data_ptr(), sc.data_ptr())
        torch.cuda.synchronize()
    finally:
        g.set_precision(E.PREC_FP32, 0)
    got, scores = out.cpu().numpy(), sc.cpu().numpy()
    assert E.recall(ref, got, 10) >= 0.9
    Xc = (X - g.center()).astype(np.float64)
    for j in range(len(Q)):
        m = min(10, len(g.last_candidates(j)))
        np.testing.assert_allclose(scores[j, :m], Xc[got[j, :m]] @ Q[j].astype(np.float64), rtol=1e-5, atol=1e-5)


def test_low_precision_rejects_k_above_refine():
    X, Q, g, o = _built("mix_d128_p2")
    g.set_precision(E.PREC_BF16, 8)
    try:
        with pytest.raises(E.FppError) as e:
            g.query(Q[:2], 10)
        assert e.value.code == E.FPP_EINVAL
    finally:
        g.set_precision(E.PREC_FP32, 0)


# ---- batched serving: per-query qProbes and k, persistent gather, graph ----

def _batch_case(seed=3):
    X, Q, g, o = _built("mix_d128_p2")
    rng = np.random.default_rng(seed)
    qps = rng.choice([20, 60, 160, 400], len(Q)).astype(np.int32)
    ks = rng.choice([1, 5, 10], len(Q)).astype(np.int32)
    return X, Q, g, qps, ks


def _run_batch(g, Q, qps, ks, kmax):
    import torch
    dQ = torch.from_numpy(np.ascontiguousarray(Q)).cuda()
    out = torch.zeros((len(Q), kmax), dtype=torch.int32, device="cuda")
    sc = torch.zeros((len(Q), kmax), dtype=torch.float32, device="cuda")
    g.query_batch_device(dQ.data_ptr(), Q.shape[1], len(Q), qps, ks, kmax, out.data_ptr(), sc.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy(), sc.cpu().numpy()


def _run_single(g, Q, qp, k):
    import torch
    g.set_qProbes(int(qp))
    dQ = torch.from_numpy(np.ascontiguousarray(Q)).cuda()
    out = torch.zeros((len(Q), k), dtype=torch.int32, device="cuda")
    sc = torch.zeros((len(Q), k), dtype=torch.float32, device="cuda")
    g.query_device(dQ.data_ptr(), Q.shape[1], len(Q), k, out.data_ptr(), sc.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy(), sc.cpu().numpy()


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("persistent", [0, 1])
def test_batch_equals_unbatched(precision, persistent):
    """Every query of a batch with mixed qProbes and k returns what the plain
    call with that query's qProbes and k returns (ids, scores, probes)."""
    X, Q, g, qps, ks = _batch_case()
    kmax = int(ks.max())
    g.set_precision(precision, 0)
    g.set_tuning(E.TUNE_PERSISTENT, persistent)
    try:
        out, sc = _run_batch(g, Q, qps, ks, kmax)
        probes = g.last_probes_batch(qps)
        for qp in np.unique(qps):
            for k in np.unique(ks):
                sel = np.where((qps == qp) & (ks == k))[0]
                if len(sel) == 0:
                    continue
                ro, rs = _run_single(g, Q[sel], qp, int(k))
                rp = g.last_probes(len(sel))
                for i, j in enumerate(sel):
                    assert np.array_equal(out[j, :k], ro[i]), (qp, k, j)
                    assert np.array_equal(sc[j, :k], rs[i]), (qp, k, j)
                    assert np.all(out[j, k:] == 0) and np.all(sc[j, k:] == 0)
                    assert np.array_equal(np.sort(probes[j]), np.sort(rp[i])), (qp, k, j)
    finally:
        g.set_precision(0, 0)
        g.set_tuning(E.TUNE_PERSISTENT, 0)


def test_persistent_equals_one_workgroup_per_query():
    X, Q, g, o = _built("multiwindow_n1.3m")
    qp, k = CASES["multiwindow_n1.3m"][6], CASES["multiwindow_n1.3m"][7]
    ref, rsc = _run_single(g, Q, qp, k)
    g.set_tuning(E.TUNE_PERSISTENT, 1)
    try:
        out, sc = _run_single(g, Q, qp, k)
    finally:
        g.set_tuning(E.TUNE_PERSISTENT, 0)
    assert np.array_equal(out, ref) and np.array_equal(sc, rsc)


def test_graph_replays_equal_eager():
    """Hashing + probe selection captured once and replayed: every replay,
    with eager calls and synchronisations in between, equals the eager run."""
    X, Q, g, qps, ks = _batch_case(seed=5)
    kmax = int(ks.max())
    ref, rsc = _run_batch(g, Q, qps, ks, kmax)
    g.set_tuning(E.TUNE_GRAPH, 1)
    try:
        for rep in range(4):
            out, sc = _run_batch(g, Q, qps, ks, kmax)
            assert np.array_equal(out, ref) and np.array_equal(sc, rsc), rep
            g.set_tuning(E.TUNE_GRAPH, 0)
            e_out, _ = _run_single(g, Q, 160, 10)     # eager work between replays
            g.set_tuning(E.TUNE_GRAPH, 1)
            # the graph was dropped by the knob and is captured again
    finally:
        g.set_tuning(E.TUNE_GRAPH, 0)
    # replays without re-capture: the same batch four times in a row, with
    # eager device work of the program and synchronisations in between
    import torch
    g.set_tuning(E.TUNE_GRAPH, 1)
    try:
        for rep in range(4):
            out, sc = _run_batch(g, Q, qps, ks, kmax)
            assert np.array_equal(out, ref) and np.array_equal(sc, rsc), rep
            t = torch.arange(1 << 20, device="cuda", dtype=torch.float32)
            assert float((t * 2).max().item()) == 2.0 * ((1 << 20) - 1)
            torch.cuda.synchronize()
    finally:
        g.set_tuning(E.TUNE_GRAPH, 0)


def test_autotune_keeps_results():
    import torch
    X, Q, g, o = _built("mix_d128_p2")
    qp, k = CASES["mix_d128_p2"][6], CASES["mix_d128_p2"][7]
    ref, rsc = _run_single(g, Q, qp, k)
    dQ = torch.from_numpy(np.ascontiguousarray(Q)).cuda()
    ch = g.autotune(dQ.data_ptr(), Q.shape[1], len(Q), k)
    assert ch["gather_threads"] in (256, 512, 1024) and ch["us_per_batch"] > 0
    try:
        out, sc = _run_single(g, Q, qp, k)
        assert np.array_equal(out, ref) and np.array_equal(sc, rsc)
    finally:
        g.set_tuning(E.TUNE_PERSISTENT, 0)
        g.set_tuning(E.TUNE_QUERY_ORDER, 1)
        g.set_tuning(E.TUNE_GATHER_THREADS, 512)


def test_autotune_batch_keeps_results():
    """The batched autotuner (sample with its own qProbes and k per query)
    changes launch shapes only: the batch answers as before."""
    import torch
    X, Q, g, qps, ks = _batch_case()
    kmax = int(ks.max())
    ref, rsc = _run_batch(g, Q, qps, ks, kmax)
    dQ = torch.from_numpy(np.ascontiguousarray(Q)).cuda()
    m = min(len(Q), 256)
    ch = g.autotune_batch(dQ.data_ptr(), Q.shape[1], m, qps[:m], ks[:m], kmax)
    assert ch["gather_threads"] in (256, 512, 1024) and ch["us_per_batch"] > 0
    assert ch["query_order"] in (1, 2) and ch["persistent"] == (1 if ch["query_order"] == 2 else 0)
    try:
        out, sc = _run_batch(g, Q, qps, ks, kmax)
        assert np.array_equal(out, ref) and np.array_equal(sc, rsc)
    finally:
        g.set_tuning(E.TUNE_PERSISTENT, 0)
        g.set_tuning(E.TUNE_QUERY_ORDER, 1)
        g.set_tuning(E.TUNE_GATHER_THREADS, 512)


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_kept_candidates_and_graph_do_not_change_results(precision):
    """The production call (each gather workgroup reuses its own scratch slot,
    nothing read back) and its hipGraph replays answer exactly like the
    inspection call that keeps every query's candidates."""
    import torch
    X, Q, g, qps, ks = _batch_case(seed=9)
    kmax = int(ks.max())
    g.set_precision(precision, 0)
    try:
        ref, rsc = _run_batch(g, Q, qps, ks, kmax)
        rst = g.stats()
        g.set_tuning(E.TUNE_KEEP_CANDIDATES, 0)
        for graph in (0, 1, 1, 1):
            g.set_tuning(E.TUNE_GRAPH, graph)
            out, sc = _run_batch(g, Q, qps, ks, kmax)
            assert np.array_equal(out, ref) and np.array_equal(sc, rsc), graph
            st = g.stats()
            assert (st["ctot"], st["cuniq"], st["query_ties"]) == (rst["ctot"], rst["cuniq"], rst["query_ties"])
            t = torch.arange(1 << 16, device="cuda", dtype=torch.float32)
            assert float((t + 1).sum().item()) == float((1 << 16) * ((1 << 16) + 1) // 2)
        with pytest.raises(E.FppError) as e:
            g.last_candidates(0)
        assert e.value.code == E.FPP_ESTATE
    finally:
        g.set_tuning(E.TUNE_GRAPH, 0)
        g.set_tuning(E.TUNE_KEEP_CANDIDATES, 1)
        g.set_precision(0, 0)


def test_workspace_bound_does_not_change_results():
    """A tight bound on the per-workgroup candidate scratch runs the gather on
    fewer workgroups (one or two at 1 MiB); the answers stay the same, for a
    plain call and a mixed batch, in FP32 and FP8."""
    X, Q, g, qps, ks = _batch_case(seed=17)
    kmax = int(ks.max())
    g.set_tuning(E.TUNE_KEEP_CANDIDATES, 0)  # kept lists bypass the bound
    try:
        for prec in (E.PREC_FP32, E.PREC_FP8):
            g.set_precision(prec, 0)
            g.set_tuning(E.TUNE_WORKSPACE_MB, 8192)
            ref = _run_batch(g, Q, qps, ks, kmax)
            ref1 = _run_single(g, Q, 400, 10)
            g.set_tuning(E.TUNE_WORKSPACE_MB, 1)
            out = _run_batch(g, Q, qps, ks, kmax)
            out1 = _run_single(g, Q, 400, 10)
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1]), prec
            assert np.array_equal(out1[0], ref1[0]) and np.array_equal(out1[1], ref1[1]), prec
    finally:
        g.set_tuning(E.TUNE_WORKSPACE_MB, 8192)
        g.set_precision(E.PREC_FP32, 0)
        g.set_tuning(E.TUNE_KEEP_CANDIDATES, 1)


def test_order_modes_and_workgroup_sizes_do_not_change_results():
    X, Q, g, qps, ks = _batch_case(seed=13)
    kmax = int(ks.max())
    ref, rsc = _run_batch(g, Q, qps, ks, kmax)
    try:
        for order in (0, 1, 2):
            for t in (256, 512, 1024):
                g.set_tuning(E.TUNE_QUERY_ORDER, order)
                g.set_tuning(E.TUNE_GATHER_THREADS, t)
                out, sc = _run_batch(g, Q, qps, ks, kmax)
                assert np.array_equal(out, ref) and np.array_equal(sc, rsc), (order, t)
    finally:
        g.set_tuning(E.TUNE_QUERY_ORDER, 1)
        g.set_tuning(E.TUNE_GATHER_THREADS, 512)
