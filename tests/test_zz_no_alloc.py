"""fpp_query_device_noalloc against fpp_query_device (needs an MI355X).

The entry point keeps every scratch buffer of a call in the library's fixed
device array.  Its answers (ids and scores) must equal the allocating call's
bit for bit, on two calls in a row, the second with a smaller batch (the
array still holds the first call's contents), for FP32, BF16 and FP8
scoring, with one dedup window and with two."""
import numpy as np
import pytest

from falconnpp_amd import engine as E
from falconnpp_amd import synth

pytestmark = pytest.mark.gpu


def _call(g, Q, k, fixed):
    import torch
    dQ = torch.from_numpy(np.ascontiguousarray(Q)).cuda()
    out = torch.full((len(Q), k), -7, dtype=torch.int32, device="cuda")
    sc = torch.full((len(Q), k), -7.0, dtype=torch.float32, device="cuda")
    fn = g.query_device_noalloc if fixed else g.query_device
    fn(dQ.data_ptr(), Q.shape[1], len(Q), k, out.data_ptr(), sc.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy(), sc.cpu().numpy()


CASES = {
    # name: (data, L, n_proj, bucketLimit, alpha, iProbes, qProbes, k)
    "one_window": (lambda: synth.gaussian_mixture(30000, 64, 60, seed=61), 8, 64, 10, 0.1, 2, 200, 20),
    "two_windows": (lambda: synth.gaussian_mixture(1_300_000, 16, 300, seed=62), 4, 16, 30, 0.05, 1, 64, 20),
}


@pytest.mark.parametrize("name", list(CASES))
def test_noalloc_equals_query_device(name):
    data, L, n_proj, blim, alpha, p, qp, k = CASES[name]
    X = data()
    n, d = X.shape
    print(f"[noalloc] case={name} n={n} d={d}", flush=True)
    rng = np.random.default_rng(5)
    Q = (X[rng.choice(n, 600, replace=False)] + np.float32(0.05) * rng.standard_normal((600, d))).astype(np.float32)
    g = E.Index(n, d)
    g.setIndexParam(L, n_proj, blim, alpha, p, 8, 42)
    g.build(X)
    g.set_qProbes(qp)
    try:
        for precision in (0, 1, 2):
            g.set_precision(precision, 0)
            for nq in (600, 150):
                ref = _call(g, Q[:nq], k, False)
                got = _call(g, Q[:nq], k, True)
                assert np.array_equal(got[0], ref[0]), (precision, nq)
                assert np.array_equal(got[1], ref[1]), (precision, nq)
    finally:
        g.set_precision(0, 0)


def test_noalloc_rejects_what_it_cannot_serve():
    import torch
    X = synth.gaussian_mixture(20000, 32, 40, seed=63)
    g = E.Index(*X.shape)
    g.setIndexParam(8, 32, 10, 0.1, 1, 8, 42)
    g.build(X)
    # the FP8 copy is made by an allocating call first
    g.set_precision(E.PREC_FP8, 0)
    with pytest.raises(E.FppError) as e:
        _call(g, X[:4], 5, True)
    assert e.value.code == E.FPP_ESTATE
    g.set_precision(E.PREC_FP32, 0)
    # a batch whose per-query arrays exceed the fixed array
    nq = 400_000
    g.set_qProbes(400)
    dQ = torch.zeros((nq, X.shape[1]), dtype=torch.float32, device="cuda")
    out = torch.zeros((nq, 5), dtype=torch.int32, device="cuda")
    with pytest.raises(E.FppError) as e:
        g.query_device_noalloc(dQ.data_ptr(), X.shape[1], nq, 5, out.data_ptr())
    assert e.value.code == E.FPP_EINVAL
    # still serves after the refusal
    g.set_qProbes(40)
    ref = _call(g, X[:50], 5, False)
    got = _call(g, X[:50], 5, True)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
