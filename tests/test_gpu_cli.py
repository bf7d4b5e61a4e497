"""The command-line driver end to end on the GPU: text data in, the
reference's qProbes sweep (src/main.cpp:80-84) with verbose queries, and one
1D_Top_*.txt file per sweep step whose ids equal the engine's own query for
that qProbes."""
import os
import subprocess

import numpy as np
import pytest

from falconnpp_amd import engine as E
from falconnpp_amd import synth

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(E.LIB_PATH), "falconnpp")


@pytest.mark.parametrize("L", [10, 100])
def test_cli_sweep_matches_engine(tmp_path, L):
    """The sweep's last steps at 100 tables take 15 probes per layer and a
    pool of 100 x 225 pairs."""
    n, d, nq, k = 3000, 20, 40, 10
    n_proj, blim, alpha, p = 128, 20, 0.01, 1
    X = synth.lognormal_sphere(n, d, 0.3, seed=31)
    Q = synth.lognormal_sphere(nq, d, 0.3, seed=32, unit=True)
    for name, M in (("x.txt", X), ("q.txt", Q)):
        with open(tmp_path / name, "w") as f:
            f.write("\n".join(" ".join("%.9g" % v for v in row) for row in M) + "\n")
    args = [CLI, "--n_points", str(n), "--n_features", str(d), "--n_tables", str(L), "--n_proj", str(n_proj),
            "--bucket_minSize", str(blim), "--bucket_scale", str(alpha), "--iProbes", str(p), "--random_seed", "42",
            "--n_queries", str(nq), "--n_neighbors", str(k), "--verbose", "--X", "x.txt", "--Q", "q.txt"]
    print(f"[cli] sweep 1000..10000 L={L}", flush=True)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("Average number of empty buckets per query") == 10
    assert r.stdout.count("Querying Time") == 10
    g = E.Index(n, d)
    g.setIndexParam(L, n_proj, blim, alpha, p, 8, 42)
    Xt, _ = E.load_txt(str(tmp_path / "x.txt"), n, d)
    Qt, _ = E.load_txt(str(tmp_path / "q.txt"), nq, d)
    assert np.array_equal(Xt, X) and np.array_equal(Qt, Q)
    g.build(Xt)
    for qp in range(1000, 10001, 1000):
        path = tmp_path / E.topk_file_name(k, n_proj, L, p, blim, alpha, qp)
        assert path.exists(), path.name
        got = np.array([[int(t) for t in line.split()] for line in open(path).read().splitlines()], np.int32)
        g.set_qProbes(qp)
        assert np.array_equal(got, g.query(Qt, k)), qp


def test_query2d_verbose_writes_its_own_file(tmp_path, monkeypatch):
    """query2D (query2Layers, src/FalconnPP.cpp:1013-1413) answers like query
    and its verbose call writes 2D_Top_...txt (:1400-1406), through the C ABI
    and through the pybind module."""
    import sys
    n, d, nq, k = 3000, 20, 12, 10
    L, n_proj, blim, alpha, p, qp = 10, 128, 20, 0.01, 1, 300
    X = synth.lognormal_sphere(n, d, 0.3, seed=33)
    Q = synth.lognormal_sphere(nq, d, 0.3, seed=34, unit=True)
    monkeypatch.chdir(tmp_path)
    g = E.Index(n, d)
    g.setIndexParam(L, n_proj, blim, alpha, p, 8, 42)
    g.build(X)
    g.set_qProbes(qp)
    ref = g.query(Q, k)
    got = g.query2D(Q, k, verbose=True)
    assert np.array_equal(got, ref)
    name2 = E.topk_file_name(k, n_proj, L, p, blim, alpha, qp, layout="2D")
    assert name2.startswith("2D_Top_") and name2[2:] == E.topk_file_name(k, n_proj, L, p, blim, alpha, qp)[2:]
    assert not (tmp_path / E.topk_file_name(k, n_proj, L, p, blim, alpha, qp)).exists()
    rows = [[int(t) for t in line.split()] for line in open(tmp_path / name2).read().splitlines()]
    assert np.array_equal(np.array(rows, np.int32), ref)
    (tmp_path / name2).unlink()
    sys.path.insert(0, os.path.dirname(E.LIB_PATH))
    import FalconnPP
    f = FalconnPP.FalconnPP(n, d)
    f.setIndexParam(L, n_proj, blim, alpha, p, 8, 42)
    f.build2D(np.asfortranarray(X.T))
    f.set_qProbes(qp)
    assert np.array_equal(f.query2D(np.asfortranarray(Q.T), k, True), ref)
    rows = [[int(t) for t in line.split()] for line in open(tmp_path / name2).read().splitlines()]
    assert np.array_equal(np.array(rows, np.int32), ref)
