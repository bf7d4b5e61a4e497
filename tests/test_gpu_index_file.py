"""Index files: an index written by save and restored by load into a new
handle answers every query as the built one does (same tables, sign bits,
centre and data; FP32, reduced precision and the exact-only probe path), and
bad files are refused."""
import os

import numpy as np
import pytest

from falconnpp_amd import engine as E
from falconnpp_amd import synth

pytestmark = pytest.mark.gpu


def _index(X, L=12, n_proj=128, blim=10, alpha=0.05, p=2):
    g = E.Index(*X.shape)
    g.setIndexParam(L, n_proj, blim, alpha, p, 8, 42)
    return g


def test_saved_index_answers_as_built(tmp_path):
    X = synth.gaussian_mixture(20000, 48, 80, seed=41)
    rng = np.random.default_rng(4)
    Q = (X[rng.choice(len(X), 64, replace=False)] + 0.05 * rng.standard_normal((64, 48))).astype(np.float32)
    g = _index(X)
    g.build(X)
    path = str(tmp_path / "idx.bin")
    print("[index-file] save/load n=20000 d=48", flush=True)
    g.save(path)
    h = _index(X, L=3, n_proj=16, blim=1, alpha=1.0, p=1)  # parameters come from the file
    h.load(path)
    assert h.shape() == g.shape()
    for a, b in zip(h.sign_bits(), g.sign_bits()):
        assert np.array_equal(a, b)
    assert np.array_equal(h.center(), g.center())
    for a, b in zip(h.buckets(), g.buckets()):
        assert np.array_equal(a, b)
    assert h.stats()["hash_ties"] == g.stats()["hash_ties"]
    # 300 probes (fast path), 9000 (more than 31 probes per layer: exact replay)
    for qp in (300, 9000):
        g.set_qProbes(qp)
        h.set_qProbes(qp)
        assert np.array_equal(h.query(Q, 10), g.query(Q, 10)), qp
    g.set_qProbes(300)
    h.set_qProbes(300)
    for mode in (E.PREC_BF16, E.PREC_FP8):
        g.set_precision(mode, 40)
        h.set_precision(mode, 40)
        assert np.array_equal(h.query(Q, 10), g.query(Q, 10)), mode


def test_bad_index_files_are_refused(tmp_path):
    X = synth.gaussian_mixture(3000, 16, 20, seed=42)
    g = _index(X, L=4, n_proj=32)
    with pytest.raises(E.FppError) as e:
        g.save(str(tmp_path / "early.bin"))
    assert e.value.code == E.FPP_ESTATE
    g.build(X)
    path = tmp_path / "idx.bin"
    g.save(str(path))
    with pytest.raises(E.FppError) as e:
        g.load(str(tmp_path / "missing.bin"))
    assert e.value.code == E.FPP_EIO
    other = E.Index(2999, 16)
    with pytest.raises(E.FppError) as e:
        other.load(str(path))
    assert e.value.code == E.FPP_EINVAL
    data = path.read_bytes()
    (tmp_path / "short.bin").write_bytes(data[:-4])
    (tmp_path / "magic.bin").write_bytes(b"X" + data[1:])
    # sections after the 160-byte header: sign bits (twice), flip masks,
    # centre, bucket offsets (u64), ids (i32), data
    hdr = np.frombuffer(data[:160], np.int64, 19, 0)
    n_ids, n_bits, n_mask, n_off = int(hdr[12]), int(hdr[16]), int(hdr[17]), int(hdr[18])
    off_pos = 160 + 2 * n_bits + 4 * n_mask + 4 * 16
    ids_pos = off_pos + 8 * n_off
    off = np.frombuffer(data[off_pos:ids_pos], np.uint64)
    assert off[-1] == n_ids and n_ids > 0
    b = int(np.nonzero(off[1:] > off[:-1])[0][0])  # a non-empty bucket
    dec = off.copy()
    dec[b + 1] = n_ids if dec[b + 2] < n_ids else 0  # offsets that do not increase
    (tmp_path / "offsets.bin").write_bytes(data[:off_pos] + dec.tobytes() + data[ids_pos:])
    ids = np.frombuffer(data[ids_pos:ids_pos + 4 * n_ids], np.int32).copy()
    ids[n_ids // 2] = 3000  # a point id past n_points
    (tmp_path / "ids.bin").write_bytes(data[:ids_pos] + ids.tobytes() + data[ids_pos + 4 * n_ids:])
    for bad in ("short.bin", "magic.bin", "offsets.bin", "ids.bin"):
        h = E.Index(3000, 16)
        with pytest.raises(E.FppError) as e:
            h.load(str(tmp_path / bad))
        assert e.value.code == E.FPP_EINVAL, bad
    # the index that failed to load is still usable after a good load
    h.load(str(path))
    g.set_qProbes(50)
    h.set_qProbes(50)
    Q = X[:8]
    assert np.array_equal(h.query(Q, 5), g.query(Q, 5))


def test_pybind_save_load(tmp_path):
    import sys
    sys.path.insert(0, os.path.dirname(E.LIB_PATH))
    import FalconnPP
    X = synth.gaussian_mixture(3000, 16, 20, seed=43)
    f = FalconnPP.FalconnPP(3000, 16)
    f.setIndexParam(4, 32, 5, 0.1, 1, 8, 7)
    f.build(np.asfortranarray(X.T))
    f.set_qProbes(40)
    ref = f.query(np.asfortranarray(X[:10].T), 5)
    path = str(tmp_path / "p.bin")
    f.save(path)
    g = FalconnPP.FalconnPP(3000, 16)
    g.load(path)
    g.set_qProbes(40)
    assert np.array_equal(g.query(np.asfortranarray(X[:10].T), 5), ref)
    with pytest.raises(OSError):
        g.load(str(tmp_path / "none.bin"))
