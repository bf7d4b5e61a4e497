"""Text formats of the reference's command-line driver, on the CPU: the data
loader (src/Utilities.cpp:41-60, fscanf "%f" into a zeroed matrix), the ids
writer (src/Utilities.cpp:13-31) and the verbose output file name
(src/FalconnPP.cpp:694-700); and the driver's argument handling up to the
first device call."""
import os
import subprocess

import numpy as np
import pytest

from falconnpp_amd import engine as E

CLI = os.path.join(os.path.dirname(E.LIB_PATH), "falconnpp")


def _write(path, text):
    with open(path, "w", newline="") as f:
        f.write(text)


def test_load_round_trips_float32_text(tmp_path):
    # 9 significant digits name every float32 exactly, so the loader must
    # return the written values bit for bit; separators vary
    rng = np.random.default_rng(3)
    X = (rng.standard_normal((200, 7)) * np.exp2(rng.integers(-20, 20, (200, 7)))).astype(np.float32)
    seps = [" ", "\t", "  ", " \t "]
    lines = [seps[i % 4].join("%.9g" % v for v in row) for i, row in enumerate(X)]
    p = tmp_path / "x.txt"
    _write(p, "\r\n".join(lines[:100]) + "\n" + "\n".join(lines[100:]) + "\n")
    got, n = E.load_txt(str(p), 200, 7)
    assert n == 1400
    assert np.array_equal(got.view(np.uint32), X.view(np.uint32))


def test_load_is_a_stream_of_numbers(tmp_path):
    # line breaks do not delimit points; signs and exponents as %f reads them
    p = tmp_path / "x.txt"
    _write(p, "1 +2.5 -3e-2\n4\n5 6\n.5 7E1\n")
    got, n = E.load_txt(str(p), 2, 4)
    assert n == 8
    assert np.array_equal(got, np.array([[1, 2.5, -0.03, 4], [5, 6, 0.5, 70]], np.float32))


def test_load_short_file_and_bad_tokens_leave_zeros(tmp_path):
    p = tmp_path / "x.txt"
    _write(p, "1 2 3\n")
    got, n = E.load_txt(str(p), 2, 3)
    assert n == 3 and np.array_equal(got, np.array([[1, 2, 3], [0, 0, 0]], np.float32))
    # a token that is not a number ends the read; one that starts with a
    # number gives that number, then ends it
    _write(p, "1 2 abc 4 5 6\n")
    got, n = E.load_txt(str(p), 2, 3)
    assert n == 2 and np.array_equal(got.ravel(), np.array([1, 2, 0, 0, 0, 0], np.float32))
    _write(p, "1 2.5xyz 3 4\n")
    got, n = E.load_txt(str(p), 1, 4)
    assert n == 2 and np.array_equal(got.ravel(), np.array([1, 2.5, 0, 0], np.float32))


def test_load_across_read_chunks(tmp_path):
    # more than the loader's 16 MiB read size, with tokens of uneven length,
    # so chunk ends fall inside tokens
    rng = np.random.default_rng(5)
    X = rng.standard_normal((300_000, 8)).astype(np.float32)
    text = "\n".join(" ".join("%.9g" % v for v in row) for row in X) + "\n"
    assert len(text) > (1 << 24)
    p = tmp_path / "big.txt"
    _write(p, text)
    got, n = E.load_txt(str(p), 300_000, 8)
    assert n == X.size
    assert np.array_equal(got.view(np.uint32), X.view(np.uint32))


def test_load_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(E.FppError) as e:
        E.load_txt(str(tmp_path / "none.txt"), 2, 2)
    assert e.value.code == E.FPP_EIO


def test_write_topk_format(tmp_path):
    p = tmp_path / "o.txt"
    E.write_topk_txt(str(p), np.array([[3, 1, 0], [7, 0, 0]], np.int32))
    assert open(p).read() == "3 1 0 \n7 0 0 \n"
    with pytest.raises(E.FppError) as e:
        E.write_topk_txt(str(tmp_path / "no_dir" / "o.txt"), np.zeros((1, 1), np.int32))
    assert e.value.code == E.FPP_EIO


@pytest.mark.parametrize("scale", [0.005, 0.01, 0.29, 0.53, 1.0, 2.0])
def test_topk_file_name(scale):
    # the scale enters as (int)(scale * 100) in float arithmetic
    pct = int(np.float32(scale) * np.float32(100))
    assert E.topk_file_name(20, 256, 100, 3, 20, scale, 4000) == \
        f"1D_Top_20_D_256_L_100_iProbe_3_bucketMinSize_20_bucketScale_{pct}_qProbe_4000.txt"
    assert int(np.float32(0.53) * np.float32(100)) == 52  # truncation of the float product


def _cli(args, cwd):
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


def test_cli_argument_errors(tmp_path):
    assert os.path.exists(CLI), "build the driver with make -C falconnpp_amd/csrc"
    r = _cli(["--n_features", "4", "--n_queries", "2", "--X", "x"], tmp_path)
    assert r.returncode == 1 and "Number of rows/points is missing" in r.stderr
    r = _cli(["--n_points", "3", "--n_features", "4", "--X", "x.txt", "--Q", "q.txt"], tmp_path)
    assert r.returncode == 1 and "Number of queries is missing" in r.stderr
    r = _cli(["--n_points", "3", "--n_features", "200", "--n_queries", "2", "--X", "none.txt", "--Q", "q.txt"],
             tmp_path)
    assert r.returncode == 1 and "cannot open data file none.txt" in r.stderr
    # defaults as the reference prints them; n_proj = max(256, 2^ceil(log2 d))
    assert "Default number of LSH tables: 10" in r.stdout
    assert "Number of projections: 256" in r.stdout
    r = _cli(["--n_points", "3", "--n_features", "300", "--n_queries", "2", "--X", "none.txt"], tmp_path)
    assert "Number of projections: 512" in r.stdout
