"""Tensor files: the parallel .tns reader (load_tensor semantics, util.h:313-370),
the .tns writer and the binary COO cache, checked by round trips and against
the oracle's single-threaded reader."""
import numpy as np
import pytest

import oracle as O
from test_gpu_builder import random_tensor

pytestmark = pytest.mark.gpu


def write_tns(path, dims, inds, vals, fmt="%.9g"):
    with open(path, "w") as fh:
        fh.write(f"{len(dims)}\n" + " ".join(str(d) for d in dims) + "\n")
        cols = [np.asarray(a, np.int64) + 1 for a in inds]
        for x in range(len(vals)):
            fh.write(" ".join(str(c[x]) for c in cols) + " " + (fmt % vals[x]) + "\n")


@pytest.mark.parametrize("dims,nnz", [([30, 20, 10], 2000), ([3000, 2000, 1000, 50], 120000)])
def test_tns_reader_matches_oracle_and_keeps_file_order(cuda, tmp_path, dims, nnz):
    import mmcsf
    rng = np.random.default_rng(nnz)
    inds, vals = random_tensor(rng, dims, nnz)
    p = tmp_path / "x.tns"
    write_tns(p, dims, inds, vals)
    assert p.stat().st_size > (1 << 20) or nnz < 10000  # the large case is parsed in several chunks
    t = mmcsf.Tensor.load_tns(p)
    d, gi, gv = t.to_numpy()
    ref = O.load_tns(str(p))
    assert list(d) == list(ref[0])
    for m in range(len(dims)):
        np.testing.assert_array_equal(gi[m], ref[1][m])
        np.testing.assert_array_equal(gi[m], inds[m])
    np.testing.assert_array_equal(gv, ref[2])
    np.testing.assert_array_equal(gv, vals)


def test_tns_writer_and_binary_cache_round_trip(cuda, tmp_path):
    import mmcsf
    rng = np.random.default_rng(3)
    dims = [500, 400, 300]
    inds, vals = random_tensor(rng, dims, 50000)
    vals = (vals * rng.standard_normal(len(vals)) * 1e3).astype(np.float32)  # signs and exponents
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    t.save_tns(tmp_path / "a.tns")
    t.save_bin(tmp_path / "a.bin")
    for u in (mmcsf.Tensor.load_tns(tmp_path / "a.tns"), mmcsf.Tensor.load_bin(tmp_path / "a.bin")):
        d, gi, gv = u.to_numpy()
        assert list(d) == dims
        for m in range(3):
            np.testing.assert_array_equal(gi[m], inds[m])
        np.testing.assert_array_equal(gv.view(np.uint32), vals.view(np.uint32))
    assert (tmp_path / "a.bin").stat().st_size == 56 + len(vals) * 16  # header + 3 coordinates + value


def test_tns_reader_errors(cuda, tmp_path):
    import mmcsf
    with pytest.raises(mmcsf.MMCSFError, match="does not exist"):
        mmcsf.Tensor.load_tns(tmp_path / "missing.tns")
    bad = tmp_path / "bad.tns"
    bad.write_text("3\n4 4 4\n1 1 1 1.0\n1 5 1 2.0\n")
    with pytest.raises(mmcsf.MMCSFError, match="out of range"):
        mmcsf.Tensor.load_tns(bad)
    bad.write_text("3\n4 4 4\n1 1 1 1.0\n2 2\n")
    with pytest.raises(mmcsf.MMCSFError):
        mmcsf.Tensor.load_tns(bad)
    bad.write_text("5\n4 4 4 4 4\n")
    with pytest.raises(mmcsf.MMCSFError, match="3 or 4"):
        mmcsf.Tensor.load_tns(bad)
    bad.write_bytes(b"not a binary tensor file at all, just some text" * 2)
    with pytest.raises(mmcsf.MMCSFError, match="not a binary COO"):
        mmcsf.Tensor.load_bin(bad)
