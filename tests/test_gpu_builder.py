"""GPU format builder vs the CPU oracle: every comparable reference array must
be equal bit for bit (fbrPtr, fbrIdx without the trailing entry,
fbrLikeSlcInds without its last entry, slcMapperBin, HB-CSF parts, and the
nonzero order, which is canonical under the full-key sort)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def random_tensor(rng, dims, nnz, skew=None):
    inds = []
    for m, d in enumerate(dims):
        if skew and skew[m]:
            r = rng.zipf(skew[m], size=nnz * 3)
            r = r[r <= d][:nnz] - 1
            if len(r) < nnz:
                r = np.concatenate([r, rng.integers(0, d, nnz - len(r))])
            inds.append(r.astype(np.int64))
        else:
            inds.append(rng.integers(0, d, nnz).astype(np.int64))
    key = np.zeros(nnz, dtype=np.int64)
    for m, d in enumerate(dims):
        key = key * d + inds[m]
    _, first = np.unique(key, return_index=True)
    first = np.sort(first)
    inds = [a[first] for a in inds]
    vals = (rng.random(len(first)) + 0.25).astype(np.float32)
    return inds, vals


def check_csf(fmt, ref, N, trailing=True):
    for l in range(N - 1):
        np.testing.assert_array_equal(fmt.export("fbrPtr", l), ref["fbrPtr"][l], err_msg=f"fbrPtr[{l}]")
        got = fmt.export("fbrIdx", l)
        exp = ref["fbrIdx"][l]
        assert len(got) == len(exp), f"fbrIdx[{l}] length"
        if trailing and l == N - 2:
            np.testing.assert_array_equal(got[:-1], exp[:-1], err_msg=f"fbrIdx[{l}]")
        else:
            np.testing.assert_array_equal(got, exp, err_msg=f"fbrIdx[{l}]")


def check_format(mmcsf, t, fmt_type, dims, inds, vals, mode, f):
    fmt = mmcsf.Format(t, fmt_type, mode=mode, fbr_threshold=f)
    ref = O.build_format(fmt_type, dims, inds, vals, mode=mode, fbr_threshold=f)
    info = fmt.info()
    N = len(dims)
    assert info["modeOrder"] == ref["modeOrder"]
    if fmt_type in (1, 3):
        for m in range(N):
            np.testing.assert_array_equal(fmt.export("inds", m), ref["keys"][ref["modeOrder"].index(m)])
        np.testing.assert_array_equal(fmt.export("vals"), ref["vals"])
        return fmt, ref
    if fmt_type in (8, 12, 13):
        check_csf(fmt, ref, N)
        assert info["nFibers"] == ref["nFibers"]
        np.testing.assert_array_equal(fmt.export("fbrLikeSlcInds")[:-1], ref["fbrLikeSlcInds"][:-1])
        assert len(fmt.export("fbrLikeSlcInds")) == ref["nFibers"]
        np.testing.assert_array_equal(fmt.export("inds", ref["modeOrder"][-1]), ref["leaf"])
        np.testing.assert_array_equal(fmt.export("vals"), ref["vals"])
        for m in range(N):  # the expanded tree restores the sorted COO
            np.testing.assert_array_equal(fmt.export("inds", m), ref["keys"][ref["modeOrder"].index(m)])
    if fmt_type == 8:
        for b in range(10):
            np.testing.assert_array_equal(fmt.export("slcMapperBin", b), ref["slcMapperBin"][b])
        assert info["droppedSlices"] == ref["dropped"]
    if fmt_type == 12:
        assert info["redunMode"] == ref["redunMode"]
        assert set(fmt.export("partPerNnz")) == {2}
    if fmt_type == 10:
        h = ref["hyb"]
        order = ref["modeOrder"]
        assert info["COOnnz"] == h["COOnnz"] and info["CSLnnz"] == h["CSLnnz"] and info["HCSRnnz"] == h["HCSRnnz"]
        for m in range(N):
            np.testing.assert_array_equal(fmt.export("COOinds", m), h["COOinds"][order.index(m)])
        np.testing.assert_array_equal(fmt.export("COOvals"), h["COOvals"])
        if N == 3:
            np.testing.assert_array_equal(fmt.export("CSLslicePtr"), h["CSLslicePtr"])
            np.testing.assert_array_equal(fmt.export("CSLsliceIdx"), h["CSLsliceIdx"])
            np.testing.assert_array_equal(fmt.export("CSLinds", order[1]), h["CSLinds1"])
            np.testing.assert_array_equal(fmt.export("CSLinds", order[2]), h["CSLinds2"])
            np.testing.assert_array_equal(fmt.export("CSLvals"), h["CSLvals"])
            for b in range(10):
                np.testing.assert_array_equal(fmt.export("CSLslcMapperBin", b), h["CSLslcMapperBin"][b])
        for l in range(N - 1):
            np.testing.assert_array_equal(fmt.export("fbrPtr", l), h["fbrPtr"][l], err_msg=f"hyb fbrPtr[{l}]")
            np.testing.assert_array_equal(fmt.export("fbrIdx", l), h["fbrIdx"][l], err_msg=f"hyb fbrIdx[{l}]")
        np.testing.assert_array_equal(fmt.export("inds", order[-1]), h["leaf"])
        np.testing.assert_array_equal(fmt.export("vals"), h["vals"])
        for b in range(10):
            np.testing.assert_array_equal(fmt.export("slcMapperBin", b), h["slcMapperBin"][b])
    return fmt, ref


def test_toy_all_types(cuda, golden_dir):
    import mmcsf
    t = mmcsf.Tensor.load_tns(f"{golden_dir}/toy.tns")
    dims, inds, vals = O.load_tns(f"{golden_dir}/toy.tns")
    for fmt_type in (1, 3, 8, 10, 12, 13):
        for mode in range(3):
            check_format(mmcsf, t, fmt_type, dims, inds, vals, mode, 128)
    # the tree each mode's MTTKRP traverses: its own root tree (-a 1), or the
    # MM-CSF tree for every mode (-a 0) and the one tree of -t 13
    for fm in (mmcsf.Format(t, 12, fbr_threshold=128), mmcsf.Format(t, 12, fbr_threshold=128, mode_trees=False),
               mmcsf.Format(t, 13, mode=1)):
        info = fm.info()
        for m in range(3):
            if info["modeTrees"]:
                assert info["modeTreeOrder"][m][0] == m
                assert sorted(info["modeTreeOrder"][m]) == [0, 1, 2]
            else:
                assert info["modeTreeOrder"][m] == info["modeOrder"]
    # the Appendix-B MM-CSF arrays, straight from the device
    fmt = mmcsf.Format(t, 12, fbr_threshold=128)
    assert list(fmt.export("fbrPtr", 0)) == [0, 2, 5, 7]
    assert list(fmt.export("fbrPtr", 1)) == [0, 2, 3, 4, 5, 6, 7, 8]
    assert list(fmt.export("fbrIdx", 1))[:-1] == [0, 1, 0, 1, 2, 1, 2]


CASES = [
    ([40, 30, 20], 4000, None),
    ([300, 40, 6], 8000, [1.3, 1.2, None]),
    ([500, 200, 50], 20000, [1.2, 1.2, None]),
    ([20, 25, 30, 15], 8000, None),
    ([60, 8, 30, 4], 9000, [1.2, None, 1.3, None]),
    ([100, 50, 20, 10], 30000, [1.1, None, 1.1, None]),
]


@pytest.mark.parametrize("dims,nnz,skew", CASES)
def test_random_tensors_all_types(cuda, dims, nnz, skew):
    import mmcsf
    rng = np.random.default_rng(nnz + len(dims))
    inds, vals = random_tensor(rng, dims, nnz, skew)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    N = len(dims)
    for f in (1, 2, 5, 128, 99999999):
        check_format(mmcsf, t, 12, dims, inds, vals, 0, f)
    for mode in range(N):
        for f in (3, 128):
            for fmt_type in (3, 8, 10, 13):
                check_format(mmcsf, t, fmt_type, dims, inds, vals, mode, f)


def test_single_nonzero_and_single_slice(cuda):
    import mmcsf
    dims = [5, 6, 7]
    for inds, vals in [([np.array([3]), np.array([2]), np.array([6])], np.array([1.5], np.float32)),
                       ([np.zeros(50, np.int64), np.arange(50) % 6, np.arange(50) // 6 % 7],
                        np.linspace(1, 2, 50).astype(np.float32))]:
        t = mmcsf.Tensor.from_numpy(dims, inds, vals)
        for fmt_type in (3, 8, 10, 12, 13):
            check_format(mmcsf, t, fmt_type, dims, inds, vals, 0, 4)


def test_dropped_heavy_slice_is_flagged(cuda):
    # one slice whose weight (fibers + nonzeros) exceeds totNnz + 1: the
    # reference's make_TiledBin puts it in no bin; the builder reports it
    import mmcsf
    n = 20000
    dims = [3, n, 2]
    inds = [np.zeros(n, np.int64), np.arange(n), np.zeros(n, np.int64)]
    vals = np.ones(n, np.float32)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    fmt, ref = check_format(mmcsf, t, 8, dims, inds, vals, 0, 128)
    assert ref["dropped"] == 1 and fmt.info()["droppedSlices"] == 1


def test_generator_properties(cuda):
    import mmcsf
    dims = [5000, 2000, 500]
    nnz = 200000
    a = mmcsf.Tensor.generate(dims, nnz, [1, 1, 0], [1.2, 1.2, 0.0], seed=7)
    b = mmcsf.Tensor.generate(dims, nnz, [1, 1, 0], [1.2, 1.2, 0.0], seed=7)
    c = mmcsf.Tensor.generate(dims, nnz, [1, 1, 0], [1.2, 1.2, 0.0], seed=8)
    da, ia, va = a.to_numpy()
    _, ib, vb = b.to_numpy()
    _, ic, _ = c.to_numpy()
    assert da == dims and len(va) == nnz
    for m in range(3):
        np.testing.assert_array_equal(ia[m], ib[m])
        assert ia[m].min() >= 0 and ia[m].max() < dims[m]
    np.testing.assert_array_equal(va, vb)
    assert not all(np.array_equal(ia[m], ic[m]) for m in range(3))
    key = (ia[0] * dims[1] + ia[1]) * dims[2] + ia[2]
    assert len(np.unique(key)) == nnz                      # distinct coordinates
    assert va.min() >= 0 and va.max() < 1
    # Zipf(1.2): index 0 is the most frequent, far above the uniform mode
    c0 = np.bincount(ia[0], minlength=dims[0])
    assert c0.argmax() == 0 and c0[0] > 20 * nnz / dims[0]
    c2 = np.bincount(ia[2], minlength=dims[2])
    assert c2.max() < 2 * nnz / dims[2]
    # the first draws are kept in draw order: a smaller request is a prefix
    d = mmcsf.Tensor.generate(dims, nnz // 2, [1, 1, 0], [1.2, 1.2, 0.0], seed=7)
    _, idd, vd = d.to_numpy()
    for m in range(3):
        np.testing.assert_array_equal(idd[m], ia[m][: nnz // 2])
    np.testing.assert_array_equal(vd, va[: nnz // 2])
