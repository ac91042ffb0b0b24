"""Pins the CPU oracle (oracle/) before any GPU result is compared with it.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong index or level, a transposed operand, an off-by-one in a split
rule) fails it:

* toy.tns known answers hand-derived from the reference's rules (MM-CSF and
  B-CSF arrays, MTTKRP with the reference's constant factors m+0.5);
* closed forms: mode orders of every benchmark shape, rank-1 factors, bins;
* brute force: float64 einsum over the dense tensor;
* an independent run/chunk formulation of the CSF builder (numpy), compared
  with the oracle's literal nonzero-walk on skewed random tensors;
* invariants: cross-mode inner-product identity, HB-CSF reconstruction.
"""
import numpy as np
import pytest

import oracle as O

NO_SPLIT = O.NO_SPLIT


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def random_tensor(rng, dims, nnz, skew=None, dup=False):
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
    if not dup:
        key = np.zeros(nnz, dtype=np.int64)
        for m, d in enumerate(dims):
            key = key * d + inds[m]
        _, first = np.unique(key, return_index=True)
        first = np.sort(first)
        inds = [a[first] for a in inds]
    vals = rng.random(len(inds[0])).astype(np.float32) + 0.5
    return inds, vals


def dense_of(dims, inds, vals):
    X = np.zeros(dims, dtype=np.float64)
    np.add.at(X, tuple(inds), vals.astype(np.float64))
    return X


def einsum_mttkrp(X, U, mode):
    N = X.ndim
    letters = "ijkl"[:N]
    ops = [X]
    spec = [letters]
    for m in range(N):
        if m != mode:
            ops.append(U[m].astype(np.float64))
            spec.append(letters[m] + "r")
    return np.einsum(",".join(spec) + "->" + letters[mode] + "r", *ops)


def csf_closed_form(keys, f, fs):
    """Independent statement of create_TiledHCSR: leaf fibers start at a key
    change of levels 0..N-2 or every f-th element of a run of equal keys;
    4-D fbrS entries start at a (k0,k1) change or, greedily, at the first leaf
    fiber whose fbrS already counts more than fs by the reference's counter
    (q - s for the tensor's first fbrS, q - s + 1 after a reset)."""
    K = np.stack([np.asarray(k, dtype=np.int64) for k in keys])
    N, n = K.shape
    d = np.full(n, N)
    d[0] = 0
    if n > 1:
        diff = K[:, 1:] != K[:, :-1]
        d[1:] = np.where(diff.any(axis=0), np.argmax(diff, axis=0), N)
    f = max(int(f), 1)
    pos = np.arange(n)
    run_start = np.maximum.accumulate(np.where(d <= N - 2, pos, 0))
    leaf = (d <= N - 2) | ((pos - run_start) % f == 0)
    lp = np.nonzero(leaf)[0]
    dF = d[lp]
    LL = N - 2
    ptr = [None] * (N - 1)
    idx = [None] * (N - 1)
    ptr[LL] = np.append(lp, n)
    idx[LL] = np.append(K[LL][lp], K[LL][lp[-1]])
    if N == 3:
        s = np.nonzero(dF == 0)[0]
        ptr[0] = np.append(s, len(lp))
        idx[0] = K[0][lp[s]]
    else:
        jflag = np.zeros(len(lp), dtype=bool)
        spos = 0
        for t in range(len(lp)):
            if dF[t] <= 1:
                jflag[t] = True
                spos = lp[t]
            elif lp[t] - spos + (1 if spos > 0 else 0) > fs:
                jflag[t] = True
                spos = lp[t]
        j = np.nonzero(jflag)[0]
        ptr[1] = np.append(j, len(lp))
        idx[1] = K[1][lp[j]]
        s = np.nonzero(dF == 0)[0]
        ptr[0] = np.append(np.searchsorted(j, s), len(j))
        idx[0] = K[0][lp[s]]
    return ptr, idx


def sorted_keys(dims, inds, order):
    perm = O.sort_perm(inds, order)
    return perm, [np.asarray(inds[m])[perm] for m in order]


# ---------------------------------------------------------------------------
# mode order
# ---------------------------------------------------------------------------
ORDER_CASES = [
    ([3, 3, 3], {0: [0, 1, 2], 1: [1, 0, 2], 2: [2, 0, 1]}),
    ([2000, 2000, 2000], {0: [0, 1, 2], 2: [2, 0, 1]}),
    ([500000, 200000, 50000], {0: [0, 2, 1], 1: [1, 2, 0], 2: [2, 1, 0]}),
    ([100000, 50000, 20000, 1000], {0: [0, 3, 2, 1], 1: [1, 3, 2, 0], 2: [2, 3, 1, 0], 3: [3, 2, 1, 0]}),
    ([200000, 100000, 20000], {2: [2, 1, 0]}),
    ([1000000, 500000, 100000], {2: [2, 1, 0]}),
    ([5, 7, 5, 7], {0: [0, 2, 1, 3], 1: [1, 0, 2, 3], 3: [3, 0, 2, 1]}),
]


@pytest.mark.parametrize("dims,expect", ORDER_CASES)
def test_order_tensormode_closed_form(dims, expect):
    for mode, order in expect.items():
        assert O.order_tensormode(dims, mode) == order
        # independent statement: the other modes by (length, index)
        rest = sorted((m for m in range(len(dims)) if m != mode), key=lambda m: (dims[m], m))
        assert O.order_tensormode(dims, mode) == [mode] + rest
        assert O.order_tensormode(dims, mode, nat=True) == [(mode + i) % len(dims) for i in range(len(dims))]


def test_redundant_mode():
    assert O.redundant_mode([3, 3, 3]) == 1
    assert O.redundant_mode([500000, 200000, 50000]) == 2
    assert O.redundant_mode([100000, 50000, 20000, 1000]) == 3


# ---------------------------------------------------------------------------
# toy.tns known answers
# ---------------------------------------------------------------------------
def test_toy_load(golden_dir):
    dims, inds, vals = O.load_tns(f"{golden_dir}/toy.tns")
    assert dims == [3, 3, 3]
    assert [list(a) for a in inds] == [[0, 0, 0, 1, 1, 1, 2, 2], [0, 1, 2, 0, 2, 2, 0, 1],
                                       [0, 1, 0, 2, 0, 1, 2, 1]]
    assert list(vals) == [1, 2, 10, 7, 6, 5, 3, 11]


def test_toy_mmcsf_arrays(golden_dir):
    dims, inds, vals = O.load_tns(f"{golden_dir}/toy.tns")
    f = O.build_format(12, dims, inds, vals, fbr_threshold=128)
    assert f["modeOrder"] == [2, 0, 1]
    assert f["redunMode"] == 1
    assert list(f["fbrIdx"][0]) == [0, 1, 2]
    assert list(f["fbrPtr"][0]) == [0, 2, 5, 7]
    assert list(f["fbrIdx"][1][:-1]) == [0, 1, 0, 1, 2, 1, 2]
    assert len(f["fbrIdx"][1]) == 8                      # trailing entry
    assert list(f["fbrPtr"][1]) == [0, 2, 3, 4, 5, 6, 7, 8]
    assert f["nFibers"] == 8
    assert list(f["fbrLikeSlcInds"][:-1]) == [0, 0, 1, 1, 1, 2, 2]
    # leaves per fiber: {0,2} | 2 | 1 | 2 | 1 | 0 | 0 with values {1,10} | 6 | 2 | 5 | 11 | 7 | 3
    assert list(f["leaf"]) == [0, 2, 2, 1, 2, 1, 0, 0]
    assert list(f["vals"]) == [1, 10, 6, 2, 5, 11, 7, 3]
    assert set(f["partPerNnz"]) == {2}


def test_toy_bcsf_arrays(golden_dir):
    dims, inds, vals = O.load_tns(f"{golden_dir}/toy.tns")
    b = O.build_format(8, dims, inds, vals, mode=0, fbr_threshold=128)
    assert b["modeOrder"] == [0, 1, 2]
    assert list(b["fbrIdx"][0]) == [0, 1, 2]
    assert list(b["fbrPtr"][0]) == [0, 3, 5, 7]
    assert list(b["fbrIdx"][1][:-1]) == [0, 1, 2, 0, 2, 0, 1]
    assert list(b["fbrPtr"][1]) == [0, 1, 2, 3, 4, 6, 7, 8]
    assert list(b["weight"]) == [6, 5, 4]
    assert list(b["binOf"]) == [0, 0, 0]
    assert b["dropped"] == 0


def test_toy_single_csf_arrays(golden_dir):
    # -t 13 on the toy at -m 0: the hand-derived create_HCSR tree of the B-CSF
    # test above, fbrLikeSlcInds = the slice of each fiber, no bins
    dims, inds, vals = O.load_tns(f"{golden_dir}/toy.tns")
    s = O.build_format(13, dims, inds, vals, mode=0, fbr_threshold=128)
    assert s["modeOrder"] == [0, 1, 2]
    assert list(s["fbrPtr"][0]) == [0, 3, 5, 7]
    assert list(s["fbrPtr"][1]) == [0, 1, 2, 3, 4, 6, 7, 8]
    assert list(s["fbrIdx"][1][:-1]) == [0, 1, 2, 0, 2, 0, 1]
    assert list(s["fbrLikeSlcInds"][:-1]) == [0, 0, 0, 1, 1, 2, 2]
    assert "slcMapperBin" not in s
    # -f 1 splits the two-nonzero fiber (slice 1, index 2) in two
    s1 = O.build_format(13, dims, inds, vals, mode=0, fbr_threshold=1)
    assert list(s1["fbrPtr"][1]) == [0, 1, 2, 3, 4, 5, 6, 7, 8]
    assert list(s1["fbrPtr"][0]) == [0, 3, 6, 8]


def test_toy_hyb_arrays(golden_dir):
    # -m 0: slices 0 and 2 have only length-1 fibers (CSL), slice 1 is CSF
    dims, inds, vals = O.load_tns(f"{golden_dir}/toy.tns")
    h = O.build_format(10, dims, inds, vals, mode=0, fbr_threshold=128)["hyb"]
    assert h["COOnnz"] == 0 and h["CSLnnz"] == 5 and h["HCSRnnz"] == 3
    assert list(h["CSLslicePtr"]) == [0, 3, 5]
    assert list(h["CSLsliceIdx"]) == [0, 2]
    assert list(h["CSLinds1"]) == [0, 1, 2, 0, 1]
    assert list(h["CSLinds2"]) == [0, 1, 0, 2, 1]
    assert list(h["CSLvals"]) == [1, 2, 10, 3, 11]
    assert list(h["fbrPtr"][0]) == [0, 2]
    assert list(h["fbrIdx"][0]) == [1]
    assert list(h["fbrPtr"][1]) == [0, 1, 3]
    assert list(h["fbrIdx"][1]) == [0, 2]
    assert list(h["leaf"]) == [2, 0, 1]
    assert list(h["vals"]) == [7, 6, 5]
    assert h["nFibers"] == 3


TOY_EXPECT = {0: [48.75, 67.5, 52.5], 1: [13.75, 16.25, 26.25], 2: [12.75, 13.5, 7.5]}


@pytest.mark.parametrize("R", [1, 5, 32])
def test_toy_mttkrp_closed_form(golden_dir, R):
    dims, inds, vals = O.load_tns(f"{golden_dir}/toy.tns")
    U = [np.full((d, R), m + 0.5, dtype=np.float32) for m, d in enumerate(dims)]
    for mode, rows in TOY_EXPECT.items():
        out = np.zeros((dims[mode], R), dtype=np.float32)
        O.mttkrp_coo_f32(inds, vals, mode, U, out)
        np.testing.assert_array_equal(out, np.repeat(np.array(rows, np.float32)[:, None], R, 1))
        o64 = O.mttkrp_coo_f64(inds, vals, mode, U)
        np.testing.assert_array_equal(o64, np.repeat(np.array(rows)[:, None], R, 1))
        # the raw -t 1 run does not zero U[mode] (mttkrp.cu:88-89)
        raw = U[mode].copy()
        O.mttkrp_coo_f32(inds, vals, mode, U, raw)
        np.testing.assert_allclose(raw, np.repeat(np.array(rows)[:, None] + mode + 0.5, R, 1))


# ---------------------------------------------------------------------------
# MTTKRP against brute force and closed forms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dims", [[7, 5, 3], [4, 9, 6], [5, 3, 4, 6]])
def test_mttkrp_matches_dense_einsum(dims):
    rng = np.random.default_rng(len(dims) * 100 + dims[0])
    nnz = int(np.prod(dims) * 0.4)
    inds, vals = random_tensor(rng, dims, nnz)
    R = 6
    U = [rng.standard_normal((d, R)).astype(np.float32) for d in dims]
    X = dense_of(dims, inds, vals)
    for mode in range(len(dims)):
        ref = einsum_mttkrp(X, U, mode)
        got = O.mttkrp_coo_f64(inds, vals, mode, U)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
        g32 = np.zeros((dims[mode], R), dtype=np.float32)
        O.mttkrp_coo_f32(inds, vals, mode, U, g32)
        np.testing.assert_allclose(g32, ref, rtol=2e-5, atol=2e-5)


def test_mttkrp_rank1_closed_form():
    # U_m = a_m b_m^T with dyadic entries (exact in float32):
    # M_n[i][r] = (sum_{x: i_n = i} v_x prod_{m != n} a_m[i_m]) * prod_{m != n} b_m[r]
    rng = np.random.default_rng(3)
    dims = [9, 8, 7]
    inds, vals = random_tensor(rng, dims, 150)
    R = 4
    a = [rng.integers(1, 16, d) / 8.0 for d in dims]
    b = [rng.integers(1, 8, R) / 4.0 for _ in dims]
    U = [np.outer(a[m], b[m]).astype(np.float32) for m in range(3)]
    for mode in range(3):
        others = [m for m in range(3) if m != mode]
        w = vals.astype(np.float64).copy()
        for m in others:
            w *= a[m][inds[m]]
        rowsum = np.bincount(inds[mode], weights=w, minlength=dims[mode])
        expect = rowsum[:, None] * np.prod([b[m] for m in others], axis=0)[None, :]
        np.testing.assert_allclose(O.mttkrp_coo_f64(inds, vals, mode, U), expect, rtol=1e-12)


def test_cross_mode_inner_product_identity():
    rng = np.random.default_rng(11)
    for dims in ([30, 20, 10], [12, 9, 8, 7]):
        inds, vals = random_tensor(rng, dims, 400)
        U = [rng.standard_normal((d, 5)).astype(np.float32) for d in dims]
        ips = [float(np.sum(O.mttkrp_coo_f64(inds, vals, m, U) * U[m])) for m in range(len(dims))]
        np.testing.assert_allclose(ips, ips[0], rtol=1e-10)


def test_f32_sequential_error_short_rows():
    # uniform 200^3 with ~25 nonzeros per row: the sequential float sum is
    # accurate (well below 1e-6 relative)
    rng = np.random.default_rng(5)
    dims = [200, 200, 200]
    inds, vals = random_tensor(rng, dims, 5000)
    U = [rng.random((d, 32)).astype(np.float32) for d in dims]
    for mode in range(3):
        r64 = O.mttkrp_coo_f64(inds, vals, mode, U)
        r32 = np.zeros_like(r64, dtype=np.float32)
        O.mttkrp_coo_f32(inds, vals, mode, U, r32)
        assert np.linalg.norm(r32 - r64) / np.linalg.norm(r64) < 1e-6


# ---------------------------------------------------------------------------
# CSF builder
# ---------------------------------------------------------------------------
BUILDER_CASES = [
    ([40, 30, 20], 3000, None, [1, 2, 3, 7, 128, 99999999, 0]),
    ([300, 40, 6], 6000, [1.3, 1.2, None], [1, 2, 5, 128]),
    ([20, 25, 30, 15], 6000, None, [1, 3, 128]),
    ([60, 8, 30, 4], 9000, [1.2, None, 1.3, None], [1, 2, 4, 16, 99999999]),
]


@pytest.mark.parametrize("dims,nnz,skew,fs", BUILDER_CASES)
def test_csf_builder_matches_closed_form(dims, nnz, skew, fs):
    rng = np.random.default_rng(nnz)
    inds, vals = random_tensor(rng, dims, nnz, skew)
    N = len(dims)
    for order in (O.order_tensormode(dims, 0), O.order_tensormode(dims, 2)):
        _, keys = sorted_keys(dims, inds, order)
        for f in fs:
            for fsth in ([NO_SPLIT] if N == 3 else [NO_SPLIT, 128, 7, 1]):
                ptr, idx, nf = O.build_csf(keys, f, fsth)
                eptr, eidx = csf_closed_form(keys, f, fsth)
                for l in range(N - 1):
                    np.testing.assert_array_equal(ptr[l], eptr[l], err_msg=f"fbrPtr[{l}] f={f} fs={fsth}")
                    if l == N - 2:  # trailing entry is unspecified
                        np.testing.assert_array_equal(idx[l][:-1], eidx[l][:-1])
                        assert len(idx[l]) == len(eidx[l])
                    else:
                        np.testing.assert_array_equal(idx[l], eidx[l])
                assert nf == len(ptr[1])


def test_csf_builder_invariants():
    rng = np.random.default_rng(17)
    dims = [50, 4, 40]
    inds, vals = random_tensor(rng, dims, 5000, [1.1, None, 1.2])
    order = O.order_tensormode(dims, 1)
    _, keys = sorted_keys(dims, inds, order)
    for f in (1, 2, 3, 10):
        ptr, idx, nf = O.build_csf(keys, f, NO_SPLIT)
        lens = np.diff(ptr[1])
        assert lens.min() >= 1 and lens.max() <= f
        assert ptr[1][0] == 0 and ptr[1][-1] == len(vals)
        assert np.all(np.diff(ptr[0]) >= 1)
        # chunk structure: each run of equal (k0,k1) holds ceil(len/f) leaf fibers
        pair = keys[0].astype(np.int64) * 10**6 + keys[1]
        _, run_len = np.unique(pair, return_counts=True)
        assert len(lens) == int(np.sum(-(-run_len // f)))


def test_fbrs_split_first_fbrs_off_by_one():
    # one (k0,k1) run of 200 leaf fibers with one nonzero each: the first fbrS
    # of the tensor closes after 129 nonzeros, the next one after 128
    n = 200
    keys = [np.zeros(n, np.uint64), np.zeros(n, np.uint64), np.arange(n, dtype=np.uint64),
            np.zeros(n, np.uint64)]
    ptr, idx, nf = O.build_csf(keys, 128, 128)
    assert list(ptr[1]) == [0, 129, 200]
    # a second run starting later counts one more
    keys2 = [np.r_[np.zeros(1), np.ones(n)].astype(np.uint64), np.zeros(n + 1, np.uint64),
             np.r_[0, np.arange(n)].astype(np.uint64), np.zeros(n + 1, np.uint64)]
    ptr2, _, _ = O.build_csf(keys2, 128, 128)
    assert list(ptr2[1]) == [0, 1, 129, 201]


def test_fbr_like_slc_inds_definition():
    rng = np.random.default_rng(2)
    dims = [30, 20, 10, 5]
    inds, vals = random_tensor(rng, dims, 1500)
    f = O.build_format(12, dims, inds, vals, fbr_threshold=4)
    ptr, idx = f["fbrPtr"], f["fbrIdx"]
    fl = f["fbrLikeSlcInds"]
    for s in range(len(idx[0])):
        assert np.all(fl[ptr[0][s]:ptr[0][s + 1]] == idx[0][s])


# ---------------------------------------------------------------------------
# bins
# ---------------------------------------------------------------------------
def closed_form_bin(w, totnnz):
    if 0 < w < 1025:
        return 0
    table = [(1024, 2049, 5), (2048, 4097, 6), (4096, 8193, 7), (8192, 16385, 8), (16384, totnnz + 1, 9)]
    for lo, hi, b in table:
        if lo < w < hi:
            return b
    return -1


def test_tiled_bin_closed_form():
    # 3-D slices with 1 fiber each and w-1 nonzeros, so weight = w
    weights = [2, 3, 1024, 1025, 2048, 2049, 4096, 4097, 8192, 8193, 16384, 16385, 20000]
    nnz_per = [w - 1 for w in weights]
    total = int(sum(nnz_per))
    p1 = np.concatenate([[0], np.cumsum(nnz_per)]).astype(np.uint64)
    p0 = np.arange(len(weights) + 1, dtype=np.uint64)
    bins, w, dropped = O.tiled_bin(3, total, [p0, p1])
    assert list(w) == weights
    assert [int(b) for b in bins] == [closed_form_bin(x, total) for x in weights]
    # a slice heavier than totNnz+1 is dropped by the reference's binning
    # (weight = fibers + nonzeros can exceed totNnz + 1 once it is above 16384)
    n = 20000
    bins2, w2, dropped2 = O.tiled_bin(3, n, [np.array([0, n], np.uint64), np.arange(n + 1, dtype=np.uint64)])
    assert list(w2) == [2 * n] and list(bins2) == [-1] and dropped2 == 1


def test_hyb_bin_uses_nnz_only():
    p0 = np.array([0, 2, 3], np.uint64)
    p1 = np.array([0, 600, 1100, 1200], np.uint64)
    bins, _ = O.hyb_bin(3, 1200, [p0, p1])
    assert [int(b) for b in bins] == [closed_form_bin(1100, 1200), closed_form_bin(100, 1200)]


# ---------------------------------------------------------------------------
# HB-CSF
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dims", [[40, 50, 30], [30, 20, 10, 8]])
def test_hyb_reconstructs_sorted_coo(dims):
    rng = np.random.default_rng(sum(dims))
    inds, vals = random_tensor(rng, dims, 900, [1.2] + [None] * (len(dims) - 1))
    # add single-nonzero slices and all-length-1 slices
    f = O.build_format(10, dims, inds, vals, mode=0, fbr_threshold=3)
    h = f["hyb"]
    N = len(dims)
    order = f["modeOrder"]
    # class rules
    p = f["fbrPtr"]
    recon = []
    for c in range(h["COOnnz"]):
        recon.append(tuple(int(h["COOinds"][l][c]) for l in range(N)) + (float(h["COOvals"][c]),))
    for s in range(len(h["CSLsliceIdx"])):
        for x in range(int(h["CSLslicePtr"][s]), int(h["CSLslicePtr"][s + 1])):
            recon.append((int(h["CSLsliceIdx"][s]), int(h["CSLinds1"][x]), int(h["CSLinds2"][x]),
                          float(h["CSLvals"][x])))
    hp, hi = h["fbrPtr"], h["fbrIdx"]
    for s in range(len(hi[0])):
        for j in range(int(hp[0][s]), int(hp[0][s + 1])):
            if N == 3:
                for x in range(int(hp[1][j]), int(hp[1][j + 1])):
                    recon.append((int(hi[0][s]), int(hi[1][j]), int(h["leaf"][x]), float(h["vals"][x])))
            else:
                for g in range(int(hp[1][j]), int(hp[1][j + 1])):
                    for x in range(int(hp[2][g]), int(hp[2][g + 1])):
                        recon.append((int(hi[0][s]), int(hi[1][j]), int(hi[2][g]), int(h["leaf"][x]),
                                      float(h["vals"][x])))
    expect = [tuple(int(k[x]) for k in f["keys"]) + (float(f["vals"][x]),) for x in range(len(vals))]
    assert sorted(recon) == sorted(expect)
    assert h["COOnnz"] + h["CSLnnz"] + h["HCSRnnz"] == len(vals)
    if N == 4:
        assert h["CSLnnz"] == 0


# ---------------------------------------------------------------------------
# float64 CP-ALS and CP-completion restatements (oracle/dense.py)
# ---------------------------------------------------------------------------
def _dense_model(U, lam):
    letters = "ijkl"[: len(U)]
    spec = ",".join(c + "r" for c in letters) + ",r->" + letters
    return np.einsum(spec, *U, lam)


def test_cpd_oracle_fit_matches_dense_residual():
    from oracle import dense as D
    rng = np.random.default_rng(4)
    dims = [12, 10, 8]
    inds, vals = random_tensor(rng, dims, 500)
    U0 = [rng.random((d, 3)) for d in dims]
    fits, U, lam = D.cpd_als(inds, vals, dims, U0, iters=4, tol=0.0)
    X = dense_of(dims, inds, vals)
    model = _dense_model(U, lam)
    fit_dense = 1 - np.linalg.norm(X - model) / np.linalg.norm(X)
    np.testing.assert_allclose(fits[-1], fit_dense, rtol=1e-10, atol=1e-12)


def test_cpd_oracle_recovers_planted_rank():
    from oracle import dense as D
    rng = np.random.default_rng(8)
    dims = [9, 8, 7]
    A = [rng.standard_normal((d, 2)) for d in dims]
    X = _dense_model(A, np.ones(2))
    inds = [a.ravel() for a in np.indices(dims)]
    vals = X.ravel().astype(np.float32)
    fits, U, lam = D.cpd_als(inds, vals, dims, [rng.random((d, 2)) for d in dims], iters=200, tol=1e-12)
    assert fits[-1] > 0.99999
    # ALS never decreases the fit (beyond round-off) once the iterations use max-norm scaling
    assert all(b >= a - 1e-9 for a, b in zip(fits[1:], fits[2:]))


def test_cp_grad_matches_finite_differences():
    from oracle import dense as D
    rng = np.random.default_rng(9)
    dims = [6, 5, 4]
    inds, vals = random_tensor(rng, dims, 60)
    U = [rng.standard_normal((d, 3)) for d in dims]
    g = D.cp_grad(inds, vals, U)
    h = 1e-6
    for m in range(3):
        for (i, r) in [(0, 0), (2, 1), (dims[m] - 1, 2)]:
            Up = [u.copy() for u in U]
            Um = [u.copy() for u in U]
            Up[m][i, r] += h
            Um[m][i, r] -= h
            fd = (D.cp_loss(inds, vals, Up)[0] - D.cp_loss(inds, vals, Um)[0]) / (2 * h)
            np.testing.assert_allclose(g[m][i, r], fd, rtol=1e-6, atol=1e-8)


# ---------------------------------------------------------------------------
# TTM restatement (oracle/dense.py ttm64)
# ---------------------------------------------------------------------------
def test_ttm_toy_closed_form(golden_dir):
    """Constant factors m + 0.5 (randomize_mats): every fiber row is the sum
    of its values times (leaf mode + 0.5)."""
    from oracle import dense as D
    dims, inds, vals = O.load_tns(f"{golden_dir}/toy.tns")
    b = O.build_format(8, dims, inds, vals, mode=0, fbr_threshold=128)
    leaf_mode = b["modeOrder"][-1]
    R = 3
    U = np.full((dims[leaf_mode], R), leaf_mode + 0.5)
    Y = D.ttm64(b["fbrPtr"][1], b["leaf"], b["vals"], U)
    ptr = b["fbrPtr"][1]
    for f in range(len(ptr) - 1):
        s = float(np.sum(b["vals"][ptr[f]:ptr[f + 1]], dtype=np.float64))
        np.testing.assert_allclose(Y[f], s * (leaf_mode + 0.5), rtol=1e-12)


@pytest.mark.parametrize("dims", [[30, 20, 10], [12, 9, 8, 7]])
def test_ttm_then_upper_levels_give_root_mttkrp(dims):
    """MTTKRP of the root mode = per slice, the sum over its leaf fibers of
    the TTM row times the rows of the levels above (a closed identity that ties
    TTM to the pinned MTTKRP restatement)."""
    from oracle import dense as D
    rng = np.random.default_rng(len(dims))
    inds, vals = random_tensor(rng, dims, 1500)
    for fthr in (2, 128):
        b = O.build_format(8, dims, inds, vals, mode=0, fbr_threshold=fthr)
        order = b["modeOrder"]
        N = len(dims)
        U = [rng.standard_normal((d, 4)) for d in dims]
        Y = D.ttm64(b["fbrPtr"][N - 2], b["leaf"], b["vals"], U[order[-1]])
        # walk the tree: leaf fibers -> their ancestors
        ptr, idx = b["fbrPtr"], b["fbrIdx"]
        out = np.zeros((dims[order[0]], 4))
        for s in range(len(idx[0])):
            for j in range(ptr[0][s], ptr[0][s + 1]):
                if N == 3:
                    out[idx[0][s]] += Y[j] * U[order[1]][idx[1][j]]
                else:
                    for f in range(ptr[1][j], ptr[1][j + 1]):
                        out[idx[0][s]] += Y[f] * U[order[2]][idx[2][f]] * U[order[1]][idx[1][j]]
        ref = O.mttkrp_coo_f64(inds, vals, order[0], [u.astype(np.float32) for u in U])
        np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-5)


# ---------------------------------------------------------------------------
# threaded float64 oracle (used at the full benchmark shapes)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dims,skew", [([300, 200, 100], [1.3, 1.2, None]), ([50, 40, 30, 20], [1.2, None, 1.1, None])])
def test_threaded_mttkrp_equals_sequential_bit_for_bit(dims, skew):
    rng = np.random.default_rng(len(dims))
    inds, vals = random_tensor(rng, dims, 60000, skew)
    U = [rng.standard_normal((d, 12)).astype(np.float32) for d in dims]
    for mode in range(len(dims)):
        seq = O.mttkrp_coo_f64(inds, vals, mode, U)
        for T in (1, 2, 5, 16, 64):
            np.testing.assert_array_equal(O.mttkrp_f64_mt(inds, vals, mode, U, threads=T), seq)
    # float64 values (residuals) take the same path: a float64 copy of the
    # float32 values gives the same result
    np.testing.assert_array_equal(O.mttkrp_f64_mt(inds, vals.astype(np.float64), 0, U),
                                  O.mttkrp_coo_f64(inds, vals, 0, U))


def test_threaded_mttkrp_toy_closed_form(golden_dir):
    dims, inds, vals = O.load_tns(f"{golden_dir}/toy.tns")
    U = [np.full((d, 3), m + 0.5, dtype=np.float32) for m, d in enumerate(dims)]
    for mode, rows in TOY_EXPECT.items():
        np.testing.assert_array_equal(O.mttkrp_f64_mt(inds, vals, mode, U, threads=4),
                                      np.repeat(np.array(rows)[:, None], 3, 1))


def test_threaded_mttkrp_empty_and_single_row():
    U = [np.ones((d, 4), np.float32) for d in (5, 6, 7)]
    empty = [np.zeros(0, np.int64)] * 3
    assert not O.mttkrp_f64_mt(empty, np.zeros(0, np.float32), 0, U, threads=8).any()
    inds = [np.full(1000, 3), np.arange(1000) % 6, np.arange(1000) % 7]
    out = O.mttkrp_f64_mt(inds, np.ones(1000, np.float32), 0, U, threads=8)
    assert out[3, 0] == 1000.0 and out.sum() == 4000.0


def test_threaded_cp_residuals_match_numpy():
    from oracle import dense as D
    rng = np.random.default_rng(9)
    for dims in ([40, 30, 20], [12, 9, 8, 7]):
        inds, vals = random_tensor(rng, dims, 3000)
        U = [rng.standard_normal((d, 6)) for d in dims]
        L, e = O.cp_resid_f64_mt(inds, vals, U, threads=7)
        L64, e64 = D.cp_loss(inds, vals, U)
        np.testing.assert_allclose(e, -e64, rtol=0, atol=1e-12)
        assert abs(L - L64) <= 1e-12 * L64
    # an exact planted model has zero residual
    dims = [10, 9, 8]
    inds, _ = random_tensor(rng, dims, 300)
    A = [rng.integers(-4, 5, (d, 3)) / 4.0 for d in dims]
    v = D.cp_model_values(inds, A).astype(np.float32)   # exact in float32 (dyadic)
    L, e = O.cp_resid_f64_mt(inds, v, A)
    assert L == 0.0 and not e.any()


# ---------------------------------------------------------------------------
# MM-CSF mixed-mode partition (mm_partition_reuseBased without the override),
# pinned by hand-derived cases
# ---------------------------------------------------------------------------
def test_mm_partition_toy_hand_derived(golden_dir):
    """toy.tns, -h 1: orders [0,1,2], [1,0,2], [2,0,1]; mode 1 is redundant
    (its fibers copy mode 0's), so a nonzero goes to mode 0 (shortMode) when
    its (i,j) fiber holds at least as many nonzeros as its (k,i) fiber, else
    to mode 2.  Walk order (k,i,j): tensor positions 0,2,4,1,5,7,3,6.  The
    sixth and eighth nonzeros sit in the last slice of a mode's CSF, whose
    denseSlcPtr end equals its start: their fibers resolve to the entry
    before that slice (counts 2 and 0)."""
    dims, inds, vals = O.load_tns(f"{golden_dir}/toy.tns")
    r = O.mm_partition(dims, inds, vals)
    assert r["redunMode"] == 1
    np.testing.assert_array_equal(r["order"], [0, 2, 4, 1, 5, 7, 3, 6])
    np.testing.assert_array_equal(r["part"], [2, 2, 0, 0, 0, 0, 0, 0])
    p0, p1, p2 = r["partitions"]
    assert p1 is None
    assert p0["modeOrder"] == [0, 1, 2] and p0["nnz"] == 6
    np.testing.assert_array_equal(p0["fbrPtr"][0], [0, 1, 3, 5])
    np.testing.assert_array_equal(p0["fbrIdx"][0], [0, 1, 2])
    np.testing.assert_array_equal(p0["fbrPtr"][1], [0, 1, 2, 4, 5, 6])
    np.testing.assert_array_equal(p0["fbrIdx"][1][:-1], [1, 0, 2, 0, 1])
    np.testing.assert_array_equal(p0["keys"][2], [1, 2, 0, 1, 2, 1])
    assert p2["modeOrder"] == [2, 0, 1] and p2["nnz"] == 2
    np.testing.assert_array_equal(p2["fbrPtr"][0], [0, 1])
    np.testing.assert_array_equal(p2["fbrPtr"][1], [0, 2])
    np.testing.assert_array_equal(p2["keys"][2], [0, 2])
    np.testing.assert_array_equal(p2["vals"], [1, 10])


def test_mm_partition_decrement_changes_a_later_choice():
    """3x3x3, nonzeros A=(0,0,0), B=(0,0,1), C=(0,1,0), G=(2,2,2).  Walk
    order A, C, B, G.  A: fibers (i,j)=(0,0) of 2 and (k,i)=(0,0) of 2 ->
    mode 0, and the (k,i) count drops to 1.  C: (0,1) of 1 against that 1 ->
    mode 0 (with the original count 2 it would be mode 2).  B and G: mode 0."""
    dims = [3, 3, 3]
    inds = [np.array([0, 0, 0, 2]), np.array([0, 0, 1, 2]), np.array([0, 1, 0, 2])]
    vals = np.ones(4, np.float32)
    r = O.mm_partition(dims, inds, vals)
    np.testing.assert_array_equal(r["order"], [0, 2, 1, 3])
    np.testing.assert_array_equal(r["part"], [0, 0, 0, 0])


def test_mm_partition_long_fiber_of_the_last_mode_wins():
    """3x3x3, nonzeros (0,j,0) for j=0..2 and G=(2,2,2): each (i,j) fiber holds
    one nonzero, the (k,i)=(0,0) fiber three -> mode 2 for all three (choosing
    mode 2 decrements only mode 0's counts); G -> mode 0."""
    dims = [3, 3, 3]
    inds = [np.array([0, 0, 0, 2]), np.array([0, 1, 2, 2]), np.array([0, 0, 0, 2])]
    r = O.mm_partition(dims, inds, np.ones(4, np.float32))
    np.testing.assert_array_equal(r["order"], [0, 1, 2, 3])
    np.testing.assert_array_equal(r["part"], [2, 2, 2, 0])
    assert r["partitions"][2]["nnz"] == 3 and r["partitions"][0]["nnz"] == 1


@pytest.mark.parametrize("dims,skew,fbth", [([60, 40, 30], [1.3, 1.2, None], 1), ([60, 40, 30], None, 2),
                                             ([30, 20, 25, 12], [1.2, None, 1.1, None], 1)])
def test_mm_partition_partitions_cover_the_tensor(dims, skew, fbth):
    rng = np.random.default_rng(len(dims) + fbth)
    inds, vals = random_tensor(rng, dims, 4000, skew)
    r = O.mm_partition(dims, inds, vals, fb_th=fbth, fbr_threshold=8)
    N = len(dims)
    assert set(r["part"].tolist()) <= set(range(N))
    got = []
    for p, pp in enumerate(r["partitions"]):
        if pp is None:
            assert not (r["part"] == p).any()
            continue
        order = pp["modeOrder"]
        assert pp["nnz"] == int((r["part"] == p).sum())
        coords = np.stack([pp["keys"][order.index(m)] for m in range(N)], 1)
        got += [tuple(c) + (v,) for c, v in zip(coords.tolist(), pp["vals"].tolist())]
    want = [tuple(int(a[x]) for a in inds) + (float(vals[x]),) for x in range(len(vals))]
    assert sorted(got) == sorted(want)
