"""GPU MTTKRP (COO / B-CSF / HB-CSF / MM-CSF, 3-D and 4-D, FP32 and BF16
factor storage) against the CPU oracle.

Tolerances (written per test):
* vs the float64 oracle: ||M - M64||_F / ||M64||_F <= 1e-5 with float32
  factors (the north star's float32 bound), and never more than 1e-6 above
  the error of the reference's own sequential float32 sum (-t 1);
* vs the reference's sequential float32 COO CPU MTTKRP (-t 1) on the
  uniform 2000^3 / 1M-nonzero / R=32 workload: <= 1e-5 relative;
* BF16 storage: <= 1e-3 relative against float64 of the float32 factors
  (quantisation included), <= 1e-5 against float64 of the bf16 factors.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

from test_gpu_builder import random_tensor  # noqa: E402


def rel(a, b):
    b = np.asarray(b, dtype=np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / (n if n > 0 else 1.0))


def factors(rng, dims, R, torch, dev):
    U = [rng.random((d, R)).astype(np.float32) for d in dims]
    return U, [torch.from_numpy(u).to(dev) for u in U]


def bf16_round(U, torch):
    return [torch.from_numpy(u).to(torch.bfloat16).to(torch.float32).numpy() for u in U]


def test_toy_closed_form_every_type(cuda, golden_dir):
    import torch
    import mmcsf
    t = mmcsf.Tensor.load_tns(f"{golden_dir}/toy.tns")
    expect = {0: [48.75, 67.5, 52.5], 1: [13.75, 16.25, 26.25], 2: [12.75, 13.5, 7.5]}
    for R in (1, 4, 32, 33):
        Ud = [torch.full((3, R), m + 0.5, device=cuda) for m in range(3)]
        fm = mmcsf.Format(t, 12, fbr_threshold=128)
        outs, _ = mmcsf.mttkrp_allmode(fm, Ud)
        for mode in range(3):
            np.testing.assert_array_equal(outs[mode].cpu().numpy(), np.repeat(np.array(expect[mode], np.float32)[:, None], R, 1))
            for ft in (3, 8, 10):
                f = mmcsf.Format(t, ft, mode=mode, fbr_threshold=128)
                o = mmcsf.mttkrp(f, mode, Ud).cpu().numpy()
                np.testing.assert_array_equal(o, np.repeat(np.array(expect[mode], np.float32)[:, None], R, 1))


CASES = [
    ([40, 30, 20], 4000, None),
    ([500, 200, 50], 30000, [1.2, 1.2, None]),
    ([300, 40, 6], 8000, [1.3, 1.2, None]),
    ([20, 25, 30, 15], 8000, None),
    ([100, 50, 20, 10], 40000, [1.1, None, 1.1, None]),
]


@pytest.mark.parametrize("dims,nnz,skew", CASES)
@pytest.mark.parametrize("R", [1, 3, 8, 16, 32, 64, 100, 256])
def test_all_types_vs_float64(cuda, dims, nnz, skew, R):
    import torch
    import mmcsf
    rng = np.random.default_rng(nnz + R)
    inds, vals = random_tensor(rng, dims, nnz, skew)
    U, Ud = factors(rng, dims, R, torch, cuda)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    N = len(dims)
    ref = [O.mttkrp_coo_f64(inds, vals, m, U) for m in range(N)]
    tol = []
    for m in range(N):
        # error of the reference's -t 1 sum (sorted like sort_COOtensor)
        order = O.order_tensormode(dims, m)
        perm = O.sort_perm(inds, order)
        t1 = np.zeros((dims[m], R), np.float32)
        O.mttkrp_coo_f32([a[perm] for a in inds], vals[perm], m, U, t1)
        tol.append(min(1e-5, max(1e-6, rel(t1, ref[m])) + 1e-6))
    for f in (2, 128):
        for trees in (True, False):
            fm = mmcsf.Format(t, 12, fbr_threshold=f, mode_trees=trees)
            assert fm.info()["modeTrees"] == (N - 1 if trees else 0)
            outs, _ = mmcsf.mttkrp_allmode(fm, Ud)
            for m in range(N):
                assert rel(outs[m].cpu().numpy(), ref[m]) < tol[m], f"MM-CSF mode {m} f={f} trees={trees}"
                one = mmcsf.mttkrp(fm, m, Ud).cpu().numpy()
                assert rel(one, ref[m]) < tol[m]
    for m in range(N):
        for ft in (3, 8, 10):
            f = mmcsf.Format(t, ft, mode=m, fbr_threshold=3)
            o = mmcsf.mttkrp(f, m, Ud).cpu().numpy()
            assert rel(o, ref[m]) < tol[m], f"type {ft} mode {m}"


@pytest.mark.parametrize("dims,nnz,skew", CASES[:2] + CASES[3:4])
@pytest.mark.parametrize("R", [16, 32, 64])
def test_bf16_factor_storage(cuda, dims, nnz, skew, R):
    import torch
    import mmcsf
    rng = np.random.default_rng(7 * nnz + R)
    inds, vals = random_tensor(rng, dims, nnz, skew)
    U, Ud = factors(rng, dims, R, torch, cuda)
    Ub = [u.to(torch.bfloat16) for u in Ud]
    Ur = bf16_round(U, torch)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    for trees in (True, False):
        fm = mmcsf.Format(t, 12, fbr_threshold=128, mode_trees=trees)
        outs, _ = mmcsf.mttkrp_allmode(fm, Ub)
        for m in range(len(dims)):
            got = outs[m].cpu().numpy()
            assert rel(got, O.mttkrp_coo_f64(inds, vals, m, U)) < 1e-3
            assert rel(got, O.mttkrp_coo_f64(inds, vals, m, Ur)) < 1e-5
    for ft in (3, 8, 10):
        f = mmcsf.Format(t, ft, mode=1, fbr_threshold=128)
        got = mmcsf.mttkrp(f, 1, Ub).cpu().numpy()
        assert rel(got, O.mttkrp_coo_f64(inds, vals, 1, Ur)) < 1e-5


@pytest.mark.parametrize("dims,nnz,skew", [CASES[1], CASES[4]])
@pytest.mark.parametrize("R", [12, 20, 128, 256])
def test_bf16_odd_and_wide_ranks(cuda, dims, nnz, skew, R):
    """bfloat16 rows of 8 bytes per lane (R % 8 == 4: R = 12, 20) and of two
    or four column blocks (R = 128, 256), on a skewed 3-D and a skewed 4-D
    tensor, with and without mode trees."""
    import torch
    import mmcsf
    rng = np.random.default_rng(11 * nnz + R)
    inds, vals = random_tensor(rng, dims, nnz, skew)
    U, Ud = factors(rng, dims, R, torch, cuda)
    Ub = [u.to(torch.bfloat16) for u in Ud]
    Ur = bf16_round(U, torch)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    for trees in (True, False):
        fm = mmcsf.Format(t, 12, fbr_threshold=128, mode_trees=trees)
        outs, _ = mmcsf.mttkrp_allmode(fm, Ub)
        for m in range(len(dims)):
            got = outs[m].cpu().numpy()
            assert rel(got, O.mttkrp_f64_mt(inds, vals, m, Ur)) < 1e-5, f"R {R} trees {trees} mode {m}"
            assert rel(got, O.mttkrp_f64_mt(inds, vals, m, U)) < 1e-3


def test_uniform_2000_cubed_vs_t1_reference(cuda):
    """The plain mode-0 workload: uniform 2000^3, 1M distinct nonzeros, R=32,
    -f 128; GPU COO and B-CSF (and MM-CSF) against the -t 1 sequential float32
    restatement within 1e-5 relative, and against float64."""
    import torch
    import mmcsf
    dims = [2000, 2000, 2000]
    t = mmcsf.Tensor.generate(dims, 1_000_000, [0, 0, 0], [0, 0, 0], seed=0x4D4D43534600)
    _, inds, vals = t.to_numpy()
    g = torch.Generator(device="cpu").manual_seed(1)
    U = [torch.rand((d, 32), generator=g).numpy() for d in dims]
    Ud = [torch.from_numpy(u).to(cuda) for u in U]
    order = O.order_tensormode(dims, 0)
    perm = O.sort_perm(inds, order)  # sort_COOtensor before -t 1 (mttkrp.cu:79)
    sinds = [a[perm] for a in inds]
    t1 = np.zeros((dims[0], 32), np.float32)
    O.mttkrp_coo_f32(sinds, vals[perm], 0, U, t1)
    r64 = O.mttkrp_coo_f64(inds, vals, 0, U)
    assert rel(t1, r64) < 1e-6  # -t 1 is itself accurate on ~500-term rows
    for ft in (3, 8, 10, 12):
        f = mmcsf.Format(t, ft, mode=0, fbr_threshold=128)
        got = mmcsf.mttkrp(f, 0, Ud).cpu().numpy()
        assert rel(got, t1) < 1e-5, f"type {ft}"
        assert rel(got, r64) < 1e-6, f"type {ft}"


def test_empty_tensor_gives_zero_output(cuda):
    import torch
    import mmcsf
    dims = [4, 5, 6]
    t = mmcsf.Tensor.from_numpy(dims, [np.zeros(0, np.int64)] * 3, np.zeros(0, np.float32))
    Ud = [torch.rand((d, 8), device=cuda) for d in dims]
    for ft in (3, 8, 12):
        f = mmcsf.Format(t, ft, mode=0)
        out = torch.full((4, 8), 7.0, device=cuda)
        mmcsf.mttkrp(f, 0, Ud, out=out)
        assert float(out.abs().max()) == 0.0


def test_repeat_runs_agree(cuda):
    import torch
    import mmcsf
    rng = np.random.default_rng(21)
    dims = [500, 200, 50]
    inds, vals = random_tensor(rng, dims, 30000, [1.2, 1.2, None])
    U, Ud = factors(rng, dims, 32, torch, cuda)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    for trees in (True, False):
        fm = mmcsf.Format(t, 12, fbr_threshold=128, mode_trees=trees)
        a, _ = mmcsf.mttkrp_allmode(fm, Ud)
        b, _ = mmcsf.mttkrp_allmode(fm, Ud)
        for m in range(3):
            assert rel(a[m].cpu().numpy(), b[m].cpu().numpy()) < 1e-6


def test_error_paths(cuda):
    import torch
    import mmcsf
    rng = np.random.default_rng(1)
    dims = [10, 9, 8]
    inds, vals = random_tensor(rng, dims, 300)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    Ud = [torch.rand((d, 8), device=cuda) for d in dims]
    fm = mmcsf.Format(t, 12)
    with pytest.raises(mmcsf.MMCSFError):
        mmcsf.mttkrp(fm, 3, Ud)                                  # mode out of range
    with pytest.raises(mmcsf.MMCSFError):
        mmcsf.mttkrp(fm, 0, [torch.rand((d, 257), device=cuda) for d in dims])   # R > 256
    with pytest.raises(mmcsf.MMCSFError):
        mmcsf.mttkrp(fm, 0, [torch.rand((d, 6), device=cuda).to(torch.bfloat16) for d in dims])
    fb = mmcsf.Format(t, 8, mode=1)
    with pytest.raises(mmcsf.MMCSFError):
        mmcsf.mttkrp(fb, 0, Ud)                                  # B-CSF serves its root mode only
    with pytest.raises(mmcsf.MMCSFError):
        mmcsf.Tensor.from_numpy([10, 9], inds[:2], vals)          # 2 modes
    with pytest.raises(mmcsf.MMCSFError):
        mmcsf.Tensor.from_numpy(dims, [inds[0] + 10, inds[1], inds[2]], vals)  # out of range
    with pytest.raises(mmcsf.MMCSFError):
        mmcsf.Tensor.load_tns("/nonexistent/file.tns")
    with pytest.raises(mmcsf.MMCSFError):
        mmcsf.Format(t, 7)                                       # unsupported type
    # misaligned factor pointer
    big = torch.rand(10 * 8 + 1, device=cuda)
    mis = big[1:].view(10, 8)
    with pytest.raises(mmcsf.MMCSFError):
        mmcsf.mttkrp(fm, 1, [mis, None, Ud[2]])
    # HB-CSF with bfloat16 rows of R = 8 (16-byte row loads in the CSF part's
    # walk): factors 8 bytes off a 16-byte boundary are refused before any
    # part runs, and the output is left untouched
    fh = mmcsf.Format(t, 10, mode=0)
    assert fh.info()["HCSRnnz"] > 0
    Ub = [torch.rand((d, 8), device=cuda).to(torch.bfloat16) for d in dims]
    for m in (1, 2):
        raw = torch.empty(dims[m] * 8 + 4, dtype=torch.bfloat16, device=cuda)
        off = raw[4:].view(dims[m], 8)
        off.copy_(Ub[m])
        assert off.data_ptr() % 16 == 8
        bad = list(Ub)
        bad[m] = off
        out = torch.full((dims[0], 8), 7.0, device=cuda)
        with pytest.raises(mmcsf.MMCSFError, match="aligned to 16"):
            mmcsf.mttkrp(fh, 0, bad, out)
        torch.cuda.synchronize()
        assert bool((out == 7.0).all())


@pytest.mark.parametrize("dims,skew", [([60, 50, 40], None), ([300, 200, 100], [1.2, 1.2, 0]),
                                       ([30, 20, 25, 15], None), ([200, 40, 150, 8], [1.1, 0, 1.1, 0])])
@pytest.mark.parametrize("fmt_type", [9, 13, 14])
def test_one_tree_types_every_mode(cuda, dims, skew, fmt_type):
    """-t 9 / -t 13 / -t 14: the B-CSF arrays of -t 8 (type 13 without the slice
    bins), every mode computed from that one tree (root walk for modeOrder[0],
    scatter roles for the other levels).  For 4-D, type 13 computes every mode
    exactly where the reference's host roles cover the root mode only."""
    import torch
    import mmcsf
    rng = np.random.default_rng(fmt_type * 7 + len(dims))
    inds, vals = random_tensor(rng, dims, 20000, skew)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    for mode in (0, len(dims) - 1):
        f8 = mmcsf.Format(t, 8, mode=mode, fbr_threshold=16)
        fo = mmcsf.Format(t, fmt_type, mode=mode, fbr_threshold=16)
        assert fo.info()["modeOrder"] == f8.info()["modeOrder"]
        for l in range(len(dims) - 1):
            np.testing.assert_array_equal(fo.export("fbrPtr", l), f8.export("fbrPtr", l))
            np.testing.assert_array_equal(fo.export("fbrIdx", l), f8.export("fbrIdx", l))
        if fmt_type == 13:
            with pytest.raises(mmcsf.MMCSFError, match="no such bin"):
                fo.export("slcMapperBin", 0)
            np.testing.assert_array_equal(fo.export("fbrLikeSlcInds")[:-1], f8.export("fbrLikeSlcInds")[:-1])
        else:
            for b in range(10):
                np.testing.assert_array_equal(fo.export("slcMapperBin", b), f8.export("slcMapperBin", b))
        for R in (4, 32, 33):
            U, Ud = factors(rng, dims, R, torch, cuda)
            outs, _ = mmcsf.mttkrp_allmode(fo, Ud)
            for m in range(len(dims)):
                ref = O.mttkrp_coo_f64(inds, vals, m, U)
                assert rel(outs[m].cpu().numpy(), ref) < 1e-5, f"type {fmt_type} mode {m} R {R}"
                one = mmcsf.mttkrp(fo, m, Ud).cpu().numpy()
                assert rel(one, ref) < 1e-5
        # bfloat16 storage
        U, Ud = factors(rng, dims, 32, torch, cuda)
        Ub = [mmcsf.to_bf16(u) for u in Ud]
        Ur = bf16_round(U, torch)
        outs, _ = mmcsf.mttkrp_allmode(fo, Ub)
        for m in range(len(dims)):
            assert rel(outs[m].cpu().numpy(), O.mttkrp_coo_f64(inds, vals, m, Ur)) < 1e-5


@pytest.mark.parametrize("fmt_type", [9, 13, 14])
def test_one_tree_cli(cuda, golden_dir, fmt_type):
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cli = os.path.join(root, "mm-csf_amd", "bin", "mttkrp")
    r = subprocess.run([cli, "-i", f"{golden_dir}/toy.tns", "-t", str(fmt_type), "-R", "4", "-c", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    tag = {9: "ONE-B-CSF-GPU-mode", 13: "sameCSF-GPU-mode", 14: "singleCSF-GPU-mode"}[fmt_type]
    assert r.stdout.count(tag) == 3
    if fmt_type == 13:
        # the role lines of MTTKRP_HCSR_CPU / _mode1 / _mode2 for the toy's order 0-1-2
        for line in ("mode 0: 0 1 2", "mode 1: 1 2 0", "mode 2: 2 0 1"):
            assert line in r.stdout
    assert "Correctness pass!" in r.stdout


@pytest.mark.parametrize("dims,skew", [([60, 50, 40], None), ([300, 200, 100], [1.2, 1.2, 0]),
                                       ([30, 20, 25, 15], None)])
@pytest.mark.parametrize("fmt_type", [8, 13, 14])
def test_ttm_matches_float64_and_host(cuda, dims, skew, fmt_type):
    """TTM per leaf fiber (TTM_GPU) against the float64 restatement of TTM_CPU
    (oracle/dense.py) on the exported tree, and against the host TTM."""
    import torch
    import mmcsf
    from oracle import dense as D
    rng = np.random.default_rng(len(dims) + fmt_type)
    inds, vals = random_tensor(rng, dims, 20000, skew)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    for mode, fthr in ((0, 2), (len(dims) - 1, 128)):
        fm = mmcsf.Format(t, fmt_type, mode=mode, fbr_threshold=fthr)
        info = fm.info()
        N = len(dims)
        leaf_mode = info["modeOrder"][-1]
        ptr = fm.export("fbrPtr", N - 2)
        leaf = fm.export("inds", leaf_mode)
        v = fm.export("vals")
        for R in (4, 32, 64):
            U = rng.standard_normal((dims[leaf_mode], R)).astype(np.float32)
            ref = D.ttm64(ptr, leaf, v, U)
            Y = mmcsf.ttm(fm, torch.from_numpy(U).to(cuda)).cpu().numpy()
            assert Y.shape == ref.shape
            assert rel(Y, ref) < 1e-5, f"mode {mode} R {R}"
            Yh = mmcsf.ttm_host(fm, U)
            assert rel(Yh, ref) < 1e-5
        Ub = mmcsf.to_bf16(torch.from_numpy(U).to(cuda))
        Ur = Ub.float().cpu().double().numpy()
        assert rel(mmcsf.ttm(fm, Ub).cpu().numpy(), D.ttm64(ptr, leaf, v, Ur)) < 1e-5


def test_ttm_cli(cuda, golden_dir, tmp_path):
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cli = os.path.join(root, "mm-csf_amd", "bin", "ttm")
    out = tmp_path / "y.txt"
    r = subprocess.run([cli, "-i", f"{golden_dir}/toy.tns", "-t", "3", "-R", "4", "-c", "1", "-o", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "TTM GPU - time" in r.stdout and "Correctness pass!" in r.stdout
    lines = out.read_text().splitlines()
    assert lines[0] == "7 x 4 semiSpTensor"  # toy.tns -m 0: seven leaf fibers
    r = subprocess.run([cli, "-i", f"{golden_dir}/toy.tns", "-t", "1", "-R", "4"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "TTM - COO CPU time" in r.stdout


@pytest.mark.parametrize("fmt_type,tag", [(1, "COO CPU - time"), (3, "Loop on. COO GPU using loop - time"),
                                          (8, "B-CSF-GPU-mode"), (10, "HYB GPU:"), (12, "Total GPU time:"),
                                          (13, "sameCSF-GPU-mode 2")])
def test_mttkrp_cli_every_type(cuda, golden_dir, tmp_path, fmt_type, tag):
    """The driver's stdout lines (mttkrp.cu), -c 1 against the COO CPU type and
    the -o matrix file (write_output, util.h:3270-3283)."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cli = os.path.join(root, "mm-csf_amd", "bin", "mttkrp")
    out = tmp_path / "m.txt"
    args = [cli, "-i", f"{golden_dir}/toy.tns", "-t", str(fmt_type), "-R", "4", "-f", "2", "-o", str(out)]
    if fmt_type != 1:
        args += ["-c", "1"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert tag in r.stdout
    if fmt_type != 1:
        assert "Correctness pass!" in r.stdout
    head = out.read_text().splitlines()[0]
    assert head.endswith(" x 4 matrix")


def test_large_skewed_tensor_every_mode(cuda):
    """2M nonzeros, Zipf(1.2) on two modes: many waves, hot slices spanning
    waves, short-slice chunks (group-run path), cost-balanced partition."""
    import torch
    import mmcsf
    rng = np.random.default_rng(123)
    dims = [200000, 80000, 5000]
    inds, vals = random_tensor(rng, dims, 2_000_000, [1.2, 1.2, None])
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    U, Ud = factors(rng, dims, 32, torch, cuda)
    for trees in (True, False):
        fm = mmcsf.Format(t, 12, fbr_threshold=128, mode_trees=trees)
        outs, _ = mmcsf.mttkrp_allmode(fm, Ud)
        for m in range(3):
            ref = O.mttkrp_coo_f64(inds, vals, m, U)
            assert rel(outs[m].cpu().numpy(), ref) < 1e-5, f"mode {m} trees {trees}"


def test_factor_table_above_4gib(cuda):
    """A 34M-row mode makes a float32 R=32 factor table of 4.35 GB, past the
    root walk's 32-bit row offsets: MTTKRP of the other modes (whose trees
    gather that table) takes the single-tree kernels and still matches
    float64; CP completion refuses with MMCSF_ERR_UNSUPPORTED."""
    import torch
    import mmcsf
    rng = np.random.default_rng(41)
    dims = [34_000_000, 64, 48]
    nnz = 200_000
    inds, vals = random_tensor(rng, dims, nnz)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    R = 32
    g = torch.Generator(device=cuda).manual_seed(5)
    U = [torch.rand((d, R), generator=g, device=cuda) for d in dims]
    ind_d = [torch.from_numpy(a).to(cuda) for a in inds]
    val_d = torch.from_numpy(vals).to(cuda).double()

    def ref(mode):
        c = val_d[:, None].expand(-1, R).clone()
        for m in range(3):
            if m != mode:
                c *= U[m][ind_d[m]].double()
        out = torch.zeros((dims[mode], R), dtype=torch.float64, device=cuda)
        return out.index_add_(0, ind_d[mode], c)

    for mode in (1, 2):
        f = mmcsf.Format(t, 8, mode=mode, fbr_threshold=128)
        got = mmcsf.mttkrp(f, mode, U).double()
        r = ref(mode)
        assert float(torch.linalg.norm(got - r) / torch.linalg.norm(r)) < 1e-5, f"mode {mode}"
    fm = mmcsf.Format(t, 12, fbr_threshold=128)
    with pytest.raises(mmcsf.MMCSFError) as e:
        mmcsf.cp_forward(fm, U)
    assert e.value.code == -4


@pytest.mark.parametrize("nnz", [1, 2, 63, 64, 65, 127, 128, 129, 191, 192, 193, 640, 1000, 4099])
def test_chunk_count_edges(cuda, nnz):
    """The walk reads the words of chunk c+2 and the level-1 entries of chunk
    c+1 while chunk c computes, with its chunk loop unrolled by three: tensors
    whose nonzero counts sit at chunk multiples, and runs of one to a few
    chunks per wave, against float64 (all modes, float32 and bfloat16, and the
    CP loss through the same traversal)."""
    import torch
    import mmcsf
    from oracle import dense as D
    rng = np.random.default_rng(nnz)
    dims = [37, 23, 11]
    inds, vals = random_tensor(rng, dims, nnz, [1.3, None, None])
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    for R in (4, 32):
        U, Ud = factors(rng, dims, R, torch, cuda)
        fm = mmcsf.Format(t, 12, fbr_threshold=128)
        outs, _ = mmcsf.mttkrp_allmode(fm, Ud)
        outb, _ = mmcsf.mttkrp_allmode(fm, [u.to(torch.bfloat16) for u in Ud])
        Ub = bf16_round(U, torch)
        for m in range(3):
            assert rel(outs[m].cpu().numpy(), O.mttkrp_coo_f64(inds, vals, m, U)) < 1e-5, f"R {R} mode {m}"
            assert rel(outb[m].cpu().numpy(), O.mttkrp_coo_f64(inds, vals, m, Ub)) < 1e-5, f"bf16 R {R} mode {m}"
        U64 = [u.astype(np.float64) for u in U]
        L64, _ = D.cp_loss(inds, vals, U64)
        loss = mmcsf.cp_forward(fm, Ud)[0].item()
        assert abs(loss - L64) <= 1e-5 * L64 + 1e-12, f"loss R {R}"
