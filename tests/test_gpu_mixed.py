"""MM-CSF with the mixed-mode partition (-M 1: the heuristic of
mm_partition_reuseBased, util.h:2386-2512, instead of the override to
partition 2): partition of every nonzero, the walk order and every partition's
arrays against the oracle (bit-exact), all-mode MTTKRP over the partitions
against float64, CP-ALS on it, and its CLI.

Tolerances: float32 factors within 1e-5 relative (Frobenius) of float64;
bfloat16 factors within 1e-5 of float64 of the bf16-rounded factors; CP-ALS
fits within 1e-4 of the float64 ALS."""
import numpy as np
import pytest

import oracle as O
from oracle import dense as D
from test_gpu_builder import random_tensor

pytestmark = pytest.mark.gpu


def rel(a, b):
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def check_arrays(fm, ref, N):
    np.testing.assert_array_equal(fm.export("partOrder"), ref["order"])
    np.testing.assert_array_equal(fm.export("partPerNnz"), ref["part"])
    info = fm.info()
    assert info["mixedPart"] == 1
    for p, pp in enumerate(ref["partitions"]):
        if pp is None:
            assert info["partNnz"][p] == 0
            continue
        assert info["partNnz"][p] == pp["nnz"]
        np.testing.assert_array_equal(fm.export("P.modeOrder", 8 * p), pp["modeOrder"])
        for l in range(N - 1):
            np.testing.assert_array_equal(fm.export("P.fbrPtr", 8 * p + l), pp["fbrPtr"][l], f"part {p} fbrPtr {l}")
            got = fm.export("P.fbrIdx", 8 * p + l)
            want = pp["fbrIdx"][l]
            if l == N - 2:  # the trailing entry is unspecified in the reference
                got, want = got[:-1], want[:-1]
            np.testing.assert_array_equal(got, want, f"part {p} fbrIdx {l}")
        for m in range(N):
            np.testing.assert_array_equal(fm.export("P.inds", 8 * p + m), pp["keys"][pp["modeOrder"].index(m)])
        np.testing.assert_array_equal(fm.export("P.vals", 8 * p), pp["vals"])
        np.testing.assert_array_equal(fm.export("P.fbrLikeSlcInds", 8 * p)[:-1], pp["fbrLikeSlcInds"][:-1])


def test_toy_partition_and_mttkrp(cuda, golden_dir):
    import torch
    import mmcsf
    dims, inds, vals = O.load_tns(f"{golden_dir}/toy.tns")
    t = mmcsf.Tensor.load_tns(f"{golden_dir}/toy.tns")
    fm = mmcsf.Format(t, 12, fbr_threshold=128, mixed=True)
    ref = O.mm_partition(dims, inds, vals, fb_th=1, fbr_threshold=128)
    np.testing.assert_array_equal(fm.export("partPerNnz"), [2, 2, 0, 0, 0, 0, 0, 0])
    check_arrays(fm, ref, 3)
    # the reference's constant factors: slice sums (README toy output)
    expect = {0: [48.75, 67.5, 52.5], 1: [13.75, 16.25, 26.25], 2: [12.75, 13.5, 7.5]}
    Ud = [torch.full((3, 4), m + 0.5, device=cuda) for m in range(3)]
    outs, _ = mmcsf.mttkrp_allmode(fm, Ud)
    for m in range(3):
        np.testing.assert_allclose(outs[m].cpu().numpy(), np.repeat(np.array(expect[m])[:, None], 4, 1), rtol=1e-6)


CASES = [([600, 400, 100], 30000, [1.2, 1.2, None]), ([300, 200, 250], 20000, [None, 1.3, None]),
         ([200, 60, 150, 12], 25000, [1.1, None, 1.1, None]), ([40, 30, 25, 20], 12000, None)]


@pytest.mark.parametrize("dims,nnz,skew", CASES)
@pytest.mark.parametrize("fbth,fthr", [(1, 128), (2, 4)])
def test_partitions_bit_exact_and_mttkrp(cuda, dims, nnz, skew, fbth, fthr):
    import torch
    import mmcsf
    rng = np.random.default_rng(nnz + fbth)
    inds, vals = random_tensor(rng, dims, nnz, skew)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    fm = mmcsf.Format(t, 12, fbr_threshold=fthr, mixed=True, fb_th=fbth)
    ref = O.mm_partition(dims, inds, vals, fb_th=fbth, fbr_threshold=fthr)
    N = len(dims)
    check_arrays(fm, ref, N)
    assert sum(p is not None for p in ref["partitions"]) >= 2, "the case should use several partitions"
    for R in (4, 32, 33, 64):
        U = [rng.random((d, R)).astype(np.float32) for d in dims]
        Ud = [torch.from_numpy(u).to(cuda) for u in U]
        outs, ms = mmcsf.mttkrp_allmode(fm, Ud, timed=True)
        for m in range(N):
            r64 = O.mttkrp_f64_mt(inds, vals, m, U)
            assert rel(outs[m].cpu().numpy(), r64) < 1e-5, f"R {R} mode {m}"
            assert rel(mmcsf.mttkrp(fm, m, Ud).cpu().numpy(), r64) < 1e-5
        if R in (32, 64):
            Ub = [u.to(torch.bfloat16) for u in Ud]
            Ur = [u.float().cpu().numpy() for u in Ub]
            ob, _ = mmcsf.mttkrp_allmode(fm, Ub)
            for m in range(N):
                assert rel(ob[m].cpu().numpy(), O.mttkrp_f64_mt(inds, vals, m, Ur)) < 1e-5, f"bf16 R {R} mode {m}"


def test_cpd_als_on_mixed_partitions(cuda):
    import torch
    import mmcsf
    from test_gpu_cpd import planted_tensor
    rng = np.random.default_rng(8)
    dims = [2000, 1000, 200]
    R = 8
    inds, vals = planted_tensor(rng, dims, 60000, R)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    fm = mmcsf.Format(t, 12, fbr_threshold=128, mixed=True)
    assert sum(1 for n in fm.info()["partNnz"] if n) >= 2
    U0 = [(rng.random((d, R)) * 6 - 3).astype(np.float32) for d in dims]
    Ud = [torch.from_numpy(u.copy()).to(cuda) for u in U0]
    _, st = mmcsf.cpd_als(fm, Ud, iters=10, tol=0.0)
    fits, _, _ = D.cpd_als(inds, vals, dims, [u.astype(np.float64) for u in U0], 10, tol=0.0,
                           mttkrp=lambda i, v, W, m: O.mttkrp_f64_mt(i, v, m, W))
    np.testing.assert_allclose(st["fit_history"], fits, atol=1e-4, rtol=0)


def test_cp_completion_refuses_mixed_partitions(cuda):
    import torch
    import mmcsf
    rng = np.random.default_rng(2)
    dims = [50, 40, 30]
    inds, vals = random_tensor(rng, dims, 3000)
    fm = mmcsf.Format(mmcsf.Tensor.from_numpy(dims, inds, vals), 12, mixed=True)
    with pytest.raises(mmcsf.MMCSFError) as e:
        mmcsf.cp_forward(fm, [torch.rand((d, 8), device=cuda) for d in dims])
    assert e.value.code == -4


def test_cli_mixed_partition(cuda, golden_dir):
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cli = os.path.join(root, "mm-csf_amd", "bin", "mttkrp")
    r = subprocess.run([cli, "-i", f"{golden_dir}/toy.tns", "-t", "12", "-M", "1", "-R", "4", "-c", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Total GPU time:" in r.stdout and "Correctness pass!" in r.stdout
