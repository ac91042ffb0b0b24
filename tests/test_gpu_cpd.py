"""GPU CP-ALS against the float64 CP-ALS restatement (oracle/dense.py) on the
same planted tensor and the same initial factors.

Tolerances: MFMA Gram vs float64 numpy <= 1e-5 relative (float32 inputs,
float32 MFMA accumulation over 1024-row chunks, float64 chunk sums); the fit
of every iteration within 1e-4 of the float64 ALS (the north star's bound);
bfloat16 MTTKRP copies within 1e-3 of the float32 run."""
import numpy as np
import pytest

from oracle import dense as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows", [1, 7, 1000, 70001])
@pytest.mark.parametrize("R", [1, 5, 16, 32, 33, 64, 100])
def test_mfma_gram_matches_numpy(cuda, rows, R):
    import torch
    import mmcsf
    g = torch.Generator(device="cpu").manual_seed(rows * 1000 + R)
    U = torch.randn((rows, R), generator=g)
    G = mmcsf.gram(U.to(cuda)).cpu().numpy()
    ref = U.double().numpy().T @ U.double().numpy()
    assert np.linalg.norm(G - ref) / np.linalg.norm(ref) < 1e-5
    np.testing.assert_allclose(G, G.T, rtol=0, atol=1e-9 * np.abs(ref).max())


def planted_tensor(rng, dims, nnz, rank, zipf0=1.2, noise=0.01):
    """Coordinates Zipf on mode 0, uniform elsewhere; values from a planted
    rank-`rank` CP model with |N(0,1)| factors plus N(0, noise) noise."""
    r0 = rng.zipf(zipf0, size=nnz * 4)
    r0 = r0[r0 <= dims[0]][: nnz * 2] - 1
    c = [r0] + [rng.integers(0, d, len(r0)) for d in dims[1:]]
    key = (c[0] * dims[1] + c[1]) * dims[2] + c[2]
    _, first = np.unique(key, return_index=True)
    first = np.sort(first)[:nnz]
    inds = [a[first] for a in c]
    A = [np.abs(rng.standard_normal((d, rank))) for d in dims]
    vals = D.cp_model_values(inds, A) + noise * rng.standard_normal(len(first))
    return inds, vals.astype(np.float32)


# R <= 36 factors the normal matrix by Cholesky in LDS, R = 48 in the global
# work space; R = 5 has a partial column quad in the row solve.  R <= 32 with
# R % 4 == 0 takes the fused row solve and normalisation + Gram: one 16-column
# tile (R = 8, 16), a partial second tile (R = 24), two full tiles (R = 32)
@pytest.mark.parametrize("R", [5, 8, 16, 24, 32, 48])
def test_cpd_als_matches_float64_als(cuda, R):
    import torch
    import mmcsf
    rng = np.random.default_rng(R)
    dims = [2000, 1000, 200]
    inds, vals = planted_tensor(rng, dims, 60000, R)
    U0 = [(rng.random((d, R)) * 6 - 3).astype(np.float32) for d in dims]
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    fmt = mmcsf.Format(t, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    Ud = [torch.from_numpy(u.copy()).to(cuda) for u in U0]
    iters = 12
    lam, st = mmcsf.cpd_als(fmt, Ud, iters=iters, tol=0.0)
    fits, U64, lam64 = D.cpd_als(inds, vals, dims, [u.astype(np.float64) for u in U0], iters, tol=0.0)
    assert st["iterations"] == iters
    np.testing.assert_allclose(st["fit_history"], fits, atol=1e-4, rtol=0)
    # the fit counts unobserved entries as zeros, so a sparse sample stays far
    # from 1; it must still improve over the iterations
    assert st["fit_history"][-1] > st["fit_history"][0]
    # final model agrees with the float64 model on the observed entries
    mg = D.cp_model_values(inds, [u.cpu().double().numpy() * (lam if m == 0 else 1.0) for m, u in enumerate(Ud)])
    m64 = D.cp_model_values(inds, [u * (lam64 if m == 0 else 1.0) for m, u in enumerate(U64)])
    assert np.linalg.norm(mg - m64) / np.linalg.norm(m64) < 1e-3


def test_cpd_als_singular_normal_equations(cuda):
    """A zero column in a factor makes every normal matrix that multiplies its
    Gram singular: the Cholesky factorization stops at that column's zero
    pivot, LU with partial pivoting finds no pivot either, and the call fails
    with an error instead of returning a model."""
    import torch
    import mmcsf
    rng = np.random.default_rng(3)
    dims = [300, 200, 100]
    R = 8
    inds, vals = planted_tensor(rng, dims, 20000, R)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    fmt = mmcsf.Format(t, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    U0 = [rng.random((d, R)).astype(np.float32) + 0.1 for d in dims]
    U0[1][:, 3] = 0.0
    Ud = [torch.from_numpy(u.copy()).to(cuda) for u in U0]
    with pytest.raises(mmcsf.MMCSFError, match="singular"):
        mmcsf.cpd_als(fmt, Ud, iters=2, tol=0.0)


@pytest.mark.parametrize("R", [16, 24])
def test_cpd_als_bf16_mttkrp(cuda, R):
    import torch
    import mmcsf
    rng = np.random.default_rng(3)
    dims = [2000, 1000, 200]
    inds, vals = planted_tensor(rng, dims, 60000, R)
    U0 = [(rng.random((d, R)) * 6 - 3).astype(np.float32) for d in dims]
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    fmt = mmcsf.Format(t, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    Ua = [torch.from_numpy(u.copy()).to(cuda) for u in U0]
    Ub = [torch.from_numpy(u.copy()).to(cuda) for u in U0]
    _, sa = mmcsf.cpd_als(fmt, Ua, iters=10, tol=0.0)
    _, sb = mmcsf.cpd_als(fmt, Ub, iters=10, tol=0.0, mttkrp_bf16=True)
    assert abs(sa["fit"] - sb["fit"]) < 1e-3


def test_cpd_als_tolerance_stops_early(cuda):
    import torch
    import mmcsf
    rng = np.random.default_rng(5)
    dims = [300, 200, 100]
    inds, vals = planted_tensor(rng, dims, 20000, 4, noise=0.0)
    U0 = [(rng.random((d, 4)) * 6 - 3).astype(np.float32) for d in dims]
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    fmt = mmcsf.Format(t, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    Ud = [torch.from_numpy(u.copy()).to(cuda) for u in U0]
    _, st = mmcsf.cpd_als(fmt, Ud, iters=60, tol=1e-5)
    fits, _, _ = D.cpd_als(inds, vals, dims, [u.astype(np.float64) for u in U0], 60, tol=1e-5)
    assert st["iterations"] < 60
    assert abs(st["iterations"] - len(fits)) <= 2
    assert abs(st["fit"] - fits[-1]) < 1e-4


def test_cpd_cli_runs(cuda, golden_dir, tmp_path):
    import os
    import subprocess
    root = os.path.dirname(golden_dir.rstrip("/").rsplit("/", 1)[0])
    cli = os.path.join(root, "mm-csf_amd", "bin", "cpd")
    r = subprocess.run([cli, "-i", f"{golden_dir}/toy.tns", "-R", "2", "-n", "5"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "fit =" in r.stdout and "Total MTTKRP time" in r.stdout
