"""The allocation-free CP-ALS entry point (mmcsf_cpd_als_noalloc), called
through the C-ABI, against mmcsf_cpd_als on the same tensors and initial
factors: two calls in a row, the second on a smaller tensor with a smaller
rank (the shared scratch keeps the first call's contents), and the refusal of
a work space above the scratch size.

Tolerances: the fit of every iteration within 1e-7 (absolute) of the
allocating entry point's and the final factors and lambda within 1e-4
relative (Frobenius); both run the same kernels, whose float atomics at wave
boundaries and float64 Gram atomics may add in a different order."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def planted(rng, dims, nnz, R):
    c = [rng.integers(0, d, nnz * 2) for d in dims]
    key = (c[0] * dims[1] + c[1]) * dims[2] + c[2]
    _, first = np.unique(key, return_index=True)
    first = np.sort(first)[:nnz]
    inds = [a[first] for a in c]
    A = [np.abs(rng.standard_normal((d, R))) for d in dims]
    p = np.ones((len(first), R))
    for m in range(3):
        p *= A[m][inds[m]]
    return inds, (p.sum(1) + 0.01 * rng.standard_normal(len(first))).astype(np.float32)


def run_both(cuda, dims, nnz, R, seed, iters=8):
    import torch
    import mmcsf
    rng = np.random.default_rng(seed)
    inds, vals = planted(rng, dims, nnz, R)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    fmt = mmcsf.Format(t, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    U0 = [(rng.random((d, R)) * 2).astype(np.float32) for d in dims]
    Ua = [torch.from_numpy(u.copy()).to(cuda) for u in U0]
    Ub = [torch.from_numpy(u.copy()).to(cuda) for u in U0]
    lam_a, st_a = mmcsf.cpd_als(fmt, Ua, iters=iters, tol=0.0)
    # the allocation-free entry point through the C ABI
    L = mmcsf.lib()
    fp = (C.c_void_p * 3)(*[C.c_void_p(u.data_ptr()) for u in Ub])
    lam_b = (C.c_double * R)()
    st_b = mmcsf.CpdStats()
    rc = L.mmcsf_cpd_als_noalloc(fmt._h, R, iters, 0.0, mmcsf.F32, fp, lam_b, C.byref(st_b),
                                 C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, L.mmcsf_last_error()
    torch.cuda.synchronize()
    assert st_b.iterations == st_a["iterations"] == iters
    assert st_b.ms_mttkrp == 0.0
    fits_b = np.array(st_b.fit_history[:iters])
    np.testing.assert_allclose(fits_b, st_a["fit_history"], atol=1e-7, rtol=0)
    assert rel(np.array(lam_b[:]), lam_a) < 1e-4
    for a, b in zip(Ua, Ub):
        assert rel(b.cpu().numpy(), a.cpu().numpy()) < 1e-4
    return fits_b


def test_noalloc_cpd_matches_allocating_entry_point_twice(cuda):
    # first call: 40000 x 20000 x 4000, 300k nonzeros, R=16; second call on a
    # smaller tensor and rank, over the contents the first call left behind
    f1 = run_both(cuda, [40000, 20000, 4000], 300_000, 16, 1)
    f2 = run_both(cuda, [3000, 2000, 500], 40_000, 8, 2)
    assert f1[-1] > f1[0] and f2[-1] > f2[0]


def test_noalloc_refuses_work_space_above_scratch(cuda):
    import torch
    import mmcsf
    # a 3M-row mode at R=8: the MTTKRP output alone needs 96 MB > 64 MiB
    dims = [3_000_000, 20, 10]
    rng = np.random.default_rng(3)
    inds = [rng.integers(0, d, 5000) for d in dims]
    t = mmcsf.Tensor.from_numpy(dims, inds, rng.random(5000).astype(np.float32))
    fmt = mmcsf.Format(t, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    U = [torch.rand((d, 8), device=cuda) for d in dims]
    with pytest.raises(mmcsf.MMCSFError) as e:
        mmcsf.cpd_als(fmt, U, iters=2, no_alloc=True)
    assert e.value.code == -4 and "scratch" in str(e.value)
