"""GPU parity at the full benchmark shapes, against the C float64 oracle split
over host threads (``oracle.mttkrp_f64_mt`` / ``oracle.cp_resid_f64_mt``,
equal bit for bit to the sequential float64 oracle).

The tensors are the ones bench.py times, generated on the GPU from the same
seeds:
* 500000x200000x50000, 100M distinct nonzeros, Zipf1.2 x Zipf1.2 x uniform,
  R=32, -f 128: every mode of MM-CSF with a root tree per mode (-a 1) and
  with the single reference tree (-a 0), and -t 8 / -t 10 rooted at each
  mode, float32 and bfloat16 factors;
* 100000x50000x20000x1000, 50M nonzeros, Zipf1.1 x U x Zipf1.1 x U, all
  modes at R=16 and R=64, float32 and bfloat16;
* CP-ALS at R=32, 20 iterations, 200000x100000x20000 with 20M nonzeros of a
  planted rank-32 |N(0,1)| model + N(0, 0.01^2) noise;
* CP completion on 1000000x500000x100000 with 50M entries of a planted
  rank-16 N(0,1)/4 model + N(0, 0.01^2) noise: loss and every gradient at
  R=16 and R=64, float32 and bfloat16; finite differences on a 10k-nonzero
  subsample; Adam at R=16.

Tolerances (written per test): float32 factors within 1e-5 relative
(Frobenius) of float64; bfloat16 factors within 1e-3 of float64 of the
float32 factors and within 1e-5 of float64 of the bf16-rounded factors; CP-ALS
fit of every iteration within 1e-4 of the float64 ALS.
"""
import numpy as np
import pytest

import oracle as O
from oracle import dense as D

pytestmark = pytest.mark.gpu

C2_DIMS = [500000, 200000, 50000]
C2_NNZ = 100_000_000
C2_SEED = 0x4D4D43534601
C3_DIMS = [100000, 50000, 20000, 1000]
C3_NNZ = 50_000_000
C3_SEED = 0x4D4D43534602
C4_DIMS = [200000, 100000, 20000]
C4_NNZ = 20_000_000
C4_SEED = 0x4D4D43534603
C5_DIMS = [1000000, 500000, 100000]
C5_NNZ = 50_000_000
C5_SEED = 0x4D4D43534604


def rel(a, b):
    b = np.asarray(b, dtype=np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / (n if n > 0 else 1.0))


def log(msg):
    print(msg, flush=True)


def host_coo(X, cuda):
    """Coordinates (uint32 per mode) and values of a device tensor, on the host."""
    inds, vals = X.to_torch(cuda)
    return [a.cpu().numpy().view(np.uint32) for a in inds], vals.cpu().numpy()


def planted_values(torch, inds, A, noise, seed, chunk=4_000_000):
    """sum_r prod_m A_m[i_m][r] + noise * N(0,1) in float64 on the GPU, rounded
    to float32 (the same values bench.py generates)."""
    n = inds[0].numel()
    g = torch.Generator(device=inds[0].device).manual_seed(seed)
    out = torch.empty(n, dtype=torch.float32, device=inds[0].device)
    for s in range(0, n, chunk):
        e = min(s + chunk, n)
        p = torch.ones((e - s, A[0].shape[1]), dtype=torch.float64, device=inds[0].device)
        for m, a in enumerate(A):
            p *= a[inds[m][s:e].long()]
        z = torch.randn(e - s, generator=g, device=inds[0].device, dtype=torch.float64)
        out[s:e] = (p.sum(1) + noise * z).float()
    return out


# ---------------------------------------------------------------------------
# 100M-nonzero all-mode MTTKRP
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2(cuda):
    import torch
    import mmcsf
    R = 32
    X = mmcsf.Tensor.generate(C2_DIMS, C2_NNZ, [1, 1, 0], [1.2, 1.2, 0.0], C2_SEED)
    inds, vals = host_coo(X, cuda)
    g = torch.Generator(device=cuda).manual_seed(C2_SEED)
    U = [torch.rand((d, R), generator=g, device=cuda) for d in C2_DIMS]
    Ub = [u.to(torch.bfloat16) for u in U]
    Uh = [u.cpu().numpy() for u in U]
    Ur = [u.float().cpu().numpy() for u in Ub]
    ref32 = [O.mttkrp_f64_mt(inds, vals, m, Uh) for m in range(3)]
    ref16 = [O.mttkrp_f64_mt(inds, vals, m, Ur) for m in range(3)]
    log("c2 oracle done")
    return {"X": X, "U": U, "Ub": Ub, "ref32": ref32, "ref16": ref16}


def _check_c2(outs32, outs16, c2, what):
    for m in range(3):
        e32 = rel(outs32[m], c2["ref32"][m])
        e16 = rel(outs16[m], c2["ref32"][m])
        e16r = rel(outs16[m], c2["ref16"][m])
        log(f"{what} mode {m}: f32 {e32:.2e}, bf16 {e16:.2e} (vs rounded {e16r:.2e})")
        assert e32 < 1e-5, f"{what} mode {m} float32"
        assert e16 < 1e-3, f"{what} mode {m} bfloat16"
        assert e16r < 1e-5, f"{what} mode {m} bfloat16 vs rounded factors"


@pytest.mark.parametrize("trees", [True, False])
def test_c2_mmcsf_every_mode(cuda, c2, trees):
    import mmcsf
    fm = mmcsf.Format(c2["X"], mmcsf.TYPE_MMCSF, fbr_threshold=128, mode_trees=trees)
    o32, _ = mmcsf.mttkrp_allmode(fm, c2["U"])
    o32 = [o.cpu().numpy() for o in o32]
    o16, _ = mmcsf.mttkrp_allmode(fm, c2["Ub"])
    o16 = [o.cpu().numpy() for o in o16]
    _check_c2(o32, o16, c2, f"-t 12 -a {int(trees)}")


@pytest.mark.parametrize("fmt_type", [8, 10])
def test_c2_bcsf_hbcsf_each_mode(cuda, c2, fmt_type):
    import mmcsf
    o32, o16 = [], []
    for m in range(3):
        f = mmcsf.Format(c2["X"], fmt_type, mode=m, fbr_threshold=128)
        o32.append(mmcsf.mttkrp(f, m, c2["U"]).cpu().numpy())
        o16.append(mmcsf.mttkrp(f, m, c2["Ub"]).cpu().numpy())
        del f
    _check_c2(o32, o16, c2, f"-t {fmt_type}")


# ---------------------------------------------------------------------------
# 4-mode precision sweep, 50M nonzeros
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c3(cuda):
    import mmcsf
    X = mmcsf.Tensor.generate(C3_DIMS, C3_NNZ, [1, 0, 1, 0], [1.1, 0.0, 1.1, 0.0], C3_SEED)
    inds, vals = host_coo(X, cuda)
    return {"X": X, "inds": inds, "vals": vals}


@pytest.mark.parametrize("R", [16, 64])
def test_c3_four_mode_precision_sweep(cuda, c3, R):
    import torch
    import mmcsf
    g = torch.Generator(device=cuda).manual_seed(C3_SEED + R)
    U = [torch.rand((d, R), generator=g, device=cuda) for d in C3_DIMS]
    Ub = [u.to(torch.bfloat16) for u in U]
    Uh = [u.cpu().numpy() for u in U]
    Ur = [u.float().cpu().numpy() for u in Ub]
    for trees in (True, False):
        fm = mmcsf.Format(c3["X"], mmcsf.TYPE_MMCSF, fbr_threshold=128, mode_trees=trees)
        o32 = [o.cpu().numpy() for o in mmcsf.mttkrp_allmode(fm, U)[0]]
        o16 = [o.cpu().numpy() for o in mmcsf.mttkrp_allmode(fm, Ub)[0]]
        del fm
        for m in range(4):
            r32 = O.mttkrp_f64_mt(c3["inds"], c3["vals"], m, Uh)
            r16 = O.mttkrp_f64_mt(c3["inds"], c3["vals"], m, Ur)
            e32, e16, e16r = rel(o32[m], r32), rel(o16[m], r32), rel(o16[m], r16)
            log(f"4-D R={R} -a {int(trees)} mode {m}: f32 {e32:.2e}, bf16 {e16:.2e} (vs rounded {e16r:.2e})")
            assert e32 < 1e-5 and e16 < 1e-3 and e16r < 1e-5, f"R {R} trees {trees} mode {m}"


# ---------------------------------------------------------------------------
# CP-ALS, R=32, 20 iterations, 20M nonzeros
# ---------------------------------------------------------------------------
def test_c4_cpd_als_every_iteration(cuda):
    import torch
    import mmcsf
    R, iters = 32, 20
    X0 = mmcsf.Tensor.generate(C4_DIMS, C4_NNZ, [1, 0, 0], [1.2, 0.0, 0.0], C4_SEED)
    inds_d, _ = X0.to_torch(cuda)
    del X0
    g = torch.Generator(device=cuda).manual_seed(C4_SEED)
    A = [torch.randn((d, R), generator=g, device=cuda, dtype=torch.float64).abs() for d in C4_DIMS]
    vals_d = planted_values(torch, inds_d, A, 0.01, C4_SEED + 1)
    X = mmcsf.Tensor.from_torch(C4_DIMS, [a.contiguous() for a in inds_d], vals_d)
    fm = mmcsf.Format(X, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    U0 = [torch.rand((d, R), generator=g, device=cuda) for d in C4_DIMS]
    U = [u.clone() for u in U0]
    _, st = mmcsf.cpd_als(fm, U, iters=iters, tol=0.0)
    Ub = [u.clone() for u in U0]
    _, stb = mmcsf.cpd_als(fm, Ub, iters=iters, tol=0.0, mttkrp_bf16=True)
    inds = [a.cpu().numpy().view(np.uint32) for a in inds_d]
    vals = vals_d.cpu().numpy()
    fits, _, _ = D.cpd_als(inds, vals, C4_DIMS, [u.cpu().double().numpy() for u in U0], iters, tol=0.0,
                           mttkrp=lambda i, v, W, m: O.mttkrp_f64_mt(i, v, m, W))
    assert st["iterations"] == iters and len(fits) == iters
    d32 = np.abs(np.array(st["fit_history"][:iters]) - fits)
    d16 = np.abs(np.array(stb["fit_history"][:iters]) - fits)
    log(f"CP-ALS fits: gpu {st['fit_history'][iters - 1]:.6e}, float64 {fits[-1]:.6e}; "
        f"max |diff| f32 {d32.max():.2e}, bf16 MTTKRP {d16.max():.2e}")
    assert d32.max() < 1e-4
    assert d16.max() < 1e-4
    assert st["fit_history"][iters - 1] > st["fit_history"][0]


# ---------------------------------------------------------------------------
# CP completion, 50M observed entries
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c5(cuda):
    import torch
    import mmcsf
    X0 = mmcsf.Tensor.generate(C5_DIMS, C5_NNZ, [1, 1, 0], [1.1, 1.1, 0.0], C5_SEED)
    inds_d, _ = X0.to_torch(cuda)
    del X0
    g = torch.Generator(device=cuda).manual_seed(C5_SEED)
    A = [torch.randn((d, 16), generator=g, device=cuda, dtype=torch.float64) / 4.0 for d in C5_DIMS]
    vals_d = planted_values(torch, inds_d, A, 0.01, C5_SEED + 1)
    del A
    X = mmcsf.Tensor.from_torch(C5_DIMS, [a.contiguous() for a in inds_d], vals_d)
    fm = mmcsf.Format(X, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    inds = [a.cpu().numpy().view(np.uint32) for a in inds_d]
    return {"X": X, "fmt": fm, "inds": inds, "vals": vals_d.cpu().numpy(), "inds_d": inds_d, "vals_d": vals_d}


@pytest.mark.parametrize("R", [16, 64])
def test_c5_loss_and_gradients(cuda, c5, R):
    import torch
    import mmcsf
    g = torch.Generator(device=cuda).manual_seed(C5_SEED + R)
    U = [0.3 * torch.randn((d, R), generator=g, device=cuda) for d in C5_DIMS]
    Ub = [u.to(torch.bfloat16) for u in U]
    for F, name in ((U, "f32"), (Ub, "bf16")):
        Fh = [f.float().cpu().numpy() for f in F]
        L64, e = O.cp_resid_f64_mt(c5["inds"], c5["vals"], Fh)
        loss, grads = mmcsf.cp_loss_grad(c5["fmt"], F)
        el = abs(loss.item() - L64) / L64
        log(f"CP completion R={R} {name}: loss {loss.item():.6e} vs {L64:.6e} ({el:.2e})")
        assert el < 1e-5
        for m in range(3):
            ref = O.mttkrp_f64_mt(c5["inds"], e, m, Fh)
            eg = rel(grads[m].cpu().numpy(), ref)
            log(f"  gradient mode {m}: {eg:.2e}")
            assert eg < 1e-5, f"R {R} {name} mode {m}"
        if name == "f32":
            _, res = mmcsf.cp_forward(c5["fmt"], F, residual=True)
            assert rel(res.cpu().numpy(), e) < 1e-5


def test_c5_finite_differences_on_subsample(cuda, c5):
    """<dL/dU, D> on a 10k-nonzero subsample of the 50M-entry tensor against
    the central difference of the float64 loss along a random direction."""
    import torch
    import mmcsf
    rng = np.random.default_rng(C5_SEED)
    sel = np.sort(rng.choice(C5_NNZ, 10000, replace=False))
    sub = [a[sel].astype(np.int64) for a in c5["inds"]]
    sv = c5["vals"][sel]
    Xs = mmcsf.Tensor.from_numpy(C5_DIMS, sub, sv)
    fs = mmcsf.Format(Xs, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    U = [(0.3 * rng.standard_normal((d, 16))).astype(np.float32) for d in C5_DIMS]
    _, grads = mmcsf.cp_loss_grad(fs, [torch.from_numpy(u).to(cuda) for u in U])
    Dir = [rng.standard_normal(u.shape) for u in U]
    U64 = [u.astype(np.float64) for u in U]
    h = 1e-4
    Lp = D.cp_loss(sub, sv, [u + h * d for u, d in zip(U64, Dir)])[0]
    Lm = D.cp_loss(sub, sv, [u - h * d for u, d in zip(U64, Dir)])[0]
    fd = (Lp - Lm) / (2 * h)
    an = sum(float((gr.cpu().double().numpy() * d).sum()) for gr, d in zip(grads, Dir))
    log(f"FD: analytic {an:.8e}, central difference {fd:.8e}")
    assert abs(an - fd) / abs(fd) < 1e-6


def test_c5_adam_r16_matches_float64(cuda, c5):
    """100 Adam steps (lr 1e-2) at R=16 on the first 1M entries of the 50M
    tensor: the loss curve of the GPU run follows a float64 run whose
    gradients come from the C oracle, with torch's Adam on both sides."""
    import torch
    import mmcsf
    n = 1_000_000
    inds = [a[:n] for a in c5["inds"]]
    vals = c5["vals"][:n]
    Xs = mmcsf.Tensor.from_numpy(C5_DIMS, [a.astype(np.int64) for a in inds], vals)
    fs = mmcsf.Format(Xs, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    rng = np.random.default_rng(C5_SEED + 16)
    U0 = [(0.3 * rng.standard_normal((d, 16))).astype(np.float32) for d in C5_DIMS]
    P = [torch.from_numpy(u).to(cuda).requires_grad_(True) for u in U0]
    opt = torch.optim.Adam(P, lr=1e-2)
    gpu = []
    for _ in range(100):
        opt.zero_grad()
        loss = mmcsf.cp_completion_loss(fs, P)
        loss.backward()
        opt.step()
        gpu.append(loss.item())
    Q = [torch.from_numpy(u.astype(np.float64)).requires_grad_(False) for u in U0]
    opt2 = torch.optim.Adam(Q, lr=1e-2)
    ref = []
    for _ in range(100):
        Wh = [q.numpy() for q in Q]
        L, e = O.cp_resid_f64_mt(inds, vals, Wh)
        for m, q in enumerate(Q):
            q.grad = torch.from_numpy(O.mttkrp_f64_mt(inds, e, m, Wh))
        opt2.step()
        ref.append(L)
    gpu, ref = np.array(gpu), np.array(ref)
    dev_ = np.max(np.abs(gpu - ref) / ref)
    log(f"Adam R=16: loss {ref[0]:.4e} -> {ref[-1]:.4e}; max relative deviation {dev_:.2e}")
    assert ref[-1] < 0.5 * ref[0]
    assert dev_ < 1e-3
