"""GPU CP completion (loss, residuals, gradients, autograd) against the float64
restatement in oracle/dense.py (cp_loss / cp_grad / mttkrp64), whose gradient
is itself pinned by finite differences in test_oracle.py.

Tolerances: float32 factors -> loss, residual and every gradient within 1e-5
relative (Frobenius) of float64; bfloat16 factors -> within 1e-5 of float64
evaluated on the bf16-rounded factors; directional derivative of the float64
loss by central differences within 1e-6 relative of <grad, direction>;
Adam training trajectories within 1e-3 relative of a float64 torch-autograd
run of the same optimiser from the same start."""
import numpy as np
import pytest

from oracle import dense as D
from test_gpu_builder import random_tensor

pytestmark = pytest.mark.gpu


def rel(a, b):
    b = np.asarray(b, dtype=np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / (n if n > 0 else 1.0))


def setup(rng, dims, nnz, R, torch, dev, skew=None, fmt_type=12, trees=True):
    import mmcsf
    inds, vals = random_tensor(rng, dims, nnz, skew)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    fmt = mmcsf.Format(t, fmt_type, fbr_threshold=128, mode_trees=trees)
    U = [(rng.standard_normal((d, R)) / np.sqrt(R)).astype(np.float32) for d in dims]
    Ud = [torch.from_numpy(u).to(dev) for u in U]
    return inds, vals, fmt, U, Ud


CASES = [([300, 200, 100], 20000, None), ([400, 300, 50], 30000, [1.2, 1.2, 0]),
         ([40, 30, 20, 25], 15000, None), ([300, 60, 200, 10], 20000, [1.1, 0, 1.1, 0])]


@pytest.mark.parametrize("dims,nnz,skew", CASES)
@pytest.mark.parametrize("R", [4, 16, 32, 64])
def test_loss_residual_gradients_float32(cuda, dims, nnz, skew, R):
    import torch
    import mmcsf
    rng = np.random.default_rng(R + len(dims))
    inds, vals, fmt, U, Ud = setup(rng, dims, nnz, R, torch, cuda, skew)
    U64 = [u.astype(np.float64) for u in U]
    L64, e64 = D.cp_loss(inds, vals, U64)
    g64 = D.cp_grad(inds, vals, U64)  # d/dU of 1/2 |v - model|^2
    loss, res = mmcsf.cp_forward(fmt, Ud, residual=True)
    assert abs(loss.item() - L64) / L64 < 1e-5
    # residual e = model - value, in the tensor's order
    assert rel(res.cpu().numpy(), -e64) < 1e-5
    loss2, grads = mmcsf.cp_loss_grad(fmt, Ud)
    assert abs(loss2.item() - L64) / L64 < 1e-5
    for m in range(len(dims)):
        assert rel(grads[m].cpu().numpy(), g64[m]) < 1e-5, f"mode {m}"
        adj = mmcsf.cp_adjoint(fmt, m, Ud, alpha=-2.0).cpu().numpy()
        assert rel(adj, -2.0 * g64[m]) < 1e-5, f"adjoint mode {m}"
    # the vector-Jacobian product with caller weights is an MTTKRP of the weights
    w = rng.standard_normal(len(vals)).astype(np.float32)
    wd = torch.from_numpy(w).to(cuda)
    for m in range(len(dims)):
        got = mmcsf.cp_adjoint(fmt, m, Ud, weights=wd, alpha=0.5).cpu().numpy()
        assert rel(got, 0.5 * D.mttkrp64(inds, w.astype(np.float64), U64, m)) < 1e-5, f"weights mode {m}"


@pytest.mark.parametrize("R", [8, 16, 64])
def test_bf16_factors(cuda, R):
    import torch
    import mmcsf
    rng = np.random.default_rng(40 + R)
    dims = [500, 300, 200]
    inds, vals, fmt, U, Ud = setup(rng, dims, 30000, R, torch, cuda, [1.2, 1.2, 0])
    Ub = [u.to(torch.bfloat16).contiguous() for u in Ud]
    Ur = [u.to(torch.float32).cpu().double().numpy() for u in Ub]
    L64, _ = D.cp_loss(inds, vals, Ur)
    g64 = D.cp_grad(inds, vals, Ur)
    loss, grads = mmcsf.cp_loss_grad(fmt, Ub)
    assert abs(loss.item() - L64) / L64 < 1e-5
    for m in range(3):
        assert rel(grads[m].cpu().numpy(), g64[m]) < 1e-5


def test_directional_derivative_matches_float64_finite_difference(cuda):
    """On a 10k-nonzero sample: <dL/dU, D> from the GPU against the central
    difference of the float64 loss along a random direction D."""
    import torch
    import mmcsf
    rng = np.random.default_rng(77)
    dims = [1000, 500, 100]
    inds, vals, fmt, U, Ud = setup(rng, dims, 10000, 16, torch, cuda, [1.1, 1.1, 0])
    _, grads = mmcsf.cp_loss_grad(fmt, Ud)
    Dir = [rng.standard_normal(u.shape) for u in U]
    h = 1e-4
    U64 = [u.astype(np.float64) for u in U]
    Lp = D.cp_loss(inds, vals, [u + h * d for u, d in zip(U64, Dir)])[0]
    Lm = D.cp_loss(inds, vals, [u - h * d for u, d in zip(U64, Dir)])[0]
    fd = (Lp - Lm) / (2 * h)
    an = sum(float((g.cpu().double().numpy() * d).sum()) for g, d in zip(grads, Dir))
    assert abs(an - fd) / abs(fd) < 1e-6


def test_autograd_matches_explicit_gradients(cuda):
    import torch
    import mmcsf
    rng = np.random.default_rng(5)
    dims = [200, 150, 100]
    inds, vals, fmt, U, Ud = setup(rng, dims, 10000, 16, torch, cuda)
    P = [u.clone().requires_grad_(True) for u in Ud]
    loss = mmcsf.cp_completion_loss(fmt, P)
    (3.0 * loss).backward()
    _, grads = mmcsf.cp_loss_grad(fmt, Ud)
    for p, g in zip(P, grads):
        assert rel(p.grad.cpu().numpy(), 3.0 * g.cpu().numpy()) < 1e-6
    # only the factors that need a gradient get one
    Q = [Ud[0].clone().requires_grad_(True), Ud[1], Ud[2]]
    mmcsf.cp_completion_loss(fmt, Q).backward()
    assert rel(Q[0].grad.cpu().numpy(), grads[0].cpu().numpy()) < 1e-6
    with torch.no_grad():
        l2 = mmcsf.cp_completion_loss(fmt, Ud)
    assert abs(l2.item() - loss.item()) / loss.item() < 1e-6


def test_adam_training_matches_float64_autograd(cuda):
    """100 Adam steps (lr 1e-2) on a planted rank-4 tensor: the GPU loss curve
    follows a float64 CPU torch-autograd run from the same start."""
    import torch
    import mmcsf
    rng = np.random.default_rng(11)
    dims = [300, 200, 100]
    R = 4
    inds, _ = random_tensor(rng, dims, 20000, [1.1, 1.1, 0])
    A = [rng.standard_normal((d, R)) / np.sqrt(R) for d in dims]
    vals = (D.cp_model_values(inds, A) + 0.01 * rng.standard_normal(len(inds[0]))).astype(np.float32)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    fmt = mmcsf.Format(t, 12, fbr_threshold=128)
    U0 = [(0.5 * rng.standard_normal((d, R))).astype(np.float32) for d in dims]

    P = [torch.from_numpy(u).to(cuda).requires_grad_(True) for u in U0]
    opt = torch.optim.Adam(P, lr=1e-2)
    gpu = []
    for _ in range(100):
        opt.zero_grad()
        loss = mmcsf.cp_completion_loss(fmt, P)
        loss.backward()
        opt.step()
        gpu.append(loss.item())

    I = [torch.from_numpy(np.asarray(i, dtype=np.int64)) for i in inds]
    v = torch.from_numpy(vals.astype(np.float64))
    Q = [torch.from_numpy(u.astype(np.float64)).requires_grad_(True) for u in U0]
    opt2 = torch.optim.Adam(Q, lr=1e-2)
    ref = []
    for _ in range(100):
        opt2.zero_grad()
        model = (Q[0][I[0]] * Q[1][I[1]] * Q[2][I[2]]).sum(1)
        loss = 0.5 * ((model - v) ** 2).sum()
        loss.backward()
        opt2.step()
        ref.append(loss.item())
    gpu, ref = np.array(gpu), np.array(ref)
    assert ref[-1] < 0.5 * ref[0]  # the run trains
    dev_ = np.max(np.abs(gpu - ref) / ref)
    print(f"max relative deviation of the loss curve: {dev_:.3e}")
    assert dev_ < 1e-3


def test_error_paths(cuda):
    import torch
    import mmcsf
    rng = np.random.default_rng(1)
    dims = [50, 40, 30]
    inds, vals, fmt, U, Ud = setup(rng, dims, 2000, 8, torch, cuda, trees=False)
    # without mode trees only the root mode (modeOrder[0]) has a tree
    root = fmt.info()["modeOrder"][0]
    other = (root + 1) % 3
    mmcsf.cp_adjoint(fmt, root, Ud)
    with pytest.raises(mmcsf.MMCSFError, match="no tree rooted"):
        mmcsf.cp_adjoint(fmt, other, Ud)
    with pytest.raises(mmcsf.MMCSFError, match="root tree per mode"):
        mmcsf.cp_loss_grad(fmt, Ud)
    # rank must be a multiple of 4
    _, _, fmt2, _, Ud2 = setup(rng, dims, 2000, 6, torch, cuda)
    with pytest.raises(mmcsf.MMCSFError, match="R % 4"):
        mmcsf.cp_forward(fmt2, Ud2)
    # HB-CSF holds no positions
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    hyb = mmcsf.Format(t, 10)
    with pytest.raises(mmcsf.MMCSFError, match="B-CSF or MM-CSF"):
        mmcsf.cp_forward(hyb, Ud)


def test_empty_tensor(cuda):
    import torch
    import mmcsf
    dims = [10, 20, 30]
    t = mmcsf.Tensor.from_numpy(dims, [np.zeros(0, np.int64)] * 3, np.zeros(0, np.float32))
    fmt = mmcsf.Format(t, 12, fbr_threshold=128)
    Ud = [torch.ones((d, 8), device=cuda) for d in dims]
    loss, grads = mmcsf.cp_loss_grad(fmt, Ud)
    assert loss.item() == 0.0
    for g in grads:
        assert float(g.abs().sum()) == 0.0
