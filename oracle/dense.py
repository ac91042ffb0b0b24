"""Float64 restatements of CP-ALS and of the CP-completion loss -- TEST
INFRASTRUCTURE ONLY (imported by tests/ and nothing else).

cpd_als follows cpd() of the reference (src/cpd_cpu.cu:484-691):
  * Grams ATA[m] = U[m]^T U[m]                         (ssyrk, 536-547, 649-652)
  * per mode: M = MTTKRP_m(X, U)                        (599)
              the last mode's M is kept for the fit     (615-616)
              U[m] = M (Hadamard_{n != m} ATA[n])^{-1}   (MatrixSolveNormals, 140-208:
                     Cholesky, LU when not SPD)
              column norms into lambda: 2-norm at iteration 0,
              max-norm (largest entry, at least 1) after (626-630, 212-331)
  * fit = 1 - sqrt(|X|^2 + lambda^T (Hadamard_m ATA[m]) lambda
                   - 2 sum_r lambda_r <U[N-1][:, r], M_last[:, r]>) / |X|
    (KruskalTensorFit, 440-465), computed from its definition with the full
    symmetric Hadamard product (the reference reads an unwritten triangle
    here, 380-395, so its own numbers cannot be the target);
  * stop when |fit - fit_prev| < tol after the first iteration (670-672);
  * GetFinalLambda (333-349) folds the final column 2-norms into lambda.
The reference's max-norm sums per-thread maxima (298-312); the definition
(largest entry per column, clamped at 1) is used here.  ALS is invariant to
column scaling, so the fit does not depend on this choice.

ttm64 states the tensor-times-matrix of the leaf mode per leaf fiber
(ttm_cpu.cpp:60-86).

cp_loss / cp_grad state the least-squares CP loss over the observed entries,
L = 1/2 sum_x (v_x - sum_r prod_n U_n[i_n(x)][r])^2, and its gradient,
dL/dU_n = -MTTKRP_n(E) with E the residual tensor on the same nonzeros.
"""
from __future__ import annotations

import numpy as np


def mttkrp64(inds, vals, U, mode):
    """Float64 MTTKRP (sum over nonzeros of v * prod of the other rows)."""
    R = U[0].shape[1]
    c = np.asarray(vals, dtype=np.float64)[:, None] * np.ones((1, R))
    for m in range(len(U)):
        if m != mode:
            c = c * np.asarray(U[m], dtype=np.float64)[np.asarray(inds[m])]
    out = np.zeros((U[mode].shape[0], R))
    np.add.at(out, np.asarray(inds[mode]), c)
    return out


def _solve_normals(G, M):
    """Rows of U with U G = M: Cholesky, LU fallback (MatrixSolveNormals)."""
    try:
        L = np.linalg.cholesky(G)
        Y = np.linalg.solve(L, M.T)
        return np.linalg.solve(L.T, Y).T
    except np.linalg.LinAlgError:
        return np.linalg.solve(G, M.T).T


def cpd_als(inds, vals, dims, U_init, iters, tol=1e-5, mttkrp=None):
    """Returns (fits per iteration, final factors, lambda after GetFinalLambda).
    ``mttkrp(inds, vals, U, mode)`` computes the float64 MTTKRP (default
    ``mttkrp64``; the threaded C oracle for tensors of 1e7+ nonzeros)."""
    mttkrp = mttkrp64 if mttkrp is None else mttkrp
    N = len(dims)
    U = [np.array(u, dtype=np.float64) for u in U_init]
    R = U[0].shape[1]
    ATA = [u.T @ u for u in U]
    xnorm = float(np.sum(np.asarray(vals, dtype=np.float64) ** 2))
    lam = np.ones(R)
    fits = []
    old = 0.0
    for it in range(iters):
        Mlast = None
        for m in range(N):
            M = mttkrp(inds, vals, U, m)
            if m == N - 1:
                Mlast = M
            G = np.ones((R, R))
            for n in range(N):
                if n != m:
                    G = G * ATA[n]
            U[m] = _solve_normals(G, M)
            if it == 0:
                lam = np.sqrt(np.sum(U[m] ** 2, axis=0))
            else:
                lam = np.maximum(U[m].max(axis=0), 1.0)
            U[m] = U[m] / lam
            ATA[m] = U[m].T @ U[m]
        H = np.ones((R, R))
        for n in range(N):
            H = H * ATA[n]
        norm_mats = abs(float(lam @ H @ lam))
        inner = float(np.sum((U[N - 1] * Mlast) @ lam))
        resid = xnorm + norm_mats - 2 * inner
        if resid > 0:
            resid = np.sqrt(resid)
        fit = 1 - resid / np.sqrt(xnorm)
        fits.append(fit)
        if it > 0 and abs(fit - old) < tol:
            break
        old = fit
    final_lam = lam.copy()
    for m in range(N):
        nrm = np.sqrt(np.sum(U[m] ** 2, axis=0))
        U[m] = U[m] / nrm
        final_lam = final_lam * nrm
    return fits, U, final_lam


def ttm64(fbr_ptr_leaf, leaf, vals, U_leaf):
    """TTM on the leaf mode (TTM_CPU, ttm_cpu.cpp:60-86) in float64:
    Y[f] = sum over the nonzeros x of leaf fiber f of v_x * U_leaf[leaf_x]."""
    ptr = np.asarray(fbr_ptr_leaf, dtype=np.int64)
    c = np.asarray(vals, dtype=np.float64)[:, None] * np.asarray(U_leaf, dtype=np.float64)[np.asarray(leaf)]
    nf = len(ptr) - 1
    fiber = np.repeat(np.arange(nf), np.diff(ptr))
    Y = np.zeros((nf, c.shape[1]))
    np.add.at(Y, fiber, c)
    return Y


def cp_model_values(inds, U):
    """sum_r prod_n U_n[i_n][r] for every nonzero (float64)."""
    p = np.ones((len(inds[0]), U[0].shape[1]))
    for m, u in enumerate(U):
        p = p * np.asarray(u, dtype=np.float64)[np.asarray(inds[m])]
    return p.sum(axis=1)


def cp_loss(inds, vals, U):
    e = np.asarray(vals, dtype=np.float64) - cp_model_values(inds, U)
    return 0.5 * float(e @ e), e


def cp_grad(inds, vals, U):
    _, e = cp_loss(inds, vals, U)
    return [-mttkrp64(inds, e, U, m) for m in range(len(U))]
