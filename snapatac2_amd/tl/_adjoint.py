"""Differentiable normalized-cosine operator and spectral-gap tuning of feature weights.

The forward map is Y = A(w) V with A(w) = D^-1/2 (X~ X~^T - I) D^-1/2,
X~ = rownorm(X diag(w)) and D = diag(X~ X~^T 1 - 1): the operator that
snapatac2-python/src/embedding.rs:131-171 builds (with the weighting of
embedding.rs:314-326) and that `spectral(feature_weights=...)` diagonalises
(snapatac2/tools/_embedding.py:139,249). The reference has no derivative of it;
the backward here is hand-written (csrc/spectral.hip, snapgpu_spectral_vjp_w):

    dV = A(w) G                      (A is symmetric)
    dw = d<G, A(w) V>/dw             (explicit w^2, row norms and degrees)

and the operator's scales and degrees are recomputed from the counts whenever
w changes: X diag(w) is never stored.
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np

from .. import _lib
from ._lanczos import eigsh_lm
from ._operator import SpectralOperator


class TuneResult(NamedTuple):
    weights: object  # FP64 device tensor, length m
    history: list  # (gap, step size) of each accepted iterate, the initial one first
    evals: np.ndarray  # n_components + 1 eigenvalues of the returned weights (LM order)
    vectors: object  # their eigenvectors, n x (n_components + 1) FP64 device tensor


__all__ = ["normalized_cosine", "spectral_gap", "spectral_gap_grad", "tune_feature_weights",
           "TuneResult"]


def _function():
    torch = _lib.require_gpu()

    class NormalizedCosine(torch.autograd.Function):
        """Y = A(w) V through the C-ABI; backward through the hand-written adjoint."""

        @staticmethod
        def forward(ctx, V, w, op):
            _sync_weights(op, w)
            V64 = V.detach().to(torch.float64).contiguous()
            Y = op.matmat(V64)
            ctx.op = op
            ctx.w = w.detach().to(torch.float64).clone()
            ctx.save_for_backward(V64)
            return Y.to(V.dtype)

        @staticmethod
        def backward(ctx, G):
            (V64,) = ctx.saved_tensors
            op = ctx.op
            _sync_weights(op, ctx.w)
            G64 = G.detach().to(torch.float64).contiguous()
            dV = op.matmat(G64).to(G.dtype) if ctx.needs_input_grad[0] else None
            dw = op.vjp_w(V64, G64) if ctx.needs_input_grad[1] else None
            return dV, dw, None

    return NormalizedCosine


_FN = None


def _sync_weights(op: SpectralOperator, w) -> None:
    """Re-prepare `op` when its device weights differ from `w`."""
    torch = op.torch
    w64 = w.detach().to(device=op.w.device, dtype=torch.float64).reshape(-1)
    if w64.numel() != op.m:
        raise ValueError(f"w has {w64.numel()} entries, expected {op.m}")
    if not torch.equal(w64, op.w):
        op.set_weights(w64)


def normalized_cosine(V, w, op: SpectralOperator):
    """Differentiable Y = A(w) V (n x b device tensor V, length-m device tensor w).

    Gradients flow to V and w through the hand-written adjoint kernels."""
    global _FN
    if _FN is None:
        _FN = _function()
    if V.dim() != 2 or V.shape[0] != op.n:
        raise ValueError(f"V must be {op.n} x b")
    return _FN.apply(V, w, op)


def spectral_gap(op: SpectralOperator, n_components: int = 30, seed: int = 0, v0=None):
    """(gap, evals, U): gap = |lambda_k| - |lambda_{k+1}| between the n_components-th
    and the next eigenvalue in the largest-magnitude order used by spectral().
    v0: eigenvectors of a nearby operator to start the eigensolver from."""
    gap, evals, U, _ = _gap_stats(op, n_components, seed, v0)
    return gap, evals, U


def _gap_stats(op: SpectralOperator, n_components: int, seed: int, v0=None, tol: float = 1e-9):
    evals, U, st = eigsh_lm(op, n_components + 1, seed=seed, v0=v0, tol=tol)
    k = n_components
    return float(abs(evals[k - 1]) - abs(evals[k])), evals, U, st


def _checked_gap(op: SpectralOperator, evals, U, st, k: int, pair=None):
    """(gap, uncertainty) of |lambda_k| - |lambda_{k+1}| for an acceptance test.
    With integer counts the two eigenvalues are re-estimated as FP64 Rayleigh
    quotients of their Lanczos vectors, so the FP32 products' rounding (about
    1e-7 of |lambda_1|, the size of the gap at a million cells) drops out; what
    remains is the Rayleigh quotient's error, at most residual^2 / gap per
    eigenvalue, plus FP64 summation rounding. `pair`: the (lambda_k,
    lambda_{k+1}) that _pair_f64 already returned for these vectors."""
    res = float(st.get("max_residual", 0.0)) if isinstance(st, dict) else 0.0
    if op.value_type != 0:
        gap = float(abs(evals[k - 1]) - abs(evals[k]))
        return gap, 2.0 * res + 1e-6 * abs(float(evals[0]))
    l0, l1 = pair if pair is not None else _pair_f64(op, U, evals, k)[1]
    gap = abs(l0) - abs(l1)
    return gap, 2.0 * res * res / max(abs(gap), 1e-300) + 1e-12 * abs(float(evals[0]))


def spectral_gap_grad(op: SpectralOperator, n_components: int = 30, seed: int = 0, v0=None):
    """(gap, dgap/dw, evals, U) by Hellmann-Feynman: d lambda_i = u_i^T dA u_i, i.e. the adjoint
    with V = [u_k, u_{k+1}] and G = [sign(lambda_k) u_k, -sign(lambda_{k+1}) u_{k+1}]."""
    gap, g, evals, U, _, _ = _gap_grad_stats(op, n_components, seed, v0)
    return gap, g, evals, U


def _pair_f64(op: SpectralOperator, U, evals, k: int):
    """The k-th and (k+1)-th Lanczos pairs refined in FP64: Rayleigh-Ritz of the
    FP64 operator on their 2-dimensional span. When the two eigenvalues are
    close, the FP32 products' rounding E mixes the two vectors by an angle of
    about |E| / gap (0.1 rad at a million cells), which the Hellmann-Feynman
    gradient of each eigenvalue would inherit; the 2 x 2 rotation removes that
    mixing. Returns (V, (lambda_k, lambda_{k+1})); V unchanged (and the Lanczos
    eigenvalues) for float-valued matrices."""
    V = U[:, k - 1:k + 1].contiguous()
    H = op.gram_f64(V)  # V^T A V (2 x 2) from one FP64 sparse product
    if H is None:
        return V, (float(evals[k - 1]), float(evals[k]))
    H = H.cpu().numpy()
    H = 0.5 * (H + H.T)
    lam, Q = np.linalg.eigh(H)
    order = np.argsort(-np.abs(lam), kind="stable")
    lam, Q = lam[order], Q[:, order]
    V = (V @ op.torch.from_numpy(np.ascontiguousarray(Q)).to(V.device)).contiguous()
    return V, (float(lam[0]), float(lam[1]))


def _gap_grad_stats(op: SpectralOperator, n_components: int, seed: int, v0=None, tol: float = 1e-9):
    torch = op.torch
    gap, evals, U, st = _gap_stats(op, n_components, seed, v0, tol)
    k = n_components
    V, (l0, l1) = _pair_f64(op, U, evals, k)
    sg = torch.tensor([np.sign(l0) or 1.0, -(np.sign(l1) or 1.0)], dtype=torch.float64, device=V.device)
    G = (V * sg).contiguous()
    return abs(l0) - abs(l1), op.vjp_w(V, G), evals, U, st, (l0, l1)


def tune_feature_weights(op: SpectralOperator, n_components: int = 30, steps: int = 10,
                         step_size: float = 0.05, seed: int = 0, min_step: float = 1e-4,
                         tol: float = 1e-9, grad_tol: float = 1e-7):
    """Gradient ascent on the eigengap |lambda_k| - |lambda_{k+1}| over w >= 0.

    Each step moves w by step_size * mean(w) * g / max|g| and is kept only if
    the gap grows by more than its uncertainty (otherwise the step size is
    halved): the gaps compared are FP64 Rayleigh quotients of the two Lanczos
    vectors, and the margin is their error bound (_checked_gap), so an accepted
    step is a real increase, not rounding of the FP32 products. The returned gap
    is never below the one of the initial weights. Returns a TuneResult: the
    weights, the (gap, step_size) history of the accepted iterates (FP64 gaps),
    and the n_components + 1 eigenpairs of largest magnitude for the returned
    weights, so a caller can embed with them without another eigensolve. `op` is
    left prepared for the returned weights. Each trial eigensolve starts from
    the eigenvectors of the current weights.

    Eigensolves whose pairs feed a gradient stop at the relative residual
    grad_tol, the last trial's at tol: the acceptance margin for a residual r is
    2 r^2 / gap (6e-9 for r = 1e-7 against the gap of 1.5e-6 at a million
    cells), far below any gap change a step is kept for, so the rounds from
    grad_tol to tol would not change a decision. The eigenpairs returned are
    always at tol: when they come from a grad_tol solve (no trial kept after
    it), a solve at tol started from them replaces them."""
    torch = op.torch
    k = n_components
    w = op.w.clone()
    first_tol = grad_tol if steps > 0 else tol
    _, g, evals, U, st, pair = _gap_grad_stats(op, n_components, seed, tol=first_tol)
    precise = first_tol <= tol  # whether evals, U were solved to tol
    gap, unc = _checked_gap(op, evals, U, st, k, pair)
    history = [(gap, 0.0)]
    for it in range(steps):
        if step_size < min_step:
            break
        gmax = float(g.abs().max().item())
        if gmax == 0.0:
            break
        scale = step_size * float(w.mean().item()) / gmax
        trial = torch.clamp(w + scale * g, min=0.0)
        last = it == steps - 1
        try:
            op.set_weights(trial)
            if last:  # no further step: the gradient at the trial point is not needed
                _, new_evals, new_U, new_st = _gap_stats(op, n_components, seed, v0=U, tol=tol)
                new_g, new_pair, new_precise = None, None, True
            else:
                _, new_g, new_evals, new_U, new_st, new_pair = _gap_grad_stats(op, n_components, seed, v0=U,
                                                                               tol=grad_tol)
                new_precise = grad_tol <= tol
            new_gap, new_unc = _checked_gap(op, new_evals, new_U, new_st, k, new_pair)
        except _lib.PanicException:
            new_gap, new_unc = -np.inf, 0.0
        if new_gap - new_unc > gap + unc:
            w, gap, unc, g, evals, U = trial, new_gap, new_unc, new_g, new_evals, new_U
            precise = new_precise
            history.append((gap, step_size))
            step_size *= 1.5
        else:
            step_size *= 0.5
    if not torch.equal(op.w, w):
        op.set_weights(w)
    if not precise:
        evals, U, _ = eigsh_lm(op, n_components + 1, seed=seed, v0=U, tol=tol)
    return TuneResult(w, history, evals, U)
