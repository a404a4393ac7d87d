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
    evals, U, _ = eigsh_lm(op, n_components + 1, seed=seed, v0=v0)
    k = n_components
    return float(abs(evals[k - 1]) - abs(evals[k])), evals, U


def spectral_gap_grad(op: SpectralOperator, n_components: int = 30, seed: int = 0, v0=None):
    """(gap, dgap/dw, evals, U) by Hellmann-Feynman: d lambda_i = u_i^T dA u_i, i.e. the adjoint
    with V = [u_k, u_{k+1}] and G = [sign(lambda_k) u_k, -sign(lambda_{k+1}) u_{k+1}]."""
    torch = op.torch
    gap, evals, U = spectral_gap(op, n_components, seed, v0)
    k = n_components
    V = U[:, k - 1:k + 1].contiguous()
    sg = torch.tensor([np.sign(evals[k - 1]) or 1.0, -(np.sign(evals[k]) or 1.0)],
                      dtype=torch.float64, device=V.device)
    G = (V * sg).contiguous()
    return gap, op.vjp_w(V, G), evals, U


def tune_feature_weights(op: SpectralOperator, n_components: int = 30, steps: int = 10,
                         step_size: float = 0.05, seed: int = 0, min_step: float = 1e-4):
    """Gradient ascent on the eigengap |lambda_k| - |lambda_{k+1}| over w >= 0.

    Each step moves w by step_size * mean(w) * g / max|g| and is kept only if
    the gap grows (otherwise the step size is halved), so the returned gap is
    never below the one of the initial weights. Returns a TuneResult: the
    weights, the (gap, step_size) history of the accepted iterates, and the
    n_components + 1 eigenpairs of largest magnitude for the returned weights,
    so a caller can embed with them without another eigensolve. `op` is left
    prepared for the returned weights. Each trial eigensolve starts from the
    eigenvectors of the current weights."""
    torch = op.torch
    w = op.w.clone()
    gap, g, evals, U = spectral_gap_grad(op, n_components, seed)
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
                new_gap, new_evals, new_U = spectral_gap(op, n_components, seed, v0=U)
                new_g = None
            else:
                new_gap, new_g, new_evals, new_U = spectral_gap_grad(op, n_components, seed, v0=U)
        except _lib.PanicException:
            new_gap = -np.inf
        if new_gap > gap:
            w, gap, g, evals, U = trial, new_gap, new_g, new_evals, new_U
            history.append((gap, step_size))
            step_size *= 1.5
        else:
            step_size *= 0.5
    if not torch.equal(op.w, w):
        op.set_weights(w)
    return TuneResult(w, history, evals, U)
