"""Block Lanczos eigensolver (FP64) for the largest-magnitude eigenpairs.

Replaces the scipy `eigsh(A, k, which='LM')` call of the reference
(snapatac2-python/src/embedding.rs:155-171): ARPACK's implicitly restarted
single-vector Lanczos becomes a block method whose blocks (16 vectors by
default) go through the operator together, so every pass over the sparse matrix
serves a whole block of Krylov vectors. Full reorthogonalisation (classical
Gram-Schmidt, repeated when a column loses most of its norm), CholeskyQR
(repeated when the block is ill-conditioned) and thick restarts keep the basis
orthonormal to FP64 working accuracy. The basis and all tall-skinny products
live on the device (libsnapgpu FP64 kernels); only the small projected matrix
(<= a few hundred square) is diagonalised on the host, as ARPACK does.
"""
from __future__ import annotations

import numpy as np

from .. import _lib


class _Linalg:
    """FP64 tall-skinny products through libsnapgpu."""

    def __init__(self, torch, n: int, max_cols: int, device):
        self.torch, self.n, self.dev = torch, n, device
        self.L = _lib.lib()
        self.ws_b = int(self.L.snapgpu_gemm_tn_workspace_bytes(n, max_cols, 32))
        self.ws = torch.empty(self.ws_b, dtype=torch.uint8, device=device)
        self.cbufs = [torch.empty(max_cols * 32, dtype=torch.float64, device=device) for _ in range(2)]
        self.bbuf = torch.empty(max_cols * 32, dtype=torch.float64, device=device)

    def tn(self, A, B, slot: int = 0):
        """A^T B for device views A (n x ka), B (n x kb <= 32) -> device (ka x kb)."""
        ka, kb = A.shape[1], B.shape[1]
        out = self.cbufs[slot]
        _lib.check(self.L.snapgpu_gemm_tn_f64(A.data_ptr(), A.stride(0), B.data_ptr(), B.stride(0),
                                              self.n, ka, kb, out.data_ptr(), self.ws.data_ptr(),
                                              self.ws_b, _lib.stream_ptr()))
        return out[:ka * kb].view(ka, kb)

    def nn_dev(self, A, Mdev, Cout, alpha=1.0, beta=0.0) -> None:
        """Cout = beta*Cout + alpha * A @ Mdev for a contiguous device matrix (ka x kb <= 32)."""
        ka, kb = Mdev.shape
        _lib.check(self.L.snapgpu_gemm_nn_f64(A.data_ptr(), A.stride(0), Mdev.data_ptr(), self.n,
                                              ka, kb, float(alpha), float(beta), Cout.data_ptr(),
                                              Cout.stride(0), _lib.stream_ptr()))

    def nn(self, A, M: np.ndarray, Cout, alpha=1.0, beta=0.0) -> None:
        """Cout = beta*Cout + alpha * A @ M with host matrix M (ka x kb), column chunks of 32."""
        ka, kb = M.shape
        for c0 in range(0, kb, 32):
            Mc = np.ascontiguousarray(M[:, c0:c0 + 32], dtype=np.float64)
            w = Mc.shape[1]
            self.bbuf[:ka * w].copy_(self.torch.from_numpy(Mc.ravel()))
            self.nn_dev(A, self.bbuf[:ka * w].view(ka, w), Cout[:, c0:c0 + w], alpha, beta)

    def random(self, out, seed: int) -> None:
        _lib.check(self.L.snapgpu_random_f64(int(seed) & 0xFFFFFFFFFFFFFFFF, out.numel(), out.data_ptr(),
                                             _lib.stream_ptr()))


# a Gram-Schmidt pass is repeated when a column keeps less than this fraction of its norm
REORTH_KEEP = 0.05


def _eigh_small(H):
    """Symmetric eigensolve of the projected matrix (single-threaded BLAS: faster at these sizes)."""
    try:
        from threadpoolctl import threadpool_limits
        with threadpool_limits(limits=1, user_api="blas"):
            return np.linalg.eigh(H)
    except ImportError:  # pragma: no cover
        return np.linalg.eigh(H)


def _lm_order(theta: np.ndarray, k: int) -> np.ndarray:
    """Indices of the k largest |theta| (ARPACK which='LM'); ties by larger value."""
    return np.lexsort((-theta, -np.abs(theta)))[:k]


def _fix_signs(U: np.ndarray) -> np.ndarray:
    """Deterministic sign convention: the largest-magnitude entry of each vector is positive."""
    idx = np.argmax(np.abs(U), axis=0)
    sg = np.sign(U[idx, np.arange(U.shape[1])])
    sg[sg == 0] = 1.0
    return U * sg


def dense_eigsh(op, k: int):
    """All-pairs path for small n: A assembled column-block by column-block on the device."""
    torch = op.torch
    n = op.n
    dev = op.r64.device
    A = torch.empty((n, n), dtype=torch.float64, device=dev)
    for c0 in range(0, n, 32):
        w = min(32, n - c0)
        E = torch.zeros((n, w), dtype=torch.float64, device=dev)
        E[c0:c0 + w] = torch.eye(w, dtype=torch.float64, device=dev)
        op.apply(E, A[:, c0:c0 + w], w)
    Ah = A.cpu().numpy()
    Ah = (Ah + Ah.T) / 2
    theta, S = _eigh_small(Ah)
    order = _lm_order(theta, k)
    U = torch.from_numpy(np.ascontiguousarray(S[:, order])).to(dev)
    return theta[order], U, {"method": "dense", "applies": op.n_applies}


def eigsh_lm(op, k: int, block: int = 16, max_basis: int | None = None, tol: float = 1e-9,
             seed: int = 0, max_restarts: int = 400, dense_below: int = 1024, check_every: int = 4,
             v0=None):
    """k largest-magnitude eigenpairs of the symmetric operator `op`.

    v0: optional n x m device tensor (e.g. eigenvectors of a nearby operator).
    The start block is then v0 times a seeded random m x block matrix, so the
    first block steps span v0's columns; otherwise the start block is random.
    Returns (evals (host), evecs (device tensor, n x k), stats) in LM order.
    """
    torch = op.torch
    n = op.n
    if n <= max(dense_below, 4 * block):
        theta, U, st = dense_eigsh(op, k)
        return theta, U, st
    b = block
    if max_basis is None:
        max_basis = max(20 * b, 3 * k + 2 * b)
    Jmax = int(min(max_basis, (n // b - 1) * b))
    Jmax = max(Jmax, 2 * b)
    dev = op.r64.device
    V = torch.zeros((n, Jmax + b), dtype=torch.float64, device=dev)
    W = torch.empty((n, b), dtype=torch.float64, device=dev)
    T = torch.empty((n, 32), dtype=torch.float64, device=dev)
    la = _Linalg(torch, n, Jmax + b, dev)
    H = np.zeros((Jmax + b, Jmax + b))
    rng_seed = int(seed) * 1000003 + 17

    def orth_against(Vb, J):
        for _ in range(2):
            C = la.tn(V[:, :J], Vb)
            la.nn_dev(V[:, :J], C, Vb, alpha=-1.0, beta=1.0)

    def cholqr2(Wsrc, dest):
        """dest = Wsrc R^-1 with R upper (b x b); returns R, or None if rank deficient.
        CholeskyQR leaves an orthogonality error of about eps * cond(Wsrc)^2, so
        the second pass runs only when cond^2 >= 100."""
        Rtot = np.eye(b)
        src = Wsrc
        for it in range(2):
            G = la.tn(src, src).cpu().numpy()
            G = (G + G.T) / 2
            ev = np.linalg.eigvalsh(G)
            if ev[0] <= 1e-26 * max(ev[-1], 1e-300):
                return None
            last = it == 1 or ev[-1] < 100.0 * ev[0]
            Lc = np.linalg.cholesky(G)
            Rinv = np.linalg.inv(Lc.T)
            out = dest if last else T[:, :b]
            la.nn(src, Rinv, out, 1.0, 0.0)
            Rtot = Lc.T @ Rtot
            if last:
                break
            src = out
        return Rtot

    def next_block(J, Wsrc):
        """Orthonormalise Wsrc (already orthogonal to V[:, :J]) into V[:, J:J+b]."""
        nonlocal rng_seed
        R = cholqr2(Wsrc, V[:, J:J + b])
        if R is not None:
            return R
        # invariant subspace reached: continue with fresh random directions
        for _ in range(5):
            rng_seed += 1
            la.random(W, rng_seed)
            orth_against(W, J)
            R2 = cholqr2(W, V[:, J:J + b])
            if R2 is not None:
                return np.zeros((b, b))
        raise RuntimeError("block Lanczos could not extend the Krylov basis")

    # start block
    la.random(W, rng_seed)
    if v0 is not None:
        m0 = int(v0.shape[1])
        M = np.random.default_rng(rng_seed).standard_normal((m0, b))
        la.nn(v0.to(torch.float64).contiguous(), M, W, 1.0, 1e-3)
    R0 = cholqr2(W, V[:, :b])
    if R0 is None:
        raise RuntimeError("degenerate start block")
    J = b
    restarts = 0
    steps = 0
    last_res = prev_res = np.inf
    since_check = 0
    prefetched = False  # the current block's application was launched before a check
    while True:
        steps += 1
        since_check += 1
        if not prefetched:
            op.apply(V[:, J - b:J], W, b)
        prefetched = False
        # classical Gram-Schmidt against the basis; one pass leaves components of
        # about eps * |W| / |W1| along it, so a second pass runs only when some
        # column lost more than 95% of its norm in the first
        n0 = torch.linalg.vector_norm(W, dim=0)
        C1 = la.tn(V[:, :J], W, slot=0)
        la.nn_dev(V[:, :J], C1, W, alpha=-1.0, beta=1.0)
        n1 = torch.linalg.vector_norm(W, dim=0)
        Hcol = C1.cpu().numpy().copy()
        if not bool((n1 >= REORTH_KEEP * n0).all()):
            C2 = la.tn(V[:, :J], W, slot=1)
            la.nn_dev(V[:, :J], C2, W, alpha=-1.0, beta=1.0)
            Hcol += C2.cpu().numpy()
        H[:J, J - b:J] = Hcol
        R = next_block(J, W)
        H[J:J + b, J - b:J] = R
        at_restart = J + b > Jmax
        if not (at_restart or since_check >= check_every or last_res < 100 * tol):
            J += b
            continue
        since_check = 0
        # the next block's application runs on the device while the host solves the
        # projected problem (a thick restart moves that block to V[:, p:p+b] without
        # changing it); a converged check would discard it, so it is skipped when the
        # residual's decay over the last two checks predicts convergence here
        predicted = last_res * (last_res / prev_res) if 0.0 < last_res < prev_res else np.inf
        if predicted > tol:
            op.apply(V[:, J:J + b], W, b)
            prefetched = True
        Hs = (H[:J, :J] + H[:J, :J].T) / 2
        theta, S = _eigh_small(Hs)
        order = _lm_order(theta, k)
        res = np.linalg.norm(R @ S[J - b:J, order], axis=0)
        scale = max(np.abs(theta).max(), 1e-300)
        prev_res, last_res = last_res, res.max() / scale
        if np.all(res <= tol * scale):
            break
        if not at_restart:
            J += b
            continue
        if restarts >= max_restarts:
            raise RuntimeError(f"block Lanczos did not converge (max residual {res.max():.3e})")
        restarts += 1
        # thick restart: keep the wanted Ritz vectors plus the next best by |theta|
        p = min(J - b, max(k + b, (J // 2) // b * b))
        keep = np.lexsort((-theta, -np.abs(theta)))[:p]
        Sk = S[:, keep]
        Y = torch.empty((n, p), dtype=torch.float64, device=dev)
        la.nn(V[:, :J], Sk, Y, 1.0, 0.0)
        F = R @ S[J - b:J, keep]
        V[:, :p].copy_(Y)
        V[:, p:p + b].copy_(V[:, J:J + b])
        del Y
        H[:] = 0.0
        H[:p, :p] = np.diag(theta[keep])
        H[p:p + b, :p] = F
        J = p + b
    U = torch.empty((n, k), dtype=torch.float64, device=dev)
    la.nn(V[:, :J], S[:, order], U, 1.0, 0.0)
    stats = {"method": "block-lanczos", "block": b, "steps": steps, "restarts": restarts,
             "basis": J, "applies": op.n_applies, "max_residual": float(res.max())}
    return theta[order], U, stats
