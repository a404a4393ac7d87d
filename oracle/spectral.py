"""ORACLE (test infrastructure only) for path B, the spectral embedding.

Restates, in numpy/scipy FP64:
  idf            snapatac2-python/src/embedding.rs:269-286
  normalize      embedding.rs:314-326
  spectral_mf    embedding.rs:131-153 (degrees, scaling) and the embedded
                 eigen() of :155-171, including its exact scipy call:
                 numpy.random.seed(seed); eigsh(A, k=k, v0=numpy.random.rand(n)),
                 sorted by descending eigenvalue
  spectral       snapatac2-python/python/snapatac2/tools/_embedding.py:223-293
                 (n_comps clamp, weighted_by_sd)
The third-party arithmetic (ARPACK via scipy.sparse.linalg.eigsh, scipy 1.15 in
this image, within the reference's pin scipy>=1.4,<2.0) is called exactly as the
reference calls it.

Pinned by tests/test_spectral_oracle.py: hand-derived weights and similarity
entries (w, not sqrt(w); self-similarity removed), the closed-form eigenpair
lambda = 1 with eigenvector D^1/2 1, dense numpy.linalg.eigh of the assembled
operator, the IDF special cases, and bitwise determinism
(snapatac2-python/tests/test_tools.py:95-122).
nystrom is pinned by the eigen-equation with every cell a landmark (each output
row parallel to D^1/2 U + D^-1/2 U L^-1; element-wise, the degree step included,
through sum_j l_j q_j q_j^T = M + I + (M - I)^-1 summed over each row's chunk,
M = S S^T) and by its distance to the exact embedding on clustered cells
(orthogonalized top-3 subspace within 0.05 rad), which an inverted landmark
degree scaling fails.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def idf(X: sp.csr_matrix) -> np.ndarray:
    n = X.shape[0]
    df = np.bincount(X.indices, minlength=X.shape[1]).astype(np.float64)
    if np.all(df == df[0]):
        return np.ones(X.shape[1])
    df = np.where(df == 0, 1.0, df)
    df = np.where(df == n, n - 1.0, df)
    return np.log(n / df)


def normalize(X: sp.csr_matrix, w: np.ndarray) -> sp.csr_matrix:
    X = sp.csr_matrix(X, dtype=np.float64, copy=True)
    X.data *= w[X.indices]
    for i in range(X.shape[0]):
        a, b = X.indptr[i], X.indptr[i + 1]
        if b > a:
            norm = np.sqrt(np.sum(X.data[a:b] * X.data[a:b]))
            X.data[a:b] /= norm
    return X


def operator_parts(X, features=None, feature_weights=None):
    """(X^, dinv, w, Xnorm): the scaled matrix and D of eigen() (embedding.rs:131-153)."""
    X = sp.csr_matrix(X)
    if features is not None:
        X = X[:, features]
    X = sp.csr_matrix(X, dtype=np.float64)
    X.sort_indices()
    w = idf(X) if feature_weights is None else np.asarray(feature_weights, dtype=np.float64)
    Xn = normalize(X, w)
    col_sum = np.asarray(Xn.sum(axis=0)).ravel()
    d = Xn @ col_sum - 1.0
    with np.errstate(divide="ignore"):
        dinv = 1.0 / d
    Xh = sp.diags(np.sqrt(np.where(np.diff(Xn.indptr) > 0, dinv, 0.0))) @ Xn
    return sp.csr_matrix(Xh), dinv, w, Xn


def dense_operator(X, features=None, feature_weights=None) -> np.ndarray:
    Xh, dinv, _, _ = operator_parts(X, features, feature_weights)
    A = (Xh @ Xh.T).toarray()
    A[np.diag_indices_from(A)] -= dinv
    return A


def eigen(Xh: sp.csr_matrix, D: np.ndarray, k: int, seed: int):
    """The embedded eigen() of embedding.rs:155-171, verbatim in behaviour."""
    from scipy.sparse.linalg import LinearOperator, eigsh
    np.random.seed(seed)

    def f(v):
        return Xh @ (v.T @ Xh).T - D * v

    n = Xh.shape[0]
    A = LinearOperator((n, n), matvec=f, dtype=np.float64)
    evals, evecs = eigsh(A, k=k, v0=np.random.rand(n))
    ix = evals.argsort()[::-1]
    return evals[ix], evecs[:, ix]


def spectral_embedding(X, features, n_components, random_state=0, feature_weights=None):
    Xh, dinv, _, _ = operator_parts(X, features, feature_weights)
    return eigen(Xh, dinv, n_components, random_state)


def spectral(X, n_comps=30, features=None, random_state=0, weighted_by_sd=True,
             feature_weights=None):
    n_obs, n_vars = X.shape
    n_comps = min(n_vars - 1, n_obs - 1, n_comps)
    evals, evecs = spectral_embedding(X, features, n_comps, random_state, feature_weights)
    if weighted_by_sd:
        idx = [i for i in range(evals.shape[0]) if evals[i] > 0]
        evals = evals[idx]
        evecs = evecs[:, idx] * np.sqrt(evals)
    return evals, evecs


def subspace_angle(U: np.ndarray, V: np.ndarray) -> float:
    """Largest principal angle (radians) between span(U) and span(V)."""
    qu, _ = np.linalg.qr(U)
    qv, _ = np.linalg.qr(V)
    s = np.linalg.svd(qu.T @ qv, compute_uv=False)
    return float(np.arccos(np.clip(s.min(), -1.0, 1.0)))


def _landmark_scale(u, v, dinv):
    """nystrom's eigenvector scaling (embedding.rs:203-214): row i times
    1/sqrt(degree_i) with degree = 1/dinv (embedding.rs:189-190), column c over v_c."""
    return u * np.sqrt(dinv).reshape((-1, 1)) / v.reshape((1, -1))


def nystrom(X, landmark_rows, n_components: int, chunk_size: int = 20000, feature_weights=None):
    """spectral_embedding_nystrom (snapatac2-python/src/embedding.rs:60-131) given
    the landmark rows, then nystrom (:193-268) over row chunks of chunk_size:
      weights = idf_from_chunks (:288-312) over all rows, or the given weights;
      seed = normalize(X[landmarks], weights); (v, u, dinv) = spectral_mf(seed);
      degrees = 1/dinv (:189-190); u rows * 1/sqrt(degrees) = sqrt(dinv)
      (:205-209), u columns / v (:211-214);
      per chunk: q = sample @ (seed.T @ u); t = q.sum(0) * v; d = q @ t;
      d[d <= 0] = min(d[d > 0]); q /= sqrt(d).
    X: cells x selected features, integer counts. Returns (v, q)."""
    X = sp.csr_matrix(X, dtype=np.float64)
    X.sort_indices()
    w = idf(X) if feature_weights is None else np.asarray(feature_weights, dtype=np.float64)
    seed = normalize(X[np.asarray(landmark_rows)], w)
    col_sum = np.asarray(seed.sum(axis=0)).ravel()
    dinv = 1.0 / (seed @ col_sum - 1.0)
    seed_h = sp.csr_matrix(sp.diags(np.sqrt(dinv)) @ seed)
    v, u = eigen(seed_h, dinv, n_components, 0)
    u = _landmark_scale(u, v, dinv)
    out = []
    for a in range(0, X.shape[0], chunk_size):
        sample = normalize(X[a:a + chunk_size], w)
        q = sample @ (seed.T @ u)
        t = q.sum(axis=0) * v
        d = q @ t.reshape((-1, 1))
        d[d <= 0] = np.min(d[d > 0])
        np.divide(q, np.sqrt(d), out=q)
        out.append(q)
    return v, np.vstack(out)


def orthogonalize(evals, evecs):
    """orthogonalize (snapatac2-python/python/snapatac2/tools/_embedding.py:397-413)."""
    _, sigma, Vt = np.linalg.svd(evecs)
    V = Vt.T
    B = np.multiply(V.T, evals.reshape((1, -1))) @ V
    np.multiply(B, sigma.reshape((-1, 1)), out=B)
    np.multiply(B, sigma.reshape((1, -1)), out=B)
    evals_new, evecs_new = np.linalg.eig(B)
    ix = evals_new.argsort()[::-1]
    evals_new = evals_new[ix]
    evecs_new = evecs_new[:, ix]
    np.divide(evecs_new, sigma.reshape((-1, 1)), out=evecs_new)
    evecs_new = evecs @ V @ evecs_new
    return (evals_new, evecs_new)


def subspace_sin(U: np.ndarray, V: np.ndarray) -> float:
    """Sine of the largest principal angle between span(U) and span(V), as the
    norm of V's orthonormal basis minus its projection on U's (accurate for
    small angles, where arccos of the cosines is not)."""
    qu, _ = np.linalg.qr(U)
    qv, _ = np.linalg.qr(V)
    r = qv - qu @ (qu.T @ qv)
    return float(np.linalg.norm(r, 2))


def eigen_device(T, w, k: int, seed: int = 0):
    """eigen() of embedding.rs:155-171 — scipy eigsh with which='LM',
    v0 = RandomState(seed).rand(n) and its default ncv and tol — on the FP64
    operator A(w) of the counts held by T (oracle.adjoint.ChunkedCounts), whose
    products run on T's device through torch: the reference's own eigensolver
    call on an operator too large for the scipy restatement above. Returns
    (eigenvalues descending, eigenvectors n x k, number of matvecs)."""
    import sys
    import torch
    from scipy.sparse.linalg import LinearOperator, eigsh
    Xh, XhT, d = T.scaled_csr(w)
    count = [0]

    def f(v):
        count[0] += 1
        if count[0] % 500 == 0:
            print(f"eigen_device: {count[0]} matvecs", file=sys.stderr, flush=True)
        x = torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 1)).to(d.device)
        y = torch.sparse.mm(Xh, torch.sparse.mm(XhT, x)) - x / d[:, None]
        return y.cpu().numpy().reshape(-1)

    n = T.n
    np.random.seed(seed)
    A = LinearOperator((n, n), matvec=f, dtype=np.float64)
    evals, evecs = eigsh(A, k=k, v0=np.random.rand(n))
    ix = evals.argsort()[::-1]
    return evals[ix], evecs[:, ix], count[0]


def eigen_parity(evals, U, ref_evals, ref_U, residuals, k: int) -> dict:
    """Eigenpair agreement of a solver (evals, U; residual norms ||A u - lambda u||
    against the FP64 operator) with the reference eigsh call (ref_evals, ref_U,
    at least k + 1 pairs, descending), for the top k pairs by magnitude:
      eig_rel_err      largest relative eigenvalue difference
      gaps             lambda_j - lambda_{j+1} of the reference at j = 10 and k
      separated        for every j <= k whose relative gap (lambda_j - lambda_{j+1})
                       / |lambda_1| is at least 1e-3: sin of the largest angle
                       between the top-j subspaces
      sin_k            the same for all k pairs
      dk_bound_k       its Davis-Kahan bound ||R||_F / delta, delta = smallest
                       |lambda_i| (i <= k) minus |lambda_{k+1}| of the reference"""
    ev = np.asarray(evals)[:k]
    rv = np.asarray(ref_evals)
    out = {"k": k, "eig_rel_err": float(np.max(np.abs(ev - rv[:k]) / np.abs(rv[:k])))}
    gaps = {}
    for j in (10, k):
        if j < rv.size:
            gaps[f"{j}-{j + 1}"] = float(rv[j - 1] - rv[j])
    out["gaps"] = gaps
    sep = {}
    for j in range(1, k):
        if (rv[j - 1] - rv[j]) / abs(rv[0]) >= 1e-3:
            sep[j] = subspace_sin(U[:, :j], ref_U[:, :j])
    out["separated"] = sep
    out["sin_k"] = subspace_sin(U[:, :k], ref_U[:, :k])
    if rv.size > k:
        delta = float(np.min(np.abs(ev)) - abs(rv[k]))
        out["dk_bound_k"] = float(np.linalg.norm(np.asarray(residuals)[:k]) / delta) if delta > 0 else float("inf")
    return out
