"""Nyström spectral embedding (spectral with sample_size < n_obs), on the GPU.

Reference: internal.spectral_embedding_nystrom and nystrom
(snapatac2-python/src/embedding.rs:60-131, 193-268), idf_from_chunks (:288-312),
compute_degrees / compute_probs (:328-365), and orthogonalize
(snapatac2-python/python/snapatac2/tools/_embedding.py:397-413).

Steps: IDF of all cells (unless weights are given); landmark cells; the
landmark matrix's eigenpairs through the ordinary spectral path (operator +
block Lanczos); both of nystrom's sparse products and its per-chunk degree
normalisation in libsnapgpu; orthogonalize's small eigenproblems on the host.

The reference draws landmarks with Rust's StdRng seeded with 2023
(rand::seq::index::sample / sample_weighted). That generator is not
available here, so landmarks come from numpy's PCG64 seeded with 2023 (uniform
without replacement, or Efraimidis-Spirakis weighted sampling with weights 1/d
for sample_method='degree'). Given the same landmarks the result follows the
reference's arithmetic; the landmark draw itself differs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from .._device import DeviceCSR
from ._lanczos import eigsh_lm
from ._operator import SpectralOperator


def _col_counts_idf(X: DeviceCSR, torch):
    """idf_from_chunks (embedding.rs:288-312): document frequency of every column
    over all rows, then the same clamps and ln(n / df) as idf."""
    L = _lib.lib()
    st = _lib.stream_ptr()
    n, m = X.shape
    dev = X.device
    colptr = torch.zeros(m + 1, dtype=torch.int64, device=dev)
    scratch = torch.empty(16, dtype=torch.int32, device=dev)
    _lib.check(L.snapgpu_csr_col_counts(X.indptr.data_ptr(), n, X.indices.data_ptr(), X.nnz, m,
                                        colptr.data_ptr() + 8, scratch.data_ptr(), st))
    wsb = int(L.snapgpu_scan_workspace_bytes(m))
    ws = torch.empty(max(wsb, 8), dtype=torch.uint8, device=dev)
    _lib.check(L.snapgpu_scan_inclusive_i64(colptr.data_ptr() + 8, m, ws.data_ptr(), wsb, st))
    w = torch.empty(m, dtype=torch.float64, device=dev)
    _lib.check(L.snapgpu_spectral_idf(colptr.data_ptr(), m, n, w.data_ptr(), scratch.data_ptr(), st))
    return w


def take_rows(X: DeviceCSR, rows: np.ndarray) -> DeviceCSR:
    """Rows `rows` of X (in that order), on the device."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    st = _lib.stream_ptr()
    dev = X.device
    d_rows = torch.from_numpy(np.ascontiguousarray(rows, dtype=np.int64)).to(dev)
    k = d_rows.numel()
    indptr = torch.empty(k + 1, dtype=torch.int64, device=dev)
    wsb = int(L.snapgpu_csr_take_rows_workspace_bytes(k))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    nnz = C.c_int64(0)
    _lib.check(L.snapgpu_csr_take_rows_count(X.indptr.data_ptr(), X.shape[0], d_rows.data_ptr(), k,
                                             indptr.data_ptr(), ws.data_ptr(), wsb, C.byref(nnz), st))
    idx = torch.empty(nnz.value, dtype=torch.int32, device=dev)
    val = torch.empty(nnz.value, dtype=torch.int32, device=dev)
    _lib.check(L.snapgpu_csr_take_rows_fill(X.indptr.data_ptr(), X.indices.data_ptr(), X.data.data_ptr(),
                                            d_rows.data_ptr(), k, indptr.data_ptr(),
                                            idx.data_ptr() if nnz.value else None,
                                            val.data_ptr() if nnz.value else None, st))
    return DeviceCSR(indptr, idx, val, (k, X.shape[1]), X.data_dtype)


def landmarks(n_obs: int, sample_size: int, probs: np.ndarray | None = None, seed: int = 2023) -> np.ndarray:
    """Landmark cells, ascending: uniform without replacement, or weighted by
    `probs` (Efraimidis-Spirakis keys u^(1/p), the largest sample_size)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if probs is None:
        idx = rng.choice(n_obs, size=sample_size, replace=False)
    else:
        with np.errstate(divide="ignore"):
            keys = np.log(rng.random(n_obs)) / np.asarray(probs, dtype=np.float64)
        idx = np.argsort(-keys, kind="stable")[:sample_size]
    return np.sort(idx).astype(np.int64)


def _degrees(X: DeviceCSR, w, torch):
    """compute_degrees (embedding.rs:328-360): d = X~ (X~^T 1) - 1 over all cells."""
    op = SpectralOperator(X, w.cpu().numpy(), use_panels=False)
    return 1.0 / op.dinv.cpu().numpy()


def spectral_embedding_nystrom(X: DeviceCSR, n_components: int, sample_size: int,
                               weighted_by_degree: bool = False, chunk_size: int = 20000,
                               feature_weights=None, landmark_rows: np.ndarray | None = None,
                               return_landmarks: bool = False):
    """(evals, evecs): the reference's nystrom() output (before orthogonalize)
    for the selected-feature count matrix X (cells x features, uint32 counts)."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    st = _lib.stream_ptr()
    if np.dtype(X.data_dtype) != np.uint32:
        raise NotImplementedError("the Nystrom path takes integer counts")
    n, m = X.shape
    dev = X.device
    if feature_weights is None:
        w = _col_counts_idf(X, torch)
    else:
        w = torch.from_numpy(np.ascontiguousarray(feature_weights, dtype=np.float64)).to(dev)
    if landmark_rows is None:
        probs = None
        if weighted_by_degree:
            d = _degrees(X, w, torch)
            probs = (1.0 / d) / np.sum(1.0 / d)  # compute_probs (embedding.rs:362-365)
        landmark_rows = landmarks(n, sample_size, probs)
    seed = take_rows(X, landmark_rows)
    # landmark embedding (spectral_mf on the normalised landmark rows)
    op = SpectralOperator(seed, w.cpu().numpy())
    evals, U, _ = eigsh_lm(op, n_components, seed=0)
    ix = np.argsort(evals, kind="stable")[::-1]
    evals = np.ascontiguousarray(evals[ix])
    U = U[:, torch.from_numpy(ix.copy()).to(dev)].contiguous()
    k = U.shape[1]
    ev = torch.from_numpy(evals).to(dev)
    # nystrom (embedding.rs:203-214): rows / sqrt(deg) with deg = 1 / dinv (embedding.rs:189-190),
    # i.e. rows * sqrt(dinv); columns / lambda
    E = torch.empty_like(U)
    _lib.check(L.snapgpu_nystrom_landmark_scale(U.data_ptr(), k, U.shape[0], k, op.dinv.data_ptr(), ev.data_ptr(),
                                                E.data_ptr(), k, st))
    # Z = seed~^T E = diag(w) X_seed^T diag(s_seed) E, on the landmark CSC; the
    # column scales are applied to the dense operand first (one gather per entry)
    sE = (op.s[:, None] * E).contiguous()
    Z = torch.empty((m, k), dtype=torch.float64, device=dev)
    _lib.check(L.snapgpu_csr_spmm_scaled_f64(op.colptr.data_ptr(), op.rowidx.data_ptr(), op.cval.data_ptr(),
                                             m, w.data_ptr(), 0, sE.data_ptr(), k, k, Z.data_ptr(), k, st))
    # q = sample~ Z = diag(s) X diag(w) Z for all cells
    s_all = torch.empty(n, dtype=torch.float64, device=dev)
    _lib.check(L.snapgpu_csr_row_inv_norm(X.indptr.data_ptr(), X.indices.data_ptr(), X.data.data_ptr(), n,
                                          w.data_ptr(), s_all.data_ptr(), st))
    Q = torch.empty((n, k), dtype=torch.float64, device=dev)
    wZ = (w[:, None] * Z).contiguous()
    _lib.check(L.snapgpu_csr_spmm_scaled_f64(X.indptr.data_ptr(), X.indices.data_ptr(), X.data.data_ptr(), n,
                                             s_all.data_ptr(), 0, wZ.data_ptr(), k, k, Q.data_ptr(), k, st))
    wsb = int(L.snapgpu_nystrom_normalize_workspace_bytes(n, k, chunk_size))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    _lib.check(L.snapgpu_nystrom_normalize(Q.data_ptr(), k, n, k, ev.data_ptr(), int(chunk_size), ws.data_ptr(),
                                           wsb, st))
    if return_landmarks:
        return evals, Q, landmark_rows
    return evals, Q


def orthogonalize(evals: np.ndarray, evecs):
    """orthogonalize (_embedding.py:397-413) without the n x n factor of a full
    SVD: sigma^2 and V from the k x k Gram matrix evecs^T evecs (FP64 MFMA
    product on the device), B = diag(sigma) V^T diag(evals) V diag(sigma),
    its eigenpairs by descending value, evecs V diag(1/sigma) W (device)."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    st = _lib.stream_ptr()
    n, k = evecs.shape
    if k > 32:
        raise NotImplementedError("orthogonalize takes at most 32 components")
    G = torch.empty((k, k), dtype=torch.float64, device=evecs.device)
    wsb = int(L.snapgpu_gemm_tn_workspace_bytes(n, k, k))
    ws = torch.empty(max(wsb, 8), dtype=torch.uint8, device=evecs.device)
    _lib.check(L.snapgpu_gemm_tn_f64(evecs.data_ptr(), k, evecs.data_ptr(), k, n, k, k, G.data_ptr(),
                                     ws.data_ptr(), wsb, st))
    Gh = G.cpu().numpy()
    lam, V = np.linalg.eigh((Gh + Gh.T) / 2)
    order = np.argsort(lam)[::-1]
    lam, V = lam[order], V[:, order]
    sigma = np.sqrt(np.maximum(lam, 0.0))
    B = (V.T * evals.reshape((1, -1))) @ V
    B *= sigma.reshape((-1, 1))
    B *= sigma.reshape((1, -1))
    evals_new, W = np.linalg.eigh((B + B.T) / 2)
    ix = np.argsort(evals_new)[::-1]
    evals_new, W = evals_new[ix], W[:, ix]
    M = np.ascontiguousarray(V @ (W / sigma.reshape((-1, 1))))
    out = torch.empty((n, k), dtype=torch.float64, device=evecs.device)
    _lib.check(L.snapgpu_gemm_nn_f64(evecs.data_ptr(), k, torch.from_numpy(M).to(evecs.device).data_ptr(), n, k, k,
                                     1.0, 0.0, out.data_ptr(), k, st))
    return evals_new, out
