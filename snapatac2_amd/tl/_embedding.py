"""Drop-in for snapatac2.tl.spectral (cosine distance, full matrix).

Reference: snapatac2-python/python/snapatac2/tools/_embedding.py:129-295 and
spectral_embedding / spectral_mf (snapatac2-python/src/embedding.rs:24-59,
131-191).
"""
from __future__ import annotations

import logging
from typing import Literal

import numpy as np

from .. import _lib
from ._lanczos import _fix_signs, eigsh_lm
from ._operator import SpectralOperator, _to_device_csr, select_columns


def _selected_columns(features, n_vars: int):
    """Ascending column ids and the permutation applied to user weights."""
    if features is None:
        return None, None
    f = np.asarray(features)
    if f.dtype == bool:
        if f.shape != (n_vars,):
            raise ValueError("boolean feature mask has the wrong length")
        return np.flatnonzero(f), None
    f = f.astype(np.int64)
    if np.any(f < 0):
        f = np.where(f < 0, f + n_vars, f)
    if len(np.unique(f)) != len(f):
        raise ValueError("duplicated feature indices")
    perm = np.argsort(f, kind="stable")
    return f[perm], perm


def spectral_embedding(adata, features, n_components: int, random_state: int = 0,
                       feature_weights=None, return_stats: bool = False, device_output: bool = False):
    """(evals, evecs) of the matrix-free normalized cosine operator, like the
    reference's internal.spectral_embedding: eigenpairs of largest magnitude,
    sorted by descending eigenvalue."""
    torch = _lib.require_gpu()
    X = _to_device_csr(adata.X, torch)
    cols, perm = _selected_columns(features, X.shape[1])
    Xs = select_columns(X, cols)
    w = None
    if feature_weights is not None:
        w = np.asarray(feature_weights, dtype=np.float64)
        if perm is not None:
            w = w[perm]
    op = SpectralOperator(Xs, w)
    evals, U, stats = eigsh_lm(op, n_components, seed=random_state)
    ix = np.argsort(evals, kind="stable")[::-1]
    evals = evals[ix]
    U = U[:, torch.from_numpy(ix.copy()).to(U.device)]
    if device_output:
        evecs = U
    else:
        evecs = _fix_signs(U.cpu().numpy())
    if return_stats:
        return evals, evecs, stats
    return evals, evecs


def spectral(
    adata,
    n_comps: int = 30,
    features: str | np.ndarray | None = "selected",
    random_state: int = 0,
    sample_size: int | float | None = None,
    sample_method: Literal["random", "degree"] = "random",
    chunk_size: int = 20000,
    distance_metric: Literal["jaccard", "cosine"] = "cosine",
    weighted_by_sd: bool = True,
    feature_weights: list[float] | None = None,
    inplace: bool = True,
) -> tuple[np.ndarray, np.ndarray] | None:
    """Laplacian eigenmaps of the cosine similarity graph (drop-in for snapatac2.tl.spectral).

    Runs on the GPU for distance_metric='cosine', with the full matrix or, for
    sample_size < n_obs, the Nystrom approximation (landmarks drawn with numpy's
    PCG64 seeded 2023 where the reference uses Rust's StdRng, so the landmark
    set differs; see tl/_nystrom.py). The Jaccard metric is not provided.
    Eigenvector signs follow a fixed convention (largest-magnitude entry
    positive), so they may differ from ARPACK's arbitrary signs.
    """
    np.random.seed(random_state)
    if isinstance(features, str):
        if features in adata.var:
            features = adata.var[features].to_numpy()
        else:
            raise NameError("Please call `select_features` first or explicitly set `features = None`")
    n_comps = min(adata.n_vars - 1, adata.n_obs - 1, n_comps)
    n_sample, _ = adata.shape
    if sample_size is None:
        sample_size = n_sample
    elif isinstance(sample_size, int):
        if sample_size <= 1:
            raise ValueError("when sample_size is an integer, it should be > 1")
        if sample_size > n_sample:
            sample_size = n_sample
    else:
        if sample_size <= 0.0 or sample_size > 1.0:
            raise ValueError("when sample_size is a float, it should be > 0 and <= 1")
        else:
            sample_size = int(sample_size * n_sample)
    if distance_metric != "cosine":
        raise NotImplementedError("only distance_metric='cosine' is provided")
    if sample_size < n_sample:
        # _embedding.py:257-265: Nystrom, then orthogonalize
        from ._nystrom import orthogonalize, spectral_embedding_nystrom
        torch = _lib.require_gpu()
        X = _to_device_csr(adata.X, torch)
        cols, _ = _selected_columns(features, X.shape[1])
        # the reference's Nystrom call takes no feature weights: always IDF
        v, u = spectral_embedding_nystrom(select_columns(X, cols), n_comps, sample_size,
                                          sample_method != "random", chunk_size)
        evals, evecs = orthogonalize(v, u)
        evecs = _fix_signs(evecs.cpu().numpy())
    else:
        evals, evecs = spectral_embedding(adata, features, n_comps, random_state, feature_weights)
    if weighted_by_sd:
        idx = [i for i in range(evals.shape[0]) if evals[i] > 0]
        evals = evals[idx]
        evecs = evecs[:, idx] * np.sqrt(evals)
    if inplace:
        adata.uns["spectral_eigenvalue"] = evals
        adata.obsm["X_spectral"] = evecs
        return None
    return (evals, evecs)
