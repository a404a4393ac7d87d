"""Drop-in for snapatac2.pp.knn (exact kd-tree method).

Reference: snapatac2-python/python/snapatac2/preprocessing/_knn.py:10-87 and
nearest_neighbour_graph (snapatac2-python/src/knn.rs:7-16,
snapatac2-core/src/utils/knn.rs:9-32,58-78).
"""
from __future__ import annotations

import ctypes as C
from typing import Literal

import numpy as np
from scipy.sparse import csr_matrix

from .. import _lib

_MODES = {"fp32": 0, "bf16": 1}


def nearest_neighbour_graph(data, k: int, precision: str = "bf16", force_fallback: bool = False,
                            return_stats: bool = False, prune: bool = True):
    """Exact kNN graph of the rows of `data` (n x d): scipy CSR n x n, float64
    Euclidean distances, k entries per row sorted by neighbour index, self
    excluded by index. `data` may be a numpy array or a CUDA torch tensor.
    For d <= 30 and k <= 100, `precision` picks the candidate arithmetic (bf16
    MFMA or fp32) ahead of the exact FP64 rescoring; wider rows or more
    neighbours take an exact FP64 all-pairs scan (both precisions).
    prune=False makes the bf16 path visit every database chunk (same result)."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    if precision not in _MODES:
        raise ValueError("precision must be 'fp32' or 'bf16'")
    if isinstance(data, torch.Tensor):
        X = data.to(device="cuda", dtype=torch.float64).contiguous()
    else:
        X = torch.from_numpy(np.ascontiguousarray(data, dtype=np.float64)).cuda()
    if X.dim() != 2:
        raise ValueError("data must be 2-dimensional")
    n, d = X.shape
    kk = min(int(k), n - 1)
    if kk < 1:
        # the kd-tree returns no neighbours for a single point
        out = csr_matrix((n, n), dtype=np.float64)
        return (out, {"fallback_rows": 0, "exact_scan_rows": 0, "chunk_visits": 0,
                      "chunk_pairs": 0}) if return_stats else out
    nbr = torch.empty((n, kk), dtype=torch.int32, device=X.device)
    dist = torch.empty((n, kk), dtype=torch.float64, device=X.device)
    ws_b = int(L.snapgpu_knn_workspace_bytes(n, d, kk))
    ws = torch.empty(ws_b, dtype=torch.uint8, device=X.device)
    nfb = C.c_int64(0)
    nex = C.c_int64(0)
    nvis = C.c_int64(0)
    flags = (1 if force_fallback else 0) | (0 if prune else 2)
    _lib.check(L.snapgpu_knn(X.data_ptr(), n, d, d, kk, _MODES[precision], nbr.data_ptr(),
                             dist.data_ptr(), ws.data_ptr(), ws_b, flags, C.byref(nfb), C.byref(nex),
                             C.byref(nvis), _lib.stream_ptr()))
    # page-locked host copies (the returned arrays keep them alive): one DMA each
    # instead of staging through pageable memory
    hd = torch.empty((n, kk), dtype=torch.float64, pin_memory=True)
    hn = torch.empty((n, kk), dtype=torch.int32, pin_memory=True)
    hd.copy_(dist, non_blocking=True)
    hn.copy_(nbr, non_blocking=True)
    torch.cuda.current_stream().synchronize()
    ip_t = np.int32 if n * kk < 2**31 else np.int64
    indptr = np.arange(0, n * kk + 1, kk, dtype=ip_t)
    out = csr_matrix((hd.numpy().ravel(), hn.numpy().ravel().astype(ip_t, copy=False), indptr),
                     shape=(n, n), copy=False)
    nwg = -(-n // 128)
    stats = {"fallback_rows": int(nfb.value), "exact_scan_rows": int(nex.value),
             "chunk_visits": int(nvis.value), "chunk_pairs": nwg * nwg if precision == "bf16" else 0}
    return (out, stats) if return_stats else out


def knn(
    adata,
    n_neighbors: int = 50,
    use_dims: int | list[int] | None = None,
    use_rep: str = "X_spectral",
    method: Literal["kdtree", "hora", "pynndescent"] = "kdtree",
    inplace: bool = True,
    random_state: int = 0,
    precision: Literal["fp32", "bf16"] = "bf16",
) -> csr_matrix | None:
    """Compute a neighborhood graph of observations (drop-in for snapatac2.pp.knn).

    method='kdtree' (exact) runs on the GPU; `precision` selects the candidate
    distance arithmetic, and both settings return the exact FP64 neighbour sets.
    The approximate methods ('hora', 'pynndescent') are not provided.
    """
    if hasattr(adata, "obsm"):
        data = adata.obsm[use_rep]
    else:
        inplace = False
        data = adata
    size = data.numel() if hasattr(data, "numel") else np.size(data)
    if size == 0:
        raise ValueError("matrix is empty")
    if use_dims is not None:
        if isinstance(use_dims, int):
            data = data[:, :use_dims]
        else:
            data = data[:, use_dims]
    if method == "kdtree":
        adj = nearest_neighbour_graph(data, n_neighbors, precision=precision)
    elif method in ("hora", "pynndescent"):
        raise NotImplementedError(f"approximate method {method!r} is not provided; use 'kdtree'")
    else:
        raise ValueError("method must be one of 'hora', 'pynndescent', 'kdtree'")
    if inplace:
        adata.obsp["distances"] = adj
        return None
    return adj
