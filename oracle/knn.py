"""ORACLE (test infrastructure only) for path C, the exact kNN graph.

knn_graph(): FP64 brute force (knn_oracle.cpp) restating
             snapatac2-core/src/utils/knn.rs:9-32,58-78 -> scipy CSR like
             nearest_neighbour_graph returns (snapatac2-python/src/knn.rs:7-16).
Pinned by tests/test_knn_oracle.py: scipy.spatial.cKDTree (an independent exact
kd-tree, the reference's algorithm class) on tie-free data, hand-derived small
cases, and the CSR layout rules of to_csr_matrix.
Parity unpinned: the order kdtree-0.7 gives to exactly equal distances (the
oracle breaks such ties by index).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import LIB_PATH, build

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        h.oracle_knn.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p,
                                 C.c_void_p]
        _lib = h
    return _lib


def knn_arrays(X: np.ndarray, k: int):
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, d = X.shape
    kk = min(k, n - 1)
    nbr = np.empty((n, kk), dtype=np.int32)
    dist = np.empty((n, kk), dtype=np.float64)
    lib().oracle_knn(X.ctypes.data, n, d, kk, nbr.ctypes.data, dist.ctypes.data)
    return nbr, dist


def knn_graph(X: np.ndarray, k: int) -> sp.csr_matrix:
    nbr, dist = knn_arrays(X, k)
    n, kk = nbr.shape
    indptr = np.arange(0, n * kk + 1, kk, dtype=np.int64)
    return sp.csr_matrix((dist.ravel(), nbr.ravel(), indptr), shape=(n, n))
