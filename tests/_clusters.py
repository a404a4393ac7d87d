"""Seeded clustered count matrices shared by the spectral and Nystrom tests."""
import numpy as np
import scipy.sparse as sp


def clustered_counts(n=3000, m=1500, clusters=4, seed=0, return_labels=False):
    """n cells x m features; each cell draws a multinomial of 80-400 counts from
    one of `clusters` Gamma(0.3) feature profiles."""
    rng = np.random.default_rng(seed)
    prof = rng.gamma(0.3, 1.0, size=(clusters, m))
    prof /= prof.sum(axis=1, keepdims=True)
    lab = rng.integers(0, clusters, n)
    depth = rng.integers(80, 400, n)
    rows = [rng.multinomial(depth[i], prof[lab[i]]) for i in range(n)]
    X = sp.csr_matrix(np.asarray(rows, dtype=np.int64))
    return (X, lab) if return_labels else X
