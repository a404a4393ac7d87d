"""Pins the path-C oracle (oracle/knn_oracle.cpp) without a GPU.

The reference's exact method is a kd-tree (snapatac2-core/src/utils/knn.rs:9-32);
scipy's cKDTree is an independent exact kd-tree, so on data without distance
ties both must give identical neighbour sets and distances. Layout rules come
from to_csr_matrix (knn.rs:58-78): k entries per row, sorted by neighbour index,
values = Euclidean distance, self excluded by index (knn.rs:27).
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import knn as oknn


def clustered(n, d=30, n_clusters=10, seed=0):
    rng = np.random.default_rng(seed)
    centres = 3.0 * rng.standard_normal((n_clusters, d))
    lab = rng.integers(0, n_clusters, n)
    X = centres[lab] + rng.standard_normal((n, d))
    return X / np.linalg.norm(X, axis=1, keepdims=True)


@pytest.mark.parametrize("n,d,k", [(500, 30, 50), (2000, 10, 25), (60, 3, 50)])
def test_matches_ckdtree(n, d, k):
    X = clustered(n, d, seed=n)
    g = oknn.knn_graph(X, k)
    kk = min(k, n - 1)
    dist, idx = cKDTree(X).query(X, k=kk + 1)
    for i in range(n):
        sel = [(j, dd) for j, dd in zip(idx[i], dist[i]) if j != i][:kk]
        sel.sort()
        row = slice(g.indptr[i], g.indptr[i + 1])
        assert g.indices[row].tolist() == [j for j, _ in sel]
        np.testing.assert_allclose(g.data[row], [dd for _, dd in sel], rtol=1e-12)


def test_layout_and_self_exclusion():
    # 4 points on a line: 0, 1, 3, 7
    X = np.array([[0.0], [1.0], [3.0], [7.0]])
    g = oknn.knn_graph(X, 2)
    assert g.indptr.tolist() == [0, 2, 4, 6, 8]
    assert g.indices.tolist() == [1, 2, 0, 2, 0, 1, 1, 2]
    assert g.data.tolist() == [1.0, 3.0, 1.0, 2.0, 3.0, 2.0, 6.0, 4.0]


def test_duplicate_points_tie_break_by_index():
    X = np.array([[0.0, 0.0], [0.0, 0.0], [0.0, 0.0], [5.0, 5.0]])
    g = oknn.knn_graph(X, 1)
    assert g.indices.tolist() == [1, 0, 0, 0]
    assert g.data[:3].tolist() == [0.0, 0.0, 0.0]
