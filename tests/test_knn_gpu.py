"""Path C on the GPU: neighbour sets and distances vs the FP64 brute-force oracle."""
import numpy as np
import pytest

from oracle import knn as oknn

pytestmark = pytest.mark.gpu


def clustered(n, d=30, n_clusters=10, seed=0):
    rng = np.random.default_rng(seed)
    centres = 3.0 * rng.standard_normal((n_clusters, d))
    lab = rng.integers(0, n_clusters, n)
    X = centres[lab] + rng.standard_normal((n, d))
    return X / np.linalg.norm(X, axis=1, keepdims=True)


def _check(g, X, k):
    o = oknn.knn_graph(X, k)
    assert g.shape == o.shape
    assert np.array_equal(g.indptr, o.indptr)
    assert np.array_equal(g.indices, o.indices)
    # distances: exact FP64 rescoring with the oracle's formula -> identical bits
    assert np.array_equal(g.data, o.data)


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
@pytest.mark.parametrize("n,d,k", [(5000, 30, 50), (3000, 17, 25), (1000, 30, 100), (100, 5, 50)])
def test_clustered_vs_oracle(torch_gpu, precision, n, d, k):
    from snapatac2_amd.pp._knn import nearest_neighbour_graph
    X = clustered(n, d, seed=n + d)
    g, st = nearest_neighbour_graph(X, k, precision=precision, return_stats=True)
    _check(g, X, k)
    if k <= 50:  # 128 candidates leave a wide margin for k <= 50
        assert st["fallback_rows"] <= n // 100 + 1


def test_spectral_scale_embedding(torch_gpu):
    # embeddings like X_spectral: tiny values (eigenvectors * sqrt(eigenvalue))
    from snapatac2_amd.pp._knn import nearest_neighbour_graph
    X = clustered(4000, 30, seed=9) * 1e-3
    X[:, 20:] *= 1e-2
    _check(nearest_neighbour_graph(X, 50), X, 50)


def test_ties_and_duplicates(torch_gpu):
    # integer lattice points: many exactly equal distances, plus duplicate rows
    from snapatac2_amd.pp._knn import nearest_neighbour_graph
    rng = np.random.default_rng(1)
    X = rng.integers(0, 4, (1500, 3)).astype(np.float64)
    for precision in ("bf16", "fp32"):
        _check(nearest_neighbour_graph(X, 30, precision=precision), X, 30)


def test_forced_fallback_is_exact(torch_gpu):
    from snapatac2_amd.pp._knn import nearest_neighbour_graph
    X = clustered(700, 30, seed=3)
    g, st = nearest_neighbour_graph(X, 50, force_fallback=True, return_stats=True)
    assert st["fallback_rows"] == 700
    _check(g, X, 50)


def test_knn_dropin(torch_gpu):
    import snapatac2_amd as snap
    X = clustered(800, 30, seed=5)
    ad = snap.AnnData(n_obs=800, n_vars=1)
    ad.obsm["X_spectral"] = X
    snap.pp.knn(ad, n_neighbors=25)
    _check(ad.obsp["distances"], X, 25)
    g2 = snap.pp.knn(X, n_neighbors=25, use_dims=10)
    _check(g2, X[:, :10], 25)
    # determinism (reference test_reproducibility, test_tools.py:95-122)
    g3 = snap.pp.knn(ad, n_neighbors=25, inplace=False)
    assert (g3 != ad.obsp["distances"]).nnz == 0
    with pytest.raises(ValueError, match="empty"):
        snap.pp.knn(np.zeros((0, 5)))
    with pytest.raises(ValueError, match="method"):
        snap.pp.knn(ad, method="bogus")
    tiny = snap.pp.knn(X[:5], n_neighbors=50)
    assert tiny.nnz == 20


def test_second_tier_on_offset_clusters(torch_gpu):
    # rows far from the origin with small spreads: |q|^2 >> d^2. The uncentred
    # FP32 bound cannot certify them, so the exact FP64 second tier takes over;
    # the BF16 path centres each query block and certifies them directly.
    from snapatac2_amd.pp._knn import nearest_neighbour_graph
    X = 1.0 + 1e-3 * clustered(3000, 20, seed=4)
    g, st = nearest_neighbour_graph(X, 50, precision="fp32", return_stats=True)
    _check(g, X, 50)
    assert st["fallback_rows"] > 0
    assert st["exact_scan_rows"] <= st["fallback_rows"]
    g, st = nearest_neighbour_graph(X, 50, precision="bf16", return_stats=True)
    _check(g, X, 50)
    assert st["fallback_rows"] <= 30


@pytest.mark.parametrize("n,d", [(2999, 7), (1001, 30), (17, 1)])
def test_second_tier_shapes(torch_gpu, n, d):
    """Offset clusters push rows to the exact FP64 tier; point counts that are
    not a multiple of the 16-point tile and odd dimensions exercise its tile
    loads and padding."""
    from snapatac2_amd.pp._knn import nearest_neighbour_graph
    X = 1.0 + 1e-3 * clustered(n, d, seed=n + d)
    k = min(50, n - 1)
    g, st = nearest_neighbour_graph(X, k, precision="fp32", return_stats=True)
    _check(g, X, k)
    assert st["exact_scan_rows"] <= st["fallback_rows"]


@pytest.mark.parametrize("n_clusters,spread", [(10, 3.0), (40, 8.0), (1, 0.0)])
def test_chunk_skipping_is_exact(torch_gpu, n_clusters, spread):
    """The bf16 path skips database chunks whose bounding sphere lies beyond every
    threshold of a workgroup: the result equals the FP64 oracle and the run that
    visits every chunk; chunks of well-separated clusters are skipped."""
    from snapatac2_amd.pp._knn import nearest_neighbour_graph
    rng = np.random.default_rng(n_clusters)
    n, d, k = 20000, 30, 50
    centres = spread * rng.standard_normal((n_clusters, d))
    X = centres[rng.integers(0, n_clusters, n)] + rng.standard_normal((n, d))
    g, st = nearest_neighbour_graph(X, k, return_stats=True)
    g_all, st_all = nearest_neighbour_graph(X, k, return_stats=True, prune=False)
    _check(g, X, k)
    assert np.array_equal(g.indices, g_all.indices) and np.array_equal(g.data, g_all.data)
    assert st_all["chunk_visits"] == st_all["chunk_pairs"]
    assert st["chunk_visits"] <= st["chunk_pairs"]
    if n_clusters == 40:
        assert st["chunk_visits"] < 0.7 * st["chunk_pairs"], st


# --- beyond d = 30 and k = 100: the exact all-pairs path ------------------------

@pytest.mark.parametrize("precision", ["bf16", "fp32"])
@pytest.mark.parametrize("n,d,k", [(4000, 50, 50), (3000, 20, 150), (2500, 64, 150), (300, 31, 101),
                                   (200, 1, 199), (20_000, 31, 101)])
def test_wide_rows_and_many_neighbours(torch_gpu, precision, n, d, k):
    """knn.rs:9-32 has no limit on d or k: n_comps = 50 or n_neighbors = 150 give
    the FP64 brute-force sets and bit-identical distances (20,000 points take
    several query blocks of the scan)."""
    from snapatac2_amd.pp._knn import nearest_neighbour_graph
    X = clustered(n, d, seed=n + d + k)
    g, st = nearest_neighbour_graph(X, k, precision=precision, return_stats=True)
    _check(g, X, k)
    assert st["exact_scan_rows"] == n


def test_wide_ties_and_duplicates(torch_gpu):
    """Integer lattice in 40 dimensions plus repeated rows: exact ties are ordered
    by neighbour index, as in the oracle."""
    from snapatac2_amd.pp._knn import nearest_neighbour_graph
    rng = np.random.default_rng(2)
    X = rng.integers(0, 2, (1200, 40)).astype(np.float64)
    X[600:700] = X[0]
    _check(nearest_neighbour_graph(X, 120), X, 120)


def test_knn_dropin_with_50_components(torch_gpu):
    import snapatac2_amd as snap
    from snapatac2_amd.anndata import AnnData
    X = clustered(3000, 50, seed=4)
    ad = AnnData(n_obs=3000, n_vars=0)
    ad.obsm["X_spectral"] = X
    snap.pp.knn(ad, n_neighbors=150)
    _check(ad.obsp["distances"], X, 150)
    g = snap.pp.knn(ad, n_neighbors=30, use_dims=45, inplace=False)
    _check(g, np.ascontiguousarray(X[:, :45]), 30)
