"""Pins the path-B oracle (oracle/spectral.py) without a GPU."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import spectral as osp


def small_counts(n=300, m=800, density=0.05, seed=0):
    rng = np.random.default_rng(seed)
    X = sp.random(n, m, density=density, random_state=rng, format="csr",
                  data_rvs=lambda k: rng.integers(1, 4, k)).astype(np.uint32)
    # make sure every row shares features with others
    X = X + sp.csr_matrix((np.ones(n, dtype=np.uint32), (np.arange(n), rng.integers(0, 20, n))),
                          shape=(n, m))
    return sp.csr_matrix(X, dtype=np.uint32)


def test_idf_special_cases():
    # df: col0 in all 4 rows (-> n-1), col1 in 1 row, col2 in none (-> 1), col3 in 2
    X = sp.csr_matrix(np.array([[1, 1, 0, 1], [2, 0, 0, 0], [1, 0, 0, 3], [5, 0, 0, 0]]))
    w = osp.idf(X)
    np.testing.assert_allclose(w, [np.log(4 / 3), np.log(4), np.log(4), np.log(2)])
    # all-equal document frequencies -> weights 1 (embedding.rs:273-274)
    Y = sp.csr_matrix(np.array([[1, 2], [3, 4], [5, 6.0]]))
    assert np.array_equal(osp.idf(Y), [1.0, 1.0])


def test_hand_derived_operator_entries():
    # A_ij = (S - I)_ij / sqrt(d_i d_j), S = cosine similarity with weights w^2
    # (values multiplied by w, then rows L2-normalised: embedding.rs:314-326)
    X = np.array([[1.0, 2.0, 0.0], [0.0, 1.0, 1.0], [3.0, 0.0, 1.0], [1.0, 1.0, 1.0]])
    w = np.array([0.5, 2.0, 1.5])
    n = X.shape[0]
    S = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            num = sum(X[i, k] * X[j, k] * w[k] * w[k] for k in range(3))
            ni = np.sqrt(sum((X[i, k] * w[k]) ** 2 for k in range(3)))
            nj = np.sqrt(sum((X[j, k] * w[k]) ** 2 for k in range(3)))
            S[i, j] = num / (ni * nj)
    d = S.sum(axis=1) - 1.0
    A = (S - np.eye(n)) / np.sqrt(np.outer(d, d))
    np.testing.assert_allclose(osp.dense_operator(sp.csr_matrix(X), feature_weights=w), A,
                               rtol=1e-13, atol=1e-15)


def test_known_eigenpair():
    X = small_counts()
    Xh, dinv, w, Xn = osp.operator_parts(X)
    d = 1.0 / dinv
    u = np.sqrt(d)
    Au = Xh @ (u @ Xh) - dinv * u
    np.testing.assert_allclose(Au, u, rtol=1e-10)
    evals, evecs = osp.spectral_embedding(X, None, 5)
    assert abs(evals[0] - 1.0) < 1e-10
    assert osp.subspace_angle(evecs[:, :1], u[:, None]) < 1e-6


def test_eigsh_matches_dense_eigh():
    X = small_counts(seed=2)
    A = osp.dense_operator(X)
    lam, U = np.linalg.eigh(A)
    k = 12
    order = np.argsort(-np.abs(lam))[:k]
    evals, evecs = osp.spectral_embedding(X, None, k)
    np.testing.assert_allclose(np.sort(evals), np.sort(lam[order]), rtol=1e-9, atol=1e-12)
    assert osp.subspace_angle(evecs, U[:, order]) < 1e-6


def test_features_mask_and_determinism():
    X = small_counts(seed=3)
    mask = np.zeros(X.shape[1], dtype=bool)
    mask[::2] = True
    a = osp.spectral(X, 8, features=mask)
    b = osp.spectral(X, 8, features=mask)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    c = osp.spectral(X[:, mask], 8)
    np.testing.assert_allclose(a[0], c[0], rtol=1e-12)


def test_orthogonalize_oracle_invariants():
    """orthogonalize (_embedding.py:397-413) returns orthonormal vectors that keep
    the low-rank matrix E diag(lam) E^T: evecs_new = U_svd W with W orthogonal,
    and U_svd W diag(lam_new) W^T U_svd^T = U_svd S V^T diag(lam) V S U_svd^T."""
    rng = np.random.default_rng(5)
    E = rng.standard_normal((400, 6)) @ rng.standard_normal((6, 6))
    lam = np.sort(rng.random(6))[::-1]
    ev, U = osp.orthogonalize(lam, E)
    ev, U = ev.real, U.real
    np.testing.assert_allclose(U.T @ U, np.eye(6), atol=1e-10)
    np.testing.assert_allclose((U * ev) @ U.T, (E * lam) @ E.T, atol=1e-9 * np.abs(E).max() ** 2)
    assert np.all(np.diff(ev) <= 0)


def _nystrom_fixture():
    from _clusters import clustered_counts
    X = clustered_counts(3000, 1500, 4, seed=3000)
    rows = np.sort(np.random.default_rng(7).choice(X.shape[0], 800, replace=False))
    return X, rows


def test_nystrom_oracle_all_landmarks_closed_form():
    """With every cell a landmark, nystrom's product is fixed by the eigen-equation:
    A U = U L with A = D^-1/2 (S - I) D^-1/2 gives
    S (D^-1/2 U L^-1) = D^1/2 U + D^-1/2 U L^-1, and nystrom (embedding.rs:193-229)
    computes q = S (U_scaled) with U_scaled = D^-1/2 U L^-1 (degrees = 1/dinv,
    embedding.rs:189-190, rows over sqrt(degree), :205-209), then rescales each row.
    So each output row is parallel to that closed form."""
    from _clusters import clustered_counts
    X = clustered_counts(600, 400, 3, seed=11)
    rows = np.arange(X.shape[0])
    v, q = osp.nystrom(X, rows, 5, chunk_size=250)
    Xh, dinv, _, _ = osp.operator_parts(X)
    ev, U = osp.eigen(Xh, dinv, 5, 0)
    np.testing.assert_array_equal(ev, v)
    d = 1.0 / dinv
    ref = np.sqrt(d)[:, None] * U + (U / np.sqrt(d)[:, None]) / ev[None, :]
    cos = np.sum(q * ref, axis=1) / (np.linalg.norm(q, axis=1) * np.linalg.norm(ref, axis=1))
    assert np.min(cos) > 1.0 - 1e-10


def test_nystrom_oracle_near_exact_embedding():
    """800 of 3,000 clustered cells as landmarks, then orthogonalize, as the
    reference's spectral() does after spectral_embedding_nystrom
    (_embedding.py:257-265): the top-3 subspace lies within 0.05 rad of the exact
    embedding (osp.spectral; measured 0.016). Inverting the landmark degree scaling
    (rows times sqrt(degree) instead of over it) lands 0.32 rad away, so this pin
    catches that mistake (next test)."""
    X, rows = _nystrom_fixture()
    _, eu = osp.spectral(X, n_comps=6)
    v, q = osp.nystrom(X, rows, 6)
    _, oq = osp.orthogonalize(v, q)
    assert osp.subspace_angle(oq.real[:, :3], eu[:, :3]) < 0.05


@pytest.mark.parametrize("n,m,seed,chunk", [(40, 60, 5, 16), (30, 80, 1, 30), (50, 40, 2, 7)])
def test_nystrom_oracle_degree_closed_form(n, m, seed, chunk):
    """nystrom's degree step (embedding.rs:222-229: t = colsum(q) * v over the
    chunk's rows, d = q t, q /= sqrt(d)) pinned element-wise. With every cell a
    landmark, M = S S^T (S the normalized rows) and A = D^-1/2 (M - I) D^-1/2,
    the product of pair j is q_j = D^1/2 u_j + D^-1/2 u_j / l_j (eigen-equation),
    and the sum of l_j q_j q_j^T over all n pairs is M + I + (M - I)^-1. With
    the n - 2 pairs of largest |l| that eigsh keeps, d_i is therefore the sum,
    over the columns of row i's chunk, of M + I + (M - I)^-1 minus the two
    dropped pairs' terms (taken from numpy's dense eigh). Dropping the
    eigenvalues from t, or summing q over all rows instead of the chunk, moves d
    by more than 1e-3 (checked below), against the 1e-10 the oracle meets."""
    from _clusters import clustered_counts
    X = clustered_counts(n, m, 3, seed=seed)
    k = n - 2
    v, q = osp.nystrom(X, np.arange(n), k, chunk_size=chunk)
    _, dinv, _, Xn = osp.operator_parts(X)
    lam, E = np.linalg.eigh(osp.dense_operator(X))
    order = np.argsort(-np.abs(lam), kind="stable")
    lam, E = lam[order], E[:, order]
    sd = np.sqrt(1.0 / dinv)
    M = (Xn @ Xn.T).toarray()
    W = M + np.eye(n) + np.linalg.inv(M - np.eye(n))
    for j in range(k, n):
        qj = sd * E[:, j] + E[:, j] / sd / lam[j]
        W -= lam[j] * np.outer(qj, qj)
    same = (np.arange(n) // chunk)[:, None] == (np.arange(n) // chunk)[None, :]
    d = (W * same).sum(axis=1)
    assert d.min() > 0.5  # the d <= 0 clamp stays out of play
    desc = np.argsort(-lam[:k], kind="stable")  # eigen() returns descending eigenvalues
    lam_k, E_k = lam[:k][desc], E[:, :k][:, desc]
    np.testing.assert_allclose(v, lam_k, rtol=0, atol=1e-12)
    qc = sd[:, None] * E_k + (E_k / sd[:, None]) / lam_k[None, :]
    sign = np.sign(np.sum(q * qc, axis=0))
    np.testing.assert_allclose(q, sign * qc / np.sqrt(d)[:, None], rtol=0, atol=1e-10 * np.abs(q).max())
    # the check separates the plausible mistakes
    no_eigenvalues = ((qc @ qc.T) * same).sum(axis=1)
    all_rows = qc @ (qc.sum(axis=0) * lam_k)
    assert np.abs(no_eigenvalues / d - 1).max() > 1e-3
    if chunk < n:
        assert np.abs(all_rows / d - 1).max() > 1e-3


def test_nystrom_oracle_pin_rejects_inverted_scaling(monkeypatch):
    X, rows = _nystrom_fixture()
    _, eu = osp.spectral(X, n_comps=6)
    monkeypatch.setattr(osp, "_landmark_scale",
                        lambda u, v, dinv: u / np.sqrt(dinv).reshape((-1, 1)) / v.reshape((1, -1)))
    v, q = osp.nystrom(X, rows, 6)
    _, oq = osp.orthogonalize(v, q)
    assert osp.subspace_angle(oq.real[:, :3], eu[:, :3]) > 0.2


def test_subspace_sin_known_rotation():
    """sin of the largest principal angle: a basis rotated by theta in one
    direction gives sin(theta), also at 1e-9 rad where arccos of the cosine
    loses its digits."""
    rng = np.random.default_rng(0)
    Q, _ = np.linalg.qr(rng.standard_normal((200, 6)))
    for th in (1e-9, 1e-6, 1e-3, 0.3):
        V = Q[:, :5].copy()
        V[:, 2] = np.cos(th) * Q[:, 2] + np.sin(th) * Q[:, 5]
        assert abs(osp.subspace_sin(Q[:, :5], V @ rng.standard_normal((5, 5))) - np.sin(th)) < 1e-12 + 1e-9 * th


def test_eigen_device_equals_scipy_restatement():
    """eigen_device (the reference's eigsh call on the torch FP64 operator)
    gives the eigenpairs of the scipy restatement on the same counts."""
    import torch
    from oracle import adjoint as oadj
    X = small_counts(seed=4)
    T = oadj.ChunkedCounts(torch.from_numpy(X.indptr.astype(np.int64)), torch.from_numpy(X.indices.astype(np.int32)),
                           torch.from_numpy(X.data.view(np.int32)), X.shape, chunk_entries=500)
    w = torch.from_numpy(osp.idf(X))
    ev, U, nmv = osp.eigen_device(T, w, 8, seed=0)
    ev2, U2 = osp.spectral_embedding(X, None, 8, 0)
    np.testing.assert_allclose(ev, ev2, rtol=1e-10, atol=1e-12)
    assert osp.subspace_sin(U, U2) < 1e-8
    assert nmv > 8
