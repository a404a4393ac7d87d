"""Pins the path-D oracle (oracle/adjoint.py) on the CPU.

Three independent FP64 derivations of d<G, A(w) V>/dw must agree (finite
differences, torch autograd through torch.sparse, the closed form), and the
mathematics fixes two more checks: A(c w) = A(w) so <dw, w> = 0, and
Hellmann-Feynman d lambda = u^T dA u against differences of dense eigenvalues.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import adjoint as adj
from oracle import spectral as osp


def _problem(n=60, m=45, b=3, seed=0, empty_row=True):
    rng = np.random.default_rng(seed)
    X = sp.random(n, m, density=0.25, random_state=seed, format="csr")
    X.data = np.ceil(X.data * 3.0)
    X = X.tolil()
    if empty_row:
        X[7, :] = 0
    X = sp.csr_matrix(X)
    X.eliminate_zeros()
    w = osp.idf(X) * rng.uniform(0.5, 1.5, m)
    V = rng.standard_normal((n, b))
    G = rng.standard_normal((n, b))
    return X, w, V, G


def test_forward_matches_spectral_oracle():
    X, w, V, G = _problem()
    A = osp.dense_operator(X, feature_weights=w)
    np.testing.assert_allclose(adj.apply(X, w, V), A @ V, rtol=1e-12, atol=1e-12)
    Y, _, _ = adj.torch_sparse_grad(X, w, V, G)
    np.testing.assert_allclose(Y, A @ V, rtol=1e-11, atol=1e-12)


def test_gradients_agree():
    X, w, V, G = _problem()
    cf = adj.closed_form_grad(X, w, V, G)
    _, dV, dw = adj.torch_sparse_grad(X, w, V, G)
    assert np.linalg.norm(cf - dw) <= 1e-11 * np.linalg.norm(dw)
    fd, _ = adj.fd_grad(X, w, V, G, h=1e-3, coords=range(X.shape[1]))
    assert np.linalg.norm(fd - cf) <= 1e-5 * np.linalg.norm(cf)
    # dV = A G (A symmetric)
    np.testing.assert_allclose(dV, adj.apply(X, w, G), rtol=1e-11, atol=1e-11)


def test_fd_along_directions():
    X, w, V, G = _problem(n=80, m=50, b=4, seed=3)
    rng = np.random.default_rng(9)
    dirs = rng.standard_normal((4, X.shape[1]))
    _, dd = adj.fd_grad(X, w, V, G, h=1e-3, directions=dirs)
    cf = adj.closed_form_grad(X, w, V, G)
    # central differences at h = 1e-3 carry an O(h^2) truncation error
    assert np.max(np.abs(dirs @ cf - dd)) <= 1e-5 * np.max(np.abs(dd))


def test_scale_invariance():
    X, w, V, G = _problem(seed=5)
    cf = adj.closed_form_grad(X, w, V, G)
    assert abs(cf @ w) <= 1e-10 * np.linalg.norm(cf) * np.linalg.norm(w)
    # and the forward is indeed scale invariant
    np.testing.assert_allclose(adj.apply(X, 3.0 * w, V), adj.apply(X, w, V), rtol=1e-12, atol=1e-12)


def test_hellmann_feynman():
    X, w, _, _ = _problem(n=50, m=40, seed=11, empty_row=False)
    lam, U = np.linalg.eigh(osp.dense_operator(X, feature_weights=w))
    i = int(np.argmax(np.minimum(np.abs(np.diff(lam))[:-1], np.abs(np.diff(lam))[1:]))) + 1
    u = U[:, i:i + 1]
    g = adj.closed_form_grad(X, w, u, u)
    p = np.random.default_rng(2).standard_normal(X.shape[1])
    h = 1e-4
    lp = np.linalg.eigvalsh(osp.dense_operator(X, feature_weights=w + h * p))[i]
    lm = np.linalg.eigvalsh(osp.dense_operator(X, feature_weights=w - h * p))[i]
    assert g @ p == pytest.approx((lp - lm) / (2 * h), rel=1e-5)


def test_torch_tensor_forms():
    import torch
    X, w, V, G = _problem(seed=13)
    coo = X.tocoo()
    T = adj.TorchCounts(torch.tensor(coo.row, dtype=torch.long), torch.tensor(coo.col, dtype=torch.long),
                        torch.tensor(coo.data, dtype=torch.float64), X.shape)
    wt, Vt, Gt = (torch.tensor(np.asarray(a, dtype=np.float64)) for a in (w, V, G))
    np.testing.assert_allclose(adj.apply_tensors(T, wt, Vt).numpy(), adj.apply(X, w, V),
                               rtol=1e-11, atol=1e-12)
    cf = adj.closed_form_grad(X, w, V, G)
    np.testing.assert_allclose(adj.closed_form_grad_tensors(T, wt, Vt, Gt).numpy(), cf,
                               rtol=1e-10, atol=1e-12 * np.abs(cf).max())


def test_chunked_counts_forward_and_idf():
    """ChunkedCounts (the FP64 checker of the 1M-cell tests) over chunks of 37
    entries equals the forward, the column sums and document frequencies equal
    scipy's, and idf_from_df equals oracle.spectral.idf (clamps, all-equal rule)."""
    import torch
    X, w, V, G = _problem(seed=21)
    X = sp.csr_matrix(X, dtype=np.uint32)
    T = adj.ChunkedCounts(torch.from_numpy(X.indptr.astype(np.int64)), torch.from_numpy(X.indices.astype(np.int32)),
                          torch.from_numpy(X.data.view(np.int32)), X.shape, chunk_entries=37)
    assert len(T.bounds) > 10
    wt, Vt = torch.tensor(w), torch.tensor(V)
    np.testing.assert_allclose(T.apply(wt, Vt).numpy(), adj.apply(X.astype(np.float64), w, V),
                               rtol=1e-11, atol=1e-12)
    np.testing.assert_array_equal(T.col_sums().numpy(), np.ravel(X.sum(axis=0)).astype(np.float64))
    np.testing.assert_array_equal(T.df().numpy(), np.diff(X.tocsc().indptr))
    np.testing.assert_allclose(adj.idf_from_df(T.df(), X.shape[0]).numpy(), osp.idf(X), rtol=1e-15)
    n = 10
    full = torch.full((5,), n)  # every column in every row: all equal -> 1
    assert torch.equal(adj.idf_from_df(full, n), torch.ones(5, dtype=torch.float64))
    got = adj.idf_from_df(torch.tensor([0, n, 2]), n).numpy()
    np.testing.assert_allclose(got, [np.log(n / 1.0), np.log(n / (n - 1.0)), np.log(n / 2.0)], rtol=1e-15)
