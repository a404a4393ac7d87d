"""Path D on the GPU: the hand-written adjoint of the spectral operator.

dw from snapgpu_spectral_vjp_w (FP32 sparse products, FP64 reductions) against
the FP64 oracle: the closed form and torch.sparse autograd within 1e-5 relative
(2-norm), central finite differences (h = 1e-3) within 1e-4 relative; dV = A G
within the operator's 1e-5. Scale invariance <dw, w> = 0 at the larger size.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import adjoint as adj
from oracle import spectral as osp

pytestmark = pytest.mark.gpu


def _counts(n, m, density, seed, binary=False):
    rng = np.random.default_rng(seed)
    X = sp.random(n, m, density=density, random_state=rng, format="csr",
                  data_rvs=(lambda k: np.ones(k)) if binary else (lambda k: rng.integers(1, 4, k)))
    X = X + sp.csr_matrix((np.ones(n), (np.arange(n), rng.integers(0, 20, n))), shape=(n, m))
    X = sp.csr_matrix(X)
    if binary:
        X.data[:] = 1
    return sp.csr_matrix(X, dtype=np.uint32)


def _op(torch, X, w):
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    return SpectralOperator(_to_device_csr(X, torch), w)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _weights(X, seed):
    return osp.idf(sp.csr_matrix(X, dtype=np.float64)) * np.random.default_rng(seed).uniform(0.5, 1.5, X.shape[1])


@pytest.mark.parametrize("b", [1, 5, 16, 30])
def test_vjp_matches_oracle(torch_gpu, b):
    torch = torch_gpu
    X = _counts(400, 300, 0.05, seed=b)
    w = _weights(X, b)
    rng = np.random.default_rng(100 + b)
    V = rng.standard_normal((400, b))
    G = rng.standard_normal((400, b))
    op = _op(torch, X, w)
    dw = op.vjp_w(torch.from_numpy(V).cuda(), torch.from_numpy(G).cuda()).cpu().numpy()
    cf = adj.closed_form_grad(X, w, V, G)
    _, dV_ref, dw_ref = adj.torch_sparse_grad(X, w, V, G)
    assert _rel(dw, cf) < 1e-5
    assert _rel(dw, dw_ref) < 1e-5
    dV = op.matmat(torch.from_numpy(G).cuda()).cpu().numpy()
    assert np.abs(dV - dV_ref).max() < 1e-5 * np.abs(dV_ref).max()


@pytest.mark.parametrize("b", [1, 2, 3, 5, 16, 30, 40])
def test_apply_f64_matches_oracle(torch_gpu, b):
    """The operator with both sparse products in FP64 (the eigengap acceptance
    check): against the FP64 oracle to summation rounding, for every column
    width of the product kernel (1, 2, 4, 8, 16 and 32-column lane groups)."""
    torch = torch_gpu
    X = _counts(500, 350, 0.04, seed=20 + b)
    w = _weights(X, b)
    V = np.random.default_rng(200 + b).standard_normal((500, b))
    op = _op(torch, X, w)
    got = op.apply_f64(torch.from_numpy(V).cuda()).cpu().numpy()
    ref = adj.apply(X, w, V)
    assert _rel(got, ref) < 1e-12
    rq = op.rayleigh_f64(torch.from_numpy(V).cuda()).cpu().numpy()
    assert np.allclose(rq, (V * ref).sum(0), rtol=1e-11, atol=1e-12 * np.abs(ref).max())
    if b <= 32:  # V^T A V from the one-product form
        H = op.gram_f64(torch.from_numpy(V).cuda()).cpu().numpy()
        Href = V.T @ ref
        assert np.abs(H - Href).max() <= 1e-11 * np.abs(Href).max()


@pytest.mark.parametrize("b", [2, 3])
def test_gram_f64_panel_layout(torch_gpu, b):
    """V^T A V over the CSC panel layout (several panels of cells, FP64 sums)
    against the FP64 oracle, and against the CSR-form product."""
    torch = torch_gpu
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    X = _counts(40000, 36000, 0.004, seed=31)
    w = _weights(X, 31)
    Xd = _to_device_csr(X, torch)
    op = SpectralOperator(Xd, w, layout="panel")
    op_csr = SpectralOperator(Xd, w, use_panels=False)
    assert op.layout == "panel" and op_csr.layout == "csr"
    V = np.random.default_rng(9).standard_normal((40000, b))
    Vt = torch.from_numpy(V).cuda()
    H = op.gram_f64(Vt).cpu().numpy()
    Href = V.T @ adj.apply(X, w, V)
    assert np.abs(H - Href).max() <= 1e-11 * np.abs(Href).max()
    assert np.abs(H - op_csr.gram_f64(Vt).cpu().numpy()).max() <= 1e-12 * np.abs(Href).max()


@pytest.mark.parametrize("b", [2, 5, 11])
def test_vjp_mixed_layout(torch_gpu, b):
    """The adjoint with its FP32 products over the mixed layout (CSC panels, CSR
    windows; V and G side by side in 16 columns, 8 columns per call) against
    the closed form and the row-parallel adjoint."""
    torch = torch_gpu
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    X = _counts(40000, 36000, 0.004, seed=41)
    w = _weights(X, 41)
    Xd = _to_device_csr(X, torch)
    op = SpectralOperator(Xd, w, layout="mixed")
    op_csr = SpectralOperator(Xd, w, layout="csr")
    assert op.layout == "mixed"
    rng = np.random.default_rng(400 + b)
    V = rng.standard_normal((40000, b))
    G = rng.standard_normal((40000, b))
    Vt, Gt = torch.from_numpy(V).cuda(), torch.from_numpy(G).cuda()
    dw = op.vjp_w(Vt, Gt).cpu().numpy()
    assert _rel(dw, adj.closed_form_grad(X, w, V, G)) < 1e-5
    assert _rel(dw, op_csr.vjp_w(Vt, Gt).cpu().numpy()) < 1e-5
    assert np.array_equal(dw, op.vjp_w(Vt, Gt).cpu().numpy())  # fixed summation order


def test_vjp_finite_differences(torch_gpu):
    torch = torch_gpu
    X = _counts(300, 200, 0.06, seed=7)
    w = _weights(X, 7)
    rng = np.random.default_rng(8)
    V = rng.standard_normal((300, 4))
    G = rng.standard_normal((300, 4))
    op = _op(torch, X, w)
    dw = op.vjp_w(torch.from_numpy(V).cuda(), torch.from_numpy(G).cuda()).cpu().numpy()
    coords = rng.choice(200, 12, replace=False)
    dirs = rng.standard_normal((3, 200))
    fc, fd = adj.fd_grad(X, w, V, G, h=1e-3, coords=coords, directions=dirs)
    scale = np.abs(dw).max()
    assert np.abs(dw[coords] - fc).max() < 1e-4 * scale
    assert np.abs(dirs @ dw - fd).max() < 1e-4 * np.abs(fd).max()


def test_autograd_function(torch_gpu):
    torch = torch_gpu
    from snapatac2_amd.tl import normalized_cosine
    X = _counts(500, 350, 0.04, seed=21)
    w0 = _weights(X, 21)
    rng = np.random.default_rng(22)
    V0 = rng.standard_normal((500, 6))
    G0 = rng.standard_normal((500, 6))
    op = _op(torch, X, None)  # IDF weights; the function re-prepares for w
    V = torch.tensor(V0, device="cuda", requires_grad=True)
    w = torch.tensor(w0, device="cuda", requires_grad=True)
    Y = normalized_cosine(V, w, op)
    (Y * torch.from_numpy(G0).cuda()).sum().backward()
    Yr, dV_ref, dw_ref = adj.torch_sparse_grad(X, w0, V0, G0)
    assert np.abs(Y.detach().cpu().numpy() - Yr).max() < 1e-5 * np.abs(Yr).max()
    assert np.abs(V.grad.cpu().numpy() - dV_ref).max() < 1e-5 * np.abs(dV_ref).max()
    assert _rel(w.grad.cpu().numpy(), dw_ref) < 1e-5
    # FP32 input: output and dV in FP32, dw in FP64
    V32 = torch.tensor(V0, device="cuda", dtype=torch.float32, requires_grad=True)
    Y32 = normalized_cosine(V32, w, op)
    assert Y32.dtype == torch.float32
    Y32.sum().backward()
    assert V32.grad.dtype == torch.float32


def test_float_values_and_wide_blocks(torch_gpu):
    torch = torch_gpu
    X = _counts(600, 400, 0.04, seed=31).astype(np.float32)
    X.data *= np.float32(0.75)
    w = _weights(X, 31)
    rng = np.random.default_rng(32)
    V = rng.standard_normal((600, 40))  # two column chunks (32 + 8)
    G = rng.standard_normal((600, 40))
    op = _op(torch, X, w)
    assert op.value_type == 1
    dw = op.vjp_w(torch.from_numpy(V).cuda(), torch.from_numpy(G).cuda()).cpu().numpy()
    assert _rel(dw, adj.closed_form_grad(X.astype(np.float64), w, V, G)) < 1e-5


def test_vjp_larger_binary(torch_gpu):
    torch = torch_gpu
    X = _counts(20000, 30000, 0.01, seed=41, binary=True)
    w = osp.idf(sp.csr_matrix(X, dtype=np.float64))
    rng = np.random.default_rng(42)
    V = rng.standard_normal((20000, 30))
    G = rng.standard_normal((20000, 30))
    op = _op(torch, X, None)
    dw = op.vjp_w(torch.from_numpy(V).cuda(), torch.from_numpy(G).cuda()).cpu().numpy()
    assert _rel(dw, adj.closed_form_grad(X, w, V, G)) < 1e-5
    assert abs(dw @ w) < 1e-5 * np.linalg.norm(dw) * np.linalg.norm(w)


def test_set_weights_matches_fresh_operator(torch_gpu):
    torch = torch_gpu
    X = _counts(500, 300, 0.05, seed=51)
    w = _weights(X, 51)
    op = _op(torch, X, None)
    op.set_weights(torch.from_numpy(w).cuda())
    fresh = _op(torch, X, w)
    for name in ("s", "dinv", "r32", "w2"):
        assert torch.equal(getattr(op, name), getattr(fresh, name)), name


def test_eigengap_ascent(torch_gpu):
    torch = torch_gpu
    from snapatac2_amd.tl import tune_feature_weights, spectral_gap_grad
    X = _counts(800, 500, 0.03, seed=61)
    op = _op(torch, X, None)
    k = 6
    gap0, g, evals, _ = spectral_gap_grad(op, k)
    # Hellmann-Feynman gradient against the oracle's closed form on the eigenvectors
    lam, U = np.linalg.eigh(osp.dense_operator(X))
    order = np.lexsort((-lam, -np.abs(lam)))
    u = U[:, order[k - 1:k + 1]]
    sg = np.array([np.sign(lam[order[k - 1]]), -np.sign(lam[order[k]])])
    cf = adj.closed_form_grad(X, op.w.cpu().numpy(), u, u * sg)
    assert _rel(g.cpu().numpy(), cf) < 1e-4
    w, hist, ev, U = tune_feature_weights(op, k, steps=6)
    assert hist[-1][0] > gap0
    # the history holds FP64 Rayleigh-quotient gaps of the Lanczos vectors
    assert abs((abs(ev[k - 1]) - abs(ev[k])) - hist[-1][0]) < 1e-6 * abs(ev[0])
    assert torch.equal(op.w, w)
    assert all(h1[0] > h0[0] for h0, h1 in zip(hist, hist[1:]))
    assert (w >= 0).all()


def test_ascent_tolerances(torch_gpu, monkeypatch):
    """Eigensolves that only feed a gradient stop at grad_tol (1e-7); the
    eigenpairs returned come from a solve at tol (1e-9): the last trial's when
    it is kept, else a solve at tol started from the kept pairs. Both agree with
    a fresh solve at tol (to the FP32 products' rounding)."""
    torch = torch_gpu
    from snapatac2_amd.tl import _adjoint as A
    X = _counts(3000, 1500, 0.02, seed=71)  # above 1024 cells: block Lanczos
    calls = []
    real = A.eigsh_lm

    def spy(op, k, **kw):
        calls.append((kw.get("tol"), kw.get("v0") is not None))
        return real(op, k, **kw)

    monkeypatch.setattr(A, "eigsh_lm", spy)
    k = 6
    for step_size, kept in ((0.01, True), (1e-13, False)):
        op = _op(torch, X, None)
        w0 = op.w.clone()
        calls.clear()
        w, hist, ev, U = A.tune_feature_weights(op, k, steps=1, step_size=step_size, min_step=0.0)
        assert (len(hist) == 2) == kept
        assert [t for t, _ in calls] == ([1e-7, 1e-9] if kept else [1e-7, 1e-9, 1e-9])
        assert calls[-1][1]  # the last solve starts from earlier eigenvectors
        if not kept:
            assert torch.equal(w, w0) and torch.equal(op.w, w0)
        ev2, _, _ = real(op, k + 1, seed=1, tol=1e-9)
        np.testing.assert_allclose(ev, ev2, rtol=0, atol=1e-6 * abs(ev2[0]))


def test_vjp_generator_200k_binary(torch_gpu):
    """200,000 generator cells, 500 bp tiles, top 500,000 features, binarised;
    V (FP32 values) and G 200,000 x 30. References in FP64 through torch's own
    sparse products on the device: the closed form (pinned on the CPU against
    torch autograd and finite differences) and a central difference (h = 1e-3)
    of the FP64 forward along one direction."""
    torch = torch_gpu
    from snapatac2_amd import synth
    from snapatac2_amd.genome import synthetic_genome
    from snapatac2_amd.pp import make_tile_matrix_chunked
    from snapatac2_amd.pp._basic import _column_sums, _find_most_accessible_features
    from snapatac2_amd.tl._operator import SpectralOperator, select_columns
    n = 200_000
    chunks = list(synth.iter_chunks(n, seed=3))
    X, _ = make_tile_matrix_chunked(chunks, synthetic_genome(), 500, ["chrM", "chrY"])
    del chunks
    sel = np.sort(_find_most_accessible_features(_column_sums(X), 0.005, 0.005, 500_000))
    Xs = select_columns(X, sel)
    del X
    Xs.data.fill_(1)
    op = SpectralOperator(Xs)
    print(f"n={n} m={op.m} nnz={Xs.nnz} density={Xs.nnz / (n * op.m):.4f}", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(5)
    V = torch.randn(n, 30, dtype=torch.float32, device="cuda", generator=gen).double()
    G = torch.randn(n, 30, dtype=torch.float64, device="cuda", generator=gen)
    dw = op.vjp_w(V, G)
    dV = op.matmat(G)
    rows = torch.repeat_interleave(torch.arange(n, device="cuda"), torch.diff(Xs.indptr))
    T = adj.TorchCounts(rows, Xs.indices.long(), torch.ones(Xs.nnz, dtype=torch.float64, device="cuda"),
                        (n, op.m))
    del rows
    w = op.w.clone()
    dw_ref = adj.closed_form_grad_tensors(T, w, V, G)
    dV_ref = adj.apply_tensors(T, w, G)
    rel_w = (torch.linalg.norm(dw - dw_ref) / torch.linalg.norm(dw_ref)).item()
    rel_v = ((dV - dV_ref).abs().max() / dV_ref.abs().max()).item()
    print(f"dw rel {rel_w:.3e}  dV rel {rel_v:.3e}", flush=True)
    assert rel_w < 1e-5
    assert rel_v < 1e-5
    p = torch.randn(op.m, dtype=torch.float64, device="cuda", generator=gen) * w
    h = 1e-3
    fd = ((adj.apply_tensors(T, w + h * p, V) * G).sum() -
          (adj.apply_tensors(T, w - h * p, V) * G).sum()).item() / (2 * h)
    an = (dw @ p).item()
    print(f"fd {fd:.9e} adjoint {an:.9e}", flush=True)
    assert abs(an - fd) < 1e-4 * abs(fd)
    assert abs((dw @ w).item()) < 1e-5 * torch.linalg.norm(dw).item() * torch.linalg.norm(w).item()


def test_eigengap_gradient_near_degenerate_pair(torch_gpu):
    """A pair in the bulk of the spectrum 7e-6 apart (|lambda_8| - |lambda_9|
    relative 7e-4, as lambda_30 - lambda_31 at a million cells): the
    Hellmann-Feynman gradient of the gap (after the FP64 Rayleigh-Ritz of the
    pair) matches central differences of the FP64 dense eigenvalues along random
    directions, and the gap itself matches them."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(__file__))
    from _clusters import clustered_counts
    from snapatac2_amd.tl import spectral_gap_grad
    torch = torch_gpu
    X = sp.csr_matrix(clustered_counts(600, 400, clusters=6, seed=1))
    op = _op(torch, X, None)
    k = 8
    gap, g, evals, _ = spectral_gap_grad(op, k)
    w = op.w.cpu().numpy()

    def dense_gap(wv):
        lam = np.linalg.eigvalsh(osp.dense_operator(X, feature_weights=wv))
        lam = lam[np.argsort(-np.abs(lam), kind="stable")]
        return abs(lam[k - 1]) - abs(lam[k])

    g0 = dense_gap(w)
    assert 1e-6 < g0 < 2e-5
    assert abs(gap - g0) < 1e-9
    gh = g.cpu().numpy()
    rng = np.random.default_rng(3)
    h = 1e-6
    for _ in range(3):
        d = rng.standard_normal(w.size) * w
        fd = (dense_gap(w + h * d) - dense_gap(w - h * d)) / (2 * h)
        assert abs(gh @ d - fd) <= 1e-3 * abs(fd) + 1e-9, (gh @ d, fd)
