"""Path B on the GPU: operator, preparation and eigenpairs vs the oracle.

Tolerances: eigenvalues 1e-5 relative (atol 1e-7), largest principal angle of
the eigenvector subspace < 1e-4 rad; operator application (FP32 sparse
products) 1e-5 relative in the max norm; preparation (FP64) 1e-12 relative.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import spectral as osp

pytestmark = pytest.mark.gpu


def small_counts(n=300, m=800, density=0.05, seed=0):
    rng = np.random.default_rng(seed)
    X = sp.random(n, m, density=density, random_state=rng, format="csr",
                  data_rvs=lambda k: rng.integers(1, 4, k)).astype(np.uint32)
    X = X + sp.csr_matrix((np.ones(n, dtype=np.uint32), (np.arange(n), rng.integers(0, 20, n))),
                          shape=(n, m))
    return sp.csr_matrix(X, dtype=np.uint32)


def _op(torch, X, weights=None):
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    return SpectralOperator(_to_device_csr(X, torch), weights)


def test_transpose_and_prepare(torch_gpu):
    torch = torch_gpu
    X = small_counts(2000, 3000, 0.02, seed=1)
    op = _op(torch, X)
    Xc = X.tocsc()
    Xc.sort_indices()
    assert np.array_equal(op.colptr.cpu().numpy(), Xc.indptr)
    assert np.array_equal(op.rowidx.cpu().numpy(), Xc.indices)
    assert np.array_equal(op.cval.cpu().numpy().view(np.uint32), Xc.data)
    Xh, dinv, w, Xn = osp.operator_parts(X)
    np.testing.assert_allclose(op.w.cpu().numpy(), w, rtol=1e-14)
    np.testing.assert_allclose(op.dinv.cpu().numpy(), dinv, rtol=1e-12)


def test_operator_application(torch_gpu):
    torch = torch_gpu
    X = small_counts(3000, 5000, 0.02, seed=4)
    op = _op(torch, X)
    Xh, dinv, _, _ = osp.operator_parts(X)
    rng = np.random.default_rng(0)
    for b in (32, 30, 7):
        V = rng.standard_normal((3000, b))
        Y = op.matmat(torch.from_numpy(V).cuda()).cpu().numpy()
        ref = Xh @ (Xh.T @ V) - dinv[:, None] * V
        err = np.abs(Y - ref).max() / np.abs(ref).max()
        assert err < 1e-5, (b, err)


def _compare(evals, evecs, oe, ov, k):
    np.testing.assert_allclose(evals[:k], oe[:k], rtol=1e-5, atol=1e-7)
    ang = osp.subspace_angle(evecs[:, :k], ov[:, :k])
    assert ang < 1e-4, ang


@pytest.mark.parametrize("n", [300, 2500])
def test_spectral_embedding_vs_oracle(torch_gpu, n):
    from snapatac2_amd.tl._embedding import spectral_embedding
    import snapatac2_amd as snap
    X = small_counts(n, 4000, 0.03, seed=n)
    ad = snap.AnnData(X=X)
    evals, evecs, st = spectral_embedding(ad, None, 20, 0, return_stats=True)
    oe, ov = osp.spectral_embedding(X, None, 20, 0)
    _compare(evals, evecs, oe, ov, 20)
    print("stats", st)


def test_synthetic_2k_cells_pipeline(torch_gpu):
    import snapatac2_amd as snap
    from snapatac2_amd import synth
    ad = synth.synthetic_anndata(2000, seed=0)
    snap.pp.add_tile_matrix(ad)
    snap.pp.select_features(ad, n_features=50000)
    from snapatac2_amd.tl._embedding import spectral_embedding
    evals, evecs, st = spectral_embedding(ad, ad.var["selected"].to_numpy(), 30, 0,
                                          return_stats=True)
    X = ad.X.to_scipy()
    oe, ov = osp.spectral_embedding(X, ad.var["selected"].to_numpy(), 30, 0)
    print("stats", st, "evals", evals[:5], evals[-3:], "oracle", oe[-3:])
    _compare(evals, evecs, oe, ov, 30)
    # drop-in: weighted_by_sd keeps positive eigenvalues, scales by sqrt
    snap.tl.spectral(ad, n_comps=30)
    E = ad.obsm["X_spectral"]
    lam = ad.uns["spectral_eigenvalue"]
    assert E.shape == (2000, lam.size) and np.all(lam > 0)
    np.testing.assert_allclose(np.linalg.norm(E, axis=0) ** 2, lam, rtol=1e-8)


def test_determinism_and_dense_float_input(torch_gpu):
    # reference test_reproducibility (test_tools.py:95-122): 50 x 100 dense floats in [1, 100]
    import snapatac2_amd as snap
    rng = np.random.default_rng(0)
    M = rng.uniform(1, 100, (50, 100))
    ad = snap.AnnData(X=sp.csr_matrix(M))
    outs = [snap.tl.spectral(ad, features=None, random_state=0, inplace=False)[1] for _ in range(3)]
    for o in outs:
        assert np.array_equal(o, outs[0])
    oe, ov = osp.spectral(sp.csr_matrix(M), 30, None, 0)
    ge, gv = snap.tl.spectral(ad, features=None, random_state=0, inplace=False)
    np.testing.assert_allclose(ge, oe, rtol=1e-5, atol=1e-7)
    snap.tl.spectral(ad, features=None, random_state=0)
    knns = [snap.pp.knn(ad, n_neighbors=25, inplace=False).toarray() for _ in range(3)]
    for k in knns:
        assert np.array_equal(k, knns[0])


def test_block_lanczos_determinism(torch_gpu):
    from snapatac2_amd.tl._embedding import spectral_embedding
    import snapatac2_amd as snap
    X = small_counts(3000, 4000, 0.03, seed=9)
    ad = snap.AnnData(X=X)
    a = spectral_embedding(ad, None, 30, 0)
    b = spectral_embedding(ad, None, 30, 0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_errors(torch_gpu):
    import snapatac2_amd as snap
    ad = snap.AnnData(X=small_counts(100, 200))
    with pytest.raises(NameError):
        snap.tl.spectral(ad)
    with pytest.raises(ValueError):
        snap.tl.spectral(ad, features=None, sample_size=1)
    with pytest.raises(NotImplementedError):
        snap.tl.spectral(ad, features=None, distance_metric="jaccard")
    # sample_size below n_obs takes the Nystrom branch
    ev, U = snap.tl.spectral(ad, features=None, n_comps=5, sample_size=0.5, inplace=False)
    assert U.shape[0] == 100 and U.shape[1] == ev.size


def test_panel_segmented_operator(torch_gpu):
    # more than one 32,768-row panel on both sides: 70,000 cells x 80,000 features
    torch = torch_gpu
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    X = small_counts(70000, 80000, 0.0015, seed=12)
    Xd = _to_device_csr(X, torch)
    op_p = SpectralOperator(Xd, layout="panel")
    op_1 = SpectralOperator(Xd, use_panels=False)
    Xh, dinv, _, _ = osp.operator_parts(X)
    rng = np.random.default_rng(1)
    for b in (16, 8, 32, 5):
        V = rng.standard_normal((70000, b))
        Vt = torch.from_numpy(V).cuda()
        Yp = op_p.matmat(Vt).cpu().numpy()
        Y1 = op_1.matmat(Vt).cpu().numpy()
        ref = Xh @ (Xh.T @ V) - dinv[:, None] * V
        scale = np.abs(ref).max()
        assert np.abs(Yp - ref).max() / scale < 1e-5, b
        assert np.abs(Yp - Y1).max() / scale < 1e-5, b
    # bitwise reproducible
    Vt = torch.from_numpy(rng.standard_normal((70000, 16))).cuda()
    assert torch.equal(op_p.matmat(Vt), op_p.matmat(Vt))


def test_panel_fast_product_bitwise(torch_gpu, monkeypatch):
    """The 16-column product over counts < 128 (packed 2: DPP broadcast, 32-bit
    gather offsets, first-of-pair sums started at -128 z) gives the same bits
    as the generic product over the same entries (packed 1), with the same
    work split and summation order; b = 8 and 32 take the generic product on
    both encodings."""
    torch = torch_gpu
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    X = small_counts(70000, 80000, 0.0015, seed=31)
    X.data[::97] = 127  # the largest packed-2 count
    Xd = _to_device_csr(X, torch)
    op2 = SpectralOperator(Xd, layout="panel")
    monkeypatch.setattr(SpectralOperator, "panel_pack_max", 1)
    op1 = SpectralOperator(Xd, layout="panel")
    assert op2.panel_csr.packed == 2 and op2.panel_csc.packed == 2
    assert op1.panel_csr.packed == 1 and op1.panel_csc.packed == 1
    rng = np.random.default_rng(4)
    for b in (16, 11, 8, 32):
        Vt = torch.from_numpy(rng.standard_normal((70000, b))).cuda()
        assert torch.equal(op2.matmat(Vt), op1.matmat(Vt)), b
    Xh, dinv, _, _ = osp.operator_parts(X)
    V = rng.standard_normal((70000, 16))
    ref = Xh @ (Xh.T @ V) - dinv[:, None] * V
    Y = op2.matmat(torch.from_numpy(V).cuda()).cpu().numpy()
    assert np.abs(Y - ref).max() / np.abs(ref).max() < 1e-5


@pytest.mark.parametrize("kind", ["small", "float", "large_counts", "mid_counts", "empty_rows"])
def test_panel_layout_variants(torch_gpu, kind):
    """Panel layout on one panel (forced) and on several: counts < 128 (packed
    2), counts in [128, 65536) (packed 1), float values and counts >= 65536
    (unpacked), rows and columns without entries."""
    torch = torch_gpu
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    n, m = (3000, 5000) if kind == "small" else (40000, 36000)
    X = small_counts(n, m, 0.004 if kind != "small" else 0.02, seed=21)
    if kind == "float":
        X = X.astype(np.float32)
        X.data *= np.float32(0.37)
    elif kind == "large_counts":
        # whole rows scaled to counts >= 65536 (the unpacked encoding); row
        # normalisation keeps the operator on the scale of the other cases
        X = X.astype(np.uint32)
        X.data[X.indptr[0]:X.indptr[50]] *= 70000
    elif kind == "mid_counts":
        X = X.astype(np.uint32)
        X.data[X.indptr[10]:X.indptr[60]] *= 200
    elif kind == "empty_rows":
        X = X.tolil()
        X[100:140, :] = 0
        X[:, 33000:34000] = 0
        X = sp.csr_matrix(X, dtype=np.uint32)
        X.eliminate_zeros()
    Xd = _to_device_csr(X, torch)
    op = SpectralOperator(Xd, layout="panel")
    ref_op = SpectralOperator(Xd, use_panels=False)
    packed = {"small": 2, "empty_rows": 2, "mid_counts": 1, "float": 0, "large_counts": 0}[kind]
    assert op.panel_csr.packed == packed and op.panel_csc.packed == packed
    Xh, dinv, _, _ = osp.operator_parts(X)
    rng = np.random.default_rng(3)
    for b in (16, 7, 32):
        V = rng.standard_normal((n, b))
        Vt = torch.from_numpy(V).cuda()
        Y = op.matmat(Vt).cpu().numpy()
        ref = Xh @ (Xh.T @ V) - dinv[:, None] * V
        scale = np.abs(ref).max()
        assert np.abs(Y - ref).max() / scale < 1e-5, (kind, b)
        assert np.abs(Y - ref_op.matmat(Vt).cpu().numpy()).max() / scale < 1e-5, (kind, b)


def test_warm_start_after_weight_change(torch_gpu):
    """eigsh_lm started from the eigenvectors of nearby weights finds the same
    eigenpairs (to the solver tolerance) as a random start."""
    torch = torch_gpu
    from snapatac2_amd.tl._lanczos import eigsh_lm
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    X = small_counts(6000, 4000, 0.02, seed=31)
    op = SpectralOperator(_to_device_csr(X, torch))
    _, U0, _ = eigsh_lm(op, 12, seed=0)
    w = op.w.clone() * torch.from_numpy(np.random.default_rng(2).uniform(0.97, 1.03, op.m)).cuda()
    op.set_weights(w)
    ev_r, U_r, st_r = eigsh_lm(op, 12, seed=0)
    ev_w, U_w, st_w = eigsh_lm(op, 12, seed=0, v0=U0)
    np.testing.assert_allclose(ev_w, ev_r, rtol=1e-7, atol=1e-9)
    oe, ov = np.linalg.eigh(osp.dense_operator(X, feature_weights=w.cpu().numpy()))
    idx = np.lexsort((-oe, -np.abs(oe)))[:12]
    np.testing.assert_allclose(ev_w, oe[idx], rtol=1e-5, atol=1e-7)
    assert osp.subspace_angle(U_w.cpu().numpy(), ov[:, idx]) < 1e-4


@pytest.mark.parametrize("kind", ["small", "float", "large_counts", "mid_counts", "empty_rows"])
def test_panel_spmm_f64(torch_gpu, kind):
    """FP64 products over both panel layouts (snapgpu_panel_spmm_f64: several
    panels each way, every packing, squared values, a row scale or none)
    against numpy's FP64 sums, to summation rounding."""
    torch = torch_gpu
    import ctypes as C
    from snapatac2_amd import _lib
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    n, m = 40000, 36000
    X = small_counts(n, m, 0.004, seed=22)
    if kind == "float":
        X = X.astype(np.float32)
        X.data *= np.float32(0.37)
    elif kind == "large_counts":
        X.data[X.indptr[0]:X.indptr[50]] *= 70000
    elif kind == "mid_counts":
        X.data[X.indptr[10]:X.indptr[60]] *= 200
    elif kind == "empty_rows":
        X = X.tolil()
        X[100:140, :] = 0
        X[:, 33000:34000] = 0
        X = sp.csr_matrix(X, dtype=np.uint32)
        X.eliminate_zeros()
    if kind == "small":
        X.data[:] = np.minimum(X.data, 3)
    op = SpectralOperator(_to_device_csr(X, torch), layout="panel")
    L = _lib.lib()
    rng = np.random.default_rng(5)
    Xf = sp.csr_matrix(X, dtype=np.float64)
    for layout, A, rows, drows in ((op.panel_csr, Xf, n, m), (op.panel_csc, sp.csr_matrix(Xf.T), m, n)):
        D = rng.standard_normal((drows, 2))
        s = rng.uniform(0.5, 2.0, rows)
        ws_b = int(L.snapgpu_panel_spmm_f64_workspace_bytes(layout.n_pairs))
        ws = torch.empty(ws_b, dtype=torch.uint8, device="cuda")
        Dd = torch.from_numpy(D).cuda()
        sd = torch.from_numpy(s).cuda()
        for square in (0, 1):
            for scaled in (False, True):
                out = torch.full((rows, 2), np.nan, dtype=torch.float64, device="cuda")
                _lib.check(L.snapgpu_panel_spmm_f64(C.byref(layout), Dd.data_ptr(), square,
                                                    sd.data_ptr() if scaled else None, out.data_ptr(),
                                                    ws.data_ptr(), ws_b, _lib.stream_ptr()))
                Aw = A.copy()
                if square:
                    Aw.data = Aw.data ** 2
                ref = Aw @ D
                if scaled:
                    ref = s[:, None] * ref
                got = out.cpu().numpy()
                assert np.isfinite(got).all()
                assert np.abs(got - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1e-300), (kind, square, scaled)
