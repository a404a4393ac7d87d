"""The window layout of the operator (csrc/window.hip) on the GPU: the built
layout equals its CPU model (tests/_window_model.py) in every deterministic part
(slots per row, workgroups, the pieces of each (workgroup, window) bucket, the
quad chunks), the products equal the model's summation to FP32 rounding and the
FP64 oracle within the operator's 1e-5, results are bitwise reproducible, and
matrices the layout cannot hold (counts >= 256, float values) take the panel
layout."""
import numpy as np
import pytest
import scipy.sparse as sp

import _window_model as wm
from oracle import spectral as osp

pytestmark = pytest.mark.gpu


def _counts(n, m, density, seed, heavy_rows=(), heavy_cols=(), maxc=4):
    """Random counts in [1, maxc) plus one entry per row; heavy rows (3,000
    entries: long runs in every window, several slots) and heavy columns (half
    the rows: long runs in the transpose). Overlaps keep the larger count."""
    rng = np.random.default_rng(seed)
    X = sp.random(n, m, density=density, random_state=rng, format="csr",
                  data_rvs=lambda k: rng.integers(1, maxc, k)).astype(np.float64)
    rr, cc = [np.arange(n)], [rng.integers(0, m, n)]
    for r in heavy_rows:
        cols = rng.choice(m, size=min(m, 3000), replace=False)
        rr.append(np.full(cols.size, r))
        cc.append(cols)
    for c in heavy_cols:
        rows = rng.choice(n, size=n // 2, replace=False)
        rr.append(rows)
        cc.append(np.full(rows.size, c))
    rr, cc = np.concatenate(rr), np.concatenate(cc)
    H = sp.csr_matrix((np.ones(rr.size), (rr, cc)), shape=(n, m))
    H.data[:] = 1.0  # duplicates summed by the constructor: back to 1
    X = sp.csr_matrix(X.maximum(H), dtype=np.uint32)
    X.sort_indices()
    return X


def _ops(torch, X, **kw):
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    Xd = _to_device_csr(X, torch)
    return SpectralOperator(Xd, layout="window", **kw), SpectralOperator(Xd, layout="panel", **kw)


def _check_direction(keep, ptr, idx, val, dense_rows):
    ent, bstart, qoff, block_row, row_slot = (t.cpu().numpy() for t in keep)
    br, rs = wm.plan(ptr, idx)
    assert np.array_equal(block_row, br)
    assert np.array_equal(row_slot.view(np.uint32), rs)
    buckets, nw = wm.bucket_pieces(ptr, idx, val, br, rs, dense_rows)
    nb = len(br) - 1
    assert bstart.size == nb * nw + 1 and bstart[-1] == idx.size
    words = ent.view(np.uint32)
    q = qoff.view(np.uint32).reshape(nb * nw, wm.QUADS + 1)
    for k in range(nb * nw):
        got = words[bstart[k]:bstart[k + 1]].tolist()
        b, w = divmod(k, nw)
        assert sorted(wm.split_pieces(got)) == buckets.get((b, w), []), (b, w)
        assert q[k].tolist() == wm.quad_offsets(got), (b, w)


@pytest.mark.parametrize("case", ["small", "heavy"])
def test_window_layout_matches_model(torch_gpu, case):
    torch = torch_gpu
    if case == "small":
        X = _counts(3000, 5000, 0.01, seed=1)
    else:
        X = _counts(5000, 4000, 0.004, seed=2, heavy_rows=(7, 100, 4999), heavy_cols=(3, 2047, 3999), maxc=256)
    op, _ = _ops(torch, X)
    assert op.layout == "window"
    n, m = X.shape
    _check_direction(op._panel_keep[0], X.indptr, X.indices, X.data, m)
    Xt = sp.csr_matrix(X.T)
    Xt.sort_indices()
    _check_direction(op._panel_keep[1], Xt.indptr, Xt.indices, Xt.data, n)


def test_window_products_match_model_summation(torch_gpu):
    """Both products in the kernel's order (pieces in entry order, slots in
    window order, a row's slots in slot order), FP32 with fused multiply-adds."""
    torch = torch_gpu
    X = _counts(2500, 3000, 0.01, seed=3, heavy_rows=(11,), heavy_cols=(5,))
    op, _ = _ops(torch, X)
    assert op.layout == "window"
    rng = np.random.default_rng(7)
    V = rng.standard_normal((2500, 16))
    Vt = torch.from_numpy(V).cuda()
    Y = op.matmat(Vt).cpu().numpy()
    r32 = op.r32.cpu().numpy()
    w2 = op.w2.cpu().numpy()
    Vs = (r32[:, None].astype(np.float64) * V).astype(np.float32)
    Xt = sp.csr_matrix(X.T)
    Xt.sort_indices()
    brc, rsc = wm.plan(Xt.indptr, Xt.indices)
    Z = wm.apply(Xt.indptr, Xt.indices, Xt.data, brc, rsc, Vs, w2)
    br, rs = wm.plan(X.indptr, X.indices)
    S = wm.apply(X.indptr, X.indices, X.data, br, rs, Z, r32)
    Yref = S.astype(np.float64) - op.dinv.cpu().numpy()[:, None] * V
    scale = np.abs(Yref).max()
    assert np.abs(Y - Yref).max() <= 1e-6 * scale


@pytest.mark.parametrize("b", [16, 5, 1, 21])
def test_window_operator_vs_oracle(torch_gpu, b):
    torch = torch_gpu
    # several workgroups and windows in both directions, heavy rows and columns
    X = _counts(70000, 80000, 0.0015, seed=12, heavy_rows=(5, 69999), heavy_cols=(0, 40000, 79999))
    op, op_p = _ops(torch, X)
    assert op.layout == "window" and op.window_csr.n_blocks > 1 and op.window_csc.n_windows > 50
    Xh, dinv, _, _ = osp.operator_parts(X)
    V = np.random.default_rng(b).standard_normal((70000, b))
    Vt = torch.from_numpy(V).cuda()
    Y = op.matmat(Vt)
    ref = Xh @ (Xh.T @ V) - dinv[:, None] * V
    scale = np.abs(ref).max()
    assert np.abs(Y.cpu().numpy() - ref).max() / scale < 1e-5
    assert np.abs(Y.cpu().numpy() - op_p.matmat(Vt).cpu().numpy()).max() / scale < 1e-5
    assert torch.equal(Y, op.matmat(Vt))  # bitwise reproducible


def test_window_fallbacks(torch_gpu):
    torch = torch_gpu
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    X = _counts(3000, 4000, 0.01, seed=4)
    X.data[::50] = 300  # a count the window entry cannot hold
    op = SpectralOperator(_to_device_csr(X, torch), layout="window")
    assert op.layout == "panel"
    Xf = sp.csr_matrix(_counts(3000, 4000, 0.01, seed=5), dtype=np.float32) * np.float32(0.5)
    opf = SpectralOperator(_to_device_csr(Xf, torch), layout="window")
    assert opf.layout == "panel"
    Xh, dinv, _, _ = osp.operator_parts(X)
    V = np.random.default_rng(0).standard_normal((3000, 8))
    ref = Xh @ (Xh.T @ V) - dinv[:, None] * V
    got = op.matmat(torch.from_numpy(V).cuda()).cpu().numpy()
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-5


def test_window_empty_rows_columns_and_one_window(torch_gpu):
    """Rows and columns without entries (zero slots; zero sums), and a matrix
    whose dense operands fit one window on both sides."""
    torch = torch_gpu
    for n, m in ((900, 700), (5000, 3000)):
        X = _counts(n, m, 0.01, seed=n).tolil()
        X[10:40, :] = 0
        X[:, 100:130] = 0
        X = sp.csr_matrix(X, dtype=np.uint32)
        X.eliminate_zeros()
        op, op_p = _ops(torch, X)
        assert op.layout == "window"
        Xh, dinv, _, _ = osp.operator_parts(X)
        V = np.random.default_rng(1).standard_normal((n, 16))
        ref = Xh @ (Xh.T @ V) - dinv[:, None] * V
        got = op.matmat(torch.from_numpy(V).cuda()).cpu().numpy()
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-5
        rs = op._panel_keep[0][4].cpu().numpy().view(np.uint32)
        assert (rs[10:40] >> 16 == 0).all()  # empty rows own no slot


def test_window_eigenpairs(torch_gpu):
    """The block Lanczos solver on the window layout gives the panel layout's
    eigenvalues (1e-7 relative; both within the operator's FP32 rounding)."""
    torch = torch_gpu
    from snapatac2_amd.tl._lanczos import eigsh_lm
    X = _counts(40000, 36000, 0.004, seed=6)
    op, op_p = _ops(torch, X)
    ev, _, _ = eigsh_lm(op, 10, seed=0)
    ev_p, _, _ = eigsh_lm(op_p, 10, seed=0)
    assert np.allclose(np.sort(ev), np.sort(ev_p), rtol=1e-7, atol=0)


@pytest.mark.parametrize("b", [16, 3, 21])
def test_mixed_layout(torch_gpu, monkeypatch, b):
    """The default layout of large operators: CSC product over panels, CSR
    product over windows. Against the FP64 oracle and the panel layout, bitwise
    reproducible; matrices the windows cannot hold take the panel layout, and so
    does the default when the CSR windows come to fewer workgroups than
    MIXED_MIN_BLOCKS (this matrix: a few dozen)."""
    torch = torch_gpu
    from snapatac2_amd.tl import _operator
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    X = _counts(70000, 80000, 0.0015, seed=13, heavy_rows=(5, 69999), heavy_cols=(0, 79999))
    Xd = _to_device_csr(X, torch)
    if b == 16:
        assert SpectralOperator(Xd).layout == "panel"
        assert SpectralOperator(Xd, layout="mixed").layout == "mixed"
    monkeypatch.setattr(_operator, "MIXED_MIN_BLOCKS", 1)
    op = SpectralOperator(Xd)
    assert op.layout == "mixed" and op.window_csr.n_blocks > 1 and op.panel_csc.n_chunks > 1
    op_p = SpectralOperator(Xd, layout="panel")
    Xh, dinv, _, _ = osp.operator_parts(X)
    V = np.random.default_rng(b).standard_normal((70000, b))
    Vt = torch.from_numpy(V).cuda()
    Y = op.matmat(Vt)
    ref = Xh @ (Xh.T @ V) - dinv[:, None] * V
    scale = np.abs(ref).max()
    assert np.abs(Y.cpu().numpy() - ref).max() / scale < 1e-5
    assert np.abs(Y.cpu().numpy() - op_p.matmat(Vt).cpu().numpy()).max() / scale < 1e-5
    assert torch.equal(Y, op.matmat(Vt))
    if b == 3:
        Xc = X.copy()
        Xc.data[::50] = 300  # a count the window entry cannot hold
        assert SpectralOperator(_to_device_csr(Xc, torch), layout="mixed").layout == "panel"
        Xf = sp.csr_matrix(X, dtype=np.float32) * np.float32(0.5)
        assert SpectralOperator(_to_device_csr(Xf, torch), layout="mixed").layout == "panel"
