"""Parity of the full pipeline at its benchmark size: 1,000,000 generator cells
(6.9e9 fragments in 26 chunks, 500 bp tiles, top 500,000 features, 30
eigenpairs, k = 50), the workload bench.py times.

The oracle cannot run on the whole input, so each stage is checked where an
independent computation can reach it:
- tile matrix (data_iter.rs:346-415): sampled rows bit-exact against the C++
  oracle on those cells' fragments, including the rows whose entries straddle
  the 2^31 and 2^32 offsets of the 6.9e9-entry row pointers, the first rows and
  the last chunk's rows;
- column sums (_basic.py:1053-1055) equal to FP64 index_add_ sums of every
  entry (exact for integers), the selected matrix's column sums and entry
  counts equal to those of the tile matrix on the selected columns, and its
  sampled rows equal to the tile rows restricted to them;
- the operator (embedding.rs:133-171, 269-326): IDF weights and degrees
  against an FP64 restatement through torch gathers and index_add_
  (oracle.adjoint.ChunkedCounts), and one application to a random block;
- eigenpairs (embedding.rs:155-171): residuals and Rayleigh quotients of the
  FP64 operator, orthonormality;
- kNN (knn.rs:9-32, 58-78) on the embedding: sampled rows identical to FP64
  brute force, distances to 1e-12.
"""
import numpy as np
import pytest

from oracle import adjoint as oadj
from oracle import tile as otile

pytestmark = pytest.mark.gpu

N = 1_000_000
EXCLUDE = ["chrM", "chrY", "M", "Y"]


@pytest.fixture(scope="module")
def state(torch_gpu):
    return {"torch": torch_gpu}


def _tile(state):
    if "X" not in state:
        from snapatac2_amd import synth
        from snapatac2_amd.genome import synthetic_genome
        from snapatac2_amd.pp import make_tile_matrix_chunked
        print("c5: generating 1M cells and tiling", flush=True)
        chunks = list(synth.iter_chunks(N, seed=0))
        X, names = make_tile_matrix_chunked(chunks, synthetic_genome(), 500, EXCLUDE)
        state.update(chunks=chunks, X=X, n_bins=len(names))
    return state["X"]


def _host_rows(X, rows):
    out = []
    for r in rows:
        a, b = int(X.indptr[r].item()), int(X.indptr[r + 1].item())
        out.append((X.indices[a:b].cpu().numpy().astype(np.int64), X.data[a:b].cpu().numpy().view(np.uint32)))
    return out


def test_c5_tile_sampled_rows(state):
    torch = state["torch"]
    from snapatac2_amd.genome import synthetic_genome
    X = _tile(state)
    chunks = state["chunks"]
    assert X.shape == (N, 6_062_095)
    assert X.nnz > 6e9 and X.indptr.dtype == torch.int64
    ip = X.indptr.cpu().numpy()
    rng = np.random.default_rng(5)
    rows = set(rng.choice(N, 300, replace=False).tolist()) | set(range(5)) | set(range(N - 5, N))
    for edge in (1 << 31, 1 << 32, 3 << 31, 6 << 30):
        r = int(np.searchsorted(ip, edge, side="right")) - 1
        assert ip[r] <= edge < ip[r + 1] or ip[r] == edge
        rows |= {r - 1, r, r + 1}
    rows |= set(range(chunks[-1][0], chunks[-1][0] + 3))
    sample = np.sort(np.asarray(sorted(rows), dtype=np.int64))
    cell_ptr, pos, ln = [0], [], []
    for c0, fr in chunks:
        sel = sample[(sample >= c0) & (sample < c0 + fr.shape[0])] - c0
        if sel.size == 0:
            continue
        cp = fr.cell_ptr.cpu().numpy()
        for s in sel:
            a, b = int(cp[s]), int(cp[s + 1])
            pos.append(fr.pos[a:b].cpu().numpy().view(np.uint32))
            ln.append(fr.values[a:b].cpu().numpy().view(np.uint32))
            cell_ptr.append(cell_ptr[-1] + (b - a))
    O = otile.tile_matrix(np.asarray(cell_ptr, dtype=np.uint64), np.concatenate(pos), np.concatenate(ln),
                          synthetic_genome(), 500, EXCLUDE)
    for k, (idx, val) in enumerate(_host_rows(X, sample)):
        a, b = O.indptr[k], O.indptr[k + 1]
        assert np.array_equal(idx, O.indices[a:b]), sample[k]
        assert np.array_equal(val, O.data[a:b]), sample[k]
    print(f"c5: {sample.size} tile rows bit-exact (row pointers up to {ip[-1]})", flush=True)


def test_c5_column_sums_and_selection(state):
    torch = state["torch"]
    from snapatac2_amd.pp._basic import _column_sums, _find_most_accessible_features
    from snapatac2_amd.tl._operator import select_columns
    X = _tile(state)
    state.pop("chunks", None)
    count = _column_sums(X)
    T = oadj.ChunkedCounts(X.indptr, X.indices, X.data, X.shape)
    ref = T.col_sums().cpu().numpy()
    assert np.array_equal(count, ref)
    df = T.df().cpu().numpy()
    del T
    sel = np.sort(_find_most_accessible_features(count, 0.005, 0.005, 500_000))
    assert sel.size == 500_000
    Xs = select_columns(X, sel)
    assert Xs.shape == (N, 500_000)
    assert Xs.nnz == int(df[sel].sum())
    Ts = oadj.ChunkedCounts(Xs.indptr, Xs.indices, Xs.data, Xs.shape)
    assert np.array_equal(Ts.col_sums().cpu().numpy(), count[sel])
    assert np.array_equal(Ts.df().cpu().numpy(), df[sel])
    del Ts
    rng = np.random.default_rng(6)
    ip = Xs.indptr.cpu().numpy()
    extra = [N - 1, int(np.searchsorted(ip, ip[-1] // 2, side="right")) - 1]
    if ip[-1] > (1 << 31):  # the row whose entries straddle offset 2^31
        extra.append(int(np.searchsorted(ip, 1 << 31, side="right")) - 1)
    rows = np.unique(np.concatenate([rng.choice(N, 200, replace=False), extra]))
    for (ti, tv), (si, sv) in zip(_host_rows(X, rows), _host_rows(Xs, rows)):
        keep = np.isin(ti, sel)
        assert np.array_equal(np.searchsorted(sel, ti[keep]), si)
        assert np.array_equal(tv[keep], sv)
    state["Xs"] = Xs
    state.pop("X")
    torch.cuda.empty_cache()
    print(f"c5: column sums exact; selected matrix {Xs.nnz} entries", flush=True)


def _op(state):
    if "op" not in state:
        from snapatac2_amd.tl._operator import SpectralOperator
        Xs = state["Xs"]
        state["op"] = SpectralOperator(Xs)
        state["T"] = oadj.ChunkedCounts(Xs.indptr, Xs.indices, Xs.data, Xs.shape, chunk_entries=1 << 26)
        u, d, a = state["T"].scales(state["op"].w)
        state["scales"] = (u, d, a)
    return state["op"], state["T"]


def _apply64(state, w, V):
    T = state["T"]
    u, d, a = state["scales"]
    r = (u * a)[:, None]
    return r * T.mm((w * w)[:, None] * T.rmm(r * V)) - V / d[:, None]


def test_c5_operator_application(state):
    torch = state["torch"]
    if "Xs" not in state:
        pytest.skip("needs the selected matrix of test_c5_column_sums_and_selection")
    op, T = _op(state)
    assert op.use_panels
    w64 = oadj.idf_from_df(T.df(), N)
    assert float(((op.w - w64).abs() / w64.abs().clamp_min(1e-300)).max()) < 1e-14
    u, d, a = state["scales"]
    assert float(((op.s - u).abs() / u).max()) < 1e-12
    assert float(((1.0 / op.dinv - d).abs() / d.abs()).max()) < 1e-9
    gen = torch.Generator(device="cuda").manual_seed(3)
    V = torch.randn(N, 16, dtype=torch.float64, device="cuda", generator=gen)
    Y = op.matmat(V)
    Y64 = _apply64(state, op.w, V)
    rel = (torch.linalg.norm(Y - Y64, dim=0) / torch.linalg.norm(Y64, dim=0)).max().item()
    # FP32 sparse products (about 2^-24 per term, ~2,000 terms per row)
    assert rel < 1e-5, rel
    print(f"c5: operator application within {rel:.2e} of FP64", flush=True)


def test_c5_eigenpairs_and_knn(state):
    torch = state["torch"]
    if "Xs" not in state:
        pytest.skip("needs the selected matrix of test_c5_column_sums_and_selection")
    from snapatac2_amd.pp._knn import nearest_neighbour_graph
    from snapatac2_amd.tl._lanczos import eigsh_lm
    op, T = _op(state)
    evals, U, st = eigsh_lm(op, 30, seed=0)
    lam = torch.from_numpy(np.asarray(evals)).to(U.device)
    AU = torch.cat([_apply64(state, op.w, U[:, c:c + 8]) for c in range(0, 30, 8)], dim=1)
    res = torch.linalg.norm(AU - U * lam, dim=0)
    scale = float(np.abs(evals).max())
    assert float(res.max()) < 1e-5 * scale, (res.max().item(), scale)
    rq = (U * AU).sum(0)
    assert float(((rq - lam).abs() / lam.abs()).max()) < 1e-5
    G = U.T @ U
    assert float((G - torch.eye(30, dtype=G.dtype, device=G.device)).abs().max()) < 1e-10
    # embedding (weighted_by_sd) and its exact kNN graph
    ix = np.argsort(evals, kind="stable")[::-1]
    ev = np.asarray(evals)[ix]
    keep = ix[ev > 0]
    E = (U[:, torch.from_numpy(keep.copy()).cuda()] * torch.from_numpy(np.sqrt(np.asarray(evals)[keep])).cuda())
    E = E.contiguous()
    k = 50
    g = nearest_neighbour_graph(E, k, precision="bf16")
    rng = np.random.default_rng(2)
    for q in rng.choice(N, 100, replace=False):
        d2 = ((E - E[q]) ** 2).sum(1)
        d2[q] = float("inf")
        # stable sort: equal distances keep index order (ties broken by index)
        exact = np.sort(torch.sort(d2, stable=True).indices[:k].cpu().numpy())
        got = g.indices[g.indptr[q]:g.indptr[q + 1]]
        assert np.array_equal(got, exact), q
        np.testing.assert_allclose(g.data[g.indptr[q]:g.indptr[q + 1]], np.sqrt(d2.cpu().numpy()[exact]),
                                   rtol=1e-12)
    print(f"c5: 30 eigenpairs (residual {res.max().item():.1e}), kNN rows exact", flush=True)
