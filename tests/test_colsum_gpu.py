"""Column sums and column counts of CSR matrices on the GPU (the windowed
kernel for rows with ascending columns, the per-entry path for rows out of
order), exact against numpy's integer sums.

The windowed kernel works on blocks of 3968 rows and windows of 16384
columns: the shapes below put rows and columns on both sides of those
boundaries, include empty rows, rows spanning several windows and long runs
of consecutive columns (many 32-entry steps inside one window), and values up
to 2^32 - 1 (64-bit sums).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from snapatac2_amd import _lib

pytestmark = pytest.mark.gpu


def _random_csr(n, m, per_row, seed, vmax=2**32 - 1):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 2 * per_row + 1, n)
    lens[rng.random(n) < 0.05] = 0  # empty rows
    lens = np.minimum(lens, m)
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=indptr[1:])
    idx = np.concatenate([np.sort(rng.choice(m, k, replace=False)) for k in lens] or [np.zeros(0)])
    val = rng.integers(1, vmax, indptr[-1], dtype=np.uint64, endpoint=True).astype(np.uint32)
    return indptr, idx.astype(np.int64), val


def _sums(indptr, idx, val, m, bits, torch):
    L = _lib.lib()
    dev = torch.device("cuda")
    ip = torch.from_numpy(indptr).to(dev)
    ix = torch.from_numpy(idx.astype(np.int32 if bits == 32 else np.int64)).to(dev)
    vv = torch.from_numpy(val.view(np.int32)).to(dev)
    out = torch.full((m,), -1, dtype=torch.int64, device=dev)
    ws = torch.empty(4, dtype=torch.int32, device=dev)
    _lib.check(L.snapgpu_csr_col_sum_rows_u32(ip.data_ptr(), len(indptr) - 1, ix.data_ptr(), bits,
                                              vv.data_ptr(), m, out.data_ptr(), ws.data_ptr(),
                                              _lib.stream_ptr()))
    return out.cpu().numpy().view(np.uint64)


def _expected(idx, val, m):
    ref = np.zeros(m, dtype=np.uint64)
    np.add.at(ref, idx, val.astype(np.uint64))
    return ref


@pytest.mark.parametrize("n,m,per_row,bits", [
    (1, 1, 1, 32),
    (3967, 100, 20, 32),
    (3969, 16384, 40, 32),
    (8000, 16385, 60, 64),
    (10001, 50000, 30, 32),
    (5000, 200000, 300, 64),
])
def test_sorted_rows_exact(torch_gpu, n, m, per_row, bits):
    indptr, idx, val = _random_csr(n, m, per_row, seed=n + m)
    got = _sums(indptr, idx, val, m, bits, torch_gpu)
    assert np.array_equal(got, _expected(idx, val, m))


def test_long_runs_and_wide_rows(torch_gpu):
    """Rows of up to 60000 consecutive columns: thousands of 32-entry steps,
    several windows per row."""
    n, m = 300, 70000
    rng = np.random.default_rng(5)
    rows = []
    for i in range(n):
        a = int(rng.integers(0, m))
        b = min(m, a + int(rng.integers(0, 60000)))
        rows.append(np.arange(a, b))
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum([len(r) for r in rows], out=indptr[1:])
    idx = np.concatenate(rows)
    val = rng.integers(0, 5, idx.size).astype(np.uint32)
    got = _sums(indptr, idx, val, m, 32, torch_gpu)
    assert np.array_equal(got, _expected(idx, val, m))


def test_rows_out_of_order_fall_back(torch_gpu):
    indptr, idx, val = _random_csr(6000, 40000, 50, seed=11, vmax=1000)
    rng = np.random.default_rng(0)
    for r in rng.choice(6000, 50, replace=False):  # shuffle some rows
        a, b = indptr[r], indptr[r + 1]
        idx[a:b] = rng.permutation(idx[a:b])
    got = _sums(indptr, idx, val, 40000, 32, torch_gpu)
    assert np.array_equal(got, _expected(idx, val, 40000))


def test_empty_matrix(torch_gpu):
    indptr = np.zeros(5, dtype=np.int64)
    got = _sums(indptr, np.zeros(0, dtype=np.int64), np.zeros(0, dtype=np.uint32), 7, 32, torch_gpu)
    assert np.array_equal(got, np.zeros(7, dtype=np.uint64))


@pytest.mark.parametrize("shuffle", [False, True])
def test_transpose_counts(torch_gpu, shuffle):
    """The CSC transpose counts columns with the windowed kernel (sorted rows)
    or per entry (rows out of order); both give scipy's CSC."""
    torch = torch_gpu
    from snapatac2_amd._device import DeviceCSR
    from snapatac2_amd.tl._operator import SpectralOperator
    indptr, idx, val = _random_csr(5000, 20000, 40, seed=3, vmax=7)
    if shuffle:
        rng = np.random.default_rng(1)
        for r in range(0, 5000, 97):
            a, b = indptr[r], indptr[r + 1]
            idx[a:b] = rng.permutation(idx[a:b])
    dev = torch.device("cuda")
    D = DeviceCSR(torch.from_numpy(indptr).to(dev), torch.from_numpy(idx.astype(np.int32)).to(dev),
                  torch.from_numpy(val.view(np.int32)).to(dev), (5000, 20000))
    op = SpectralOperator(D)
    Xc = sp.csr_matrix((val, idx, indptr), shape=(5000, 20000)).tocsc()
    Xc.sort_indices()
    assert np.array_equal(op.colptr.cpu().numpy(), Xc.indptr)
    assert np.array_equal(op.rowidx.cpu().numpy(), Xc.indices)
    assert np.array_equal(op.cval.cpu().numpy().view(np.uint32), Xc.data)


@pytest.mark.parametrize("m,frac,bits", [(1, 1.0, 32), (31, 0.5, 32), (33, 0.3, 64), (70001, 0.1, 32),
                                         (100000, 0.0, 32), (65, 1.0, 64)])
def test_select_columns_matches_scipy(torch_gpu, m, frac, bits):
    """select_columns (selection table of 2 words per 32 columns) against
    scipy column slicing: same rows, renumbered columns, same values."""
    torch = torch_gpu
    from snapatac2_amd._device import DeviceCSR
    from snapatac2_amd.tl._operator import select_columns
    n = 3000
    indptr, idx, val = _random_csr(n, m, min(m, 40), seed=m, vmax=100)
    rng = np.random.default_rng(m)
    cols = np.flatnonzero(rng.random(m) < frac)
    if frac == 1.0:
        cols = np.arange(m)
    dev = torch.device("cuda")
    D = DeviceCSR(torch.from_numpy(indptr).to(dev),
                  torch.from_numpy(idx.astype(np.int32 if bits == 32 else np.int64)).to(dev),
                  torch.from_numpy(val.view(np.int32)).to(dev), (n, m))
    S = select_columns(D, cols)
    ref = sp.csr_matrix((val, idx, indptr), shape=(n, m))[:, cols]
    ref.sort_indices()
    assert S.shape == ref.shape
    assert np.array_equal(S.indptr.cpu().numpy(), ref.indptr)
    assert np.array_equal(S.indices.cpu().numpy(), ref.indices)
    assert np.array_equal(S.data.cpu().numpy().view(np.uint32), ref.data)
