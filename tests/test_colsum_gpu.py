"""Column sums and column counts of CSR matrices on the GPU (the windowed
kernel for rows with ascending columns, the per-entry path for rows out of
order), exact against numpy's integer sums.

The windowed kernel works on blocks of 3968 rows and windows of 32768
columns with 32-bit bins: the shapes below put rows and columns on both sides
of those boundaries, include empty rows, rows spanning several windows and
long runs of consecutive columns (many 32-entry steps inside one window). The
status word tells which path ran: 0 = windows; bit 2 = a block's values add up
to 2^32 or more (values up to 2^32 - 1), where the 64-bit per-entry path takes
over.
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


def _sums(indptr, idx, val, m, bits, torch, status=None):
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
    if status is not None:
        status.append(int(ws[0].item()))
    return out.cpu().numpy().view(np.uint64)


def _expected(idx, val, m):
    ref = np.zeros(m, dtype=np.uint64)
    np.add.at(ref, idx, val.astype(np.uint64))
    return ref


@pytest.mark.parametrize("n,m,per_row,bits,vmax", [
    (1, 1, 1, 32, 1000),
    (3967, 100, 20, 32, 1000),
    (3969, 16384, 40, 32, 1000),
    (8000, 16385, 60, 64, 1000),
    (3969, 32768, 40, 32, 1000),
    (8000, 32769, 60, 64, 1000),
    (10001, 50000, 30, 32, 1000),
    (5000, 200000, 300, 64, 1000),
    (5000, 200000, 300, 32, 2**32 - 1),
    (8000, 32769, 60, 64, 2**32 - 1),
])
def test_sorted_rows_exact(torch_gpu, n, m, per_row, bits, vmax):
    indptr, idx, val = _random_csr(n, m, per_row, seed=n + m, vmax=vmax)
    status = []
    got = _sums(indptr, idx, val, m, bits, torch_gpu, status)
    assert np.array_equal(got, _expected(idx, val, m))
    assert status == [0 if vmax < 2**16 else 4]


def test_block_sum_at_the_wrap(torch_gpu):
    """A block whose values add up to just below 2^32 stays on the windows; one
    more unit sends the sums to the per-entry path; both exact."""
    n, m = 100, 5000
    indptr, idx, val = _random_csr(n, m, 20, seed=4, vmax=1)
    val[:] = 0
    for extra, want in ((0, 0), (1, 4)):
        v = val.copy()
        v[0] = 2**32 - 2
        v[1] = 1 + extra
        status = []
        got = _sums(indptr, idx, v, m, 32, torch_gpu, status)
        assert np.array_equal(got, _expected(idx, v, m))
        assert status == [want]


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


def _select_raw(indptr, idx, val, m, cols, torch, bits=32):
    """select_count, scan and select_fill through the C-ABI."""
    L = _lib.lib()
    dev = torch.device("cuda")
    n = len(indptr) - 1
    st = _lib.stream_ptr()
    ip = torch.from_numpy(indptr).to(dev)
    ix = torch.from_numpy(idx.astype(np.int32 if bits == 32 else np.int64)).to(dev)
    vv = torch.from_numpy(val.view(np.int32)).to(dev)
    dc = torch.from_numpy(np.asarray(cols, dtype=np.int64)).to(dev)
    sel = torch.empty(2 * (-(-m // 32)), dtype=torch.int32, device=dev)
    _lib.check(L.snapgpu_col_select_table(dc.data_ptr(), len(cols), m, sel.data_ptr(), st))
    new_ip = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    _lib.check(L.snapgpu_csr_select_count(ip.data_ptr(), ix.data_ptr(), bits, n, sel.data_ptr(), m,
                                          new_ip.data_ptr() + 8, st))
    ws_b = int(L.snapgpu_scan_workspace_bytes(n))
    ws = torch.empty(max(ws_b, 8), dtype=torch.uint8, device=dev)
    _lib.check(L.snapgpu_scan_inclusive_i64(new_ip.data_ptr() + 8, n, ws.data_ptr(), ws_b, st))
    nnz = int(new_ip[-1].item())
    ni = torch.full((max(nnz, 1),), -7, dtype=torch.int32, device=dev)
    nv = torch.full((max(nnz, 1),), -7, dtype=torch.int32, device=dev)
    _lib.check(L.snapgpu_csr_select_fill(ip.data_ptr(), ix.data_ptr(), bits, vv.data_ptr(), n, sel.data_ptr(),
                                         m, new_ip.data_ptr(), ni.data_ptr(), nv.data_ptr(), st))
    return new_ip.cpu().numpy(), ni.cpu().numpy()[:nnz], nv.cpu().numpy()[:nnz].view(np.uint32)


def _select_model(indptr, idx, val, m, cols):
    """Entries whose column is selected (columns outside [0, m) select nothing),
    in row order, renumbered by rank among the selected columns."""
    rank = np.full(m, -1, dtype=np.int64)
    rank[np.asarray(cols, dtype=np.int64)] = np.arange(len(cols))
    ok = (idx >= 0) & (idx < m)
    nc = np.full(idx.size, -1, dtype=np.int64)
    nc[ok] = rank[idx[ok]]
    keep = nc >= 0
    row = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    new_ip = np.zeros(len(indptr), dtype=np.int64)
    np.cumsum(np.bincount(row[keep], minlength=len(indptr) - 1), out=new_ip[1:])
    return new_ip, nc[keep], val[keep]


@pytest.mark.parametrize("case", ["sorted", "shuffled"])
def test_select_windows(torch_gpu, case):
    """The windowed selection passes (262,144 columns of table per window)
    over 1.1M columns (five windows, the last one partial): rows with entries
    in every window, empty rows, blocks of 512 rows on both sides of n; rows
    out of order and columns outside [0, m) (negative, past the table) in the
    shuffled case, where every entry must still be visited once, in row order."""
    torch = torch_gpu
    n, m = 1500, 1_100_000
    indptr, idx, val = _random_csr(n, m, 150, seed=11, vmax=1000)
    rng = np.random.default_rng(12)
    cols = np.flatnonzero(rng.random(m) < 0.3)
    if case == "shuffled":
        for r in range(0, n, 7):
            a, b = indptr[r], indptr[r + 1]
            idx[a:b] = rng.permutation(idx[a:b])
        bad = rng.random(idx.size) < 0.01
        idx[bad] = rng.choice(np.array([-1, -5000, m, m + 31, 2**31 - 1]), bad.sum())
    got = _select_raw(indptr, idx, val, m, cols, torch)
    ref = _select_model(indptr, idx, val, m, cols)
    assert np.array_equal(got[0], ref[0])
    assert np.array_equal(got[1], ref[1])
    assert np.array_equal(got[2], ref[2])
    if case == "sorted":  # the per-entry kernels (64-bit indices) agree
        got64 = _select_raw(indptr, idx, val, m, cols, torch, bits=64)
        assert all(np.array_equal(a, b) for a, b in zip(got, got64))
