"""Drop-ins for snapatac2.pp.add_tile_matrix and snapatac2.pp.select_features.

Reference: snapatac2-python/python/snapatac2/preprocessing/_basic.py
  add_tile_matrix            :454-575
  _find_most_accessible_features :952-967
  select_features            :969-1095
and the Rust it calls: mk_tile_matrix (snapatac2-python/src/preprocessing.rs:286-363),
create_tile_matrix (snapatac2-core/src/feature_count/matrix.rs:26-93).
"""
from __future__ import annotations

import ctypes as C
import logging
import threading
from typing import Literal

import numpy as np

from .. import _lib
from .._device import DeviceCSR, DeviceFragments
from ..anndata import AnnData
from ..genome import BinIndex, read_chrom_sizes
from ._regions import intersect_bed

FRAGMENT_SINGLE = "fragment_single"  # snapatac2-core/src/feature_count/mod.rs:20
FRAGMENT_PAIRED = "fragment_paired"  # mod.rs:23

_STRATEGY = {"insertion": 0, "fragment": 1, "paired-insertion": 2}
_SINGLE = 3
_VALUE_TYPES = ("target", "total", "fraction")        # preprocessing.rs:347-354
_SUMMARY_TYPES = ("sum", "mean", "count")             # preprocessing.rs:356-363


def _sorted_rows(frags: DeviceFragments) -> DeviceFragments:
    """A copy of `frags` with every row sorted by (position, value)
    (snapgpu_frag_sort_rows), for the tile passes, which need rows in position
    order; the reference's per-row BTreeMap (data_iter.rs:362-411) gives the same
    matrix for any order."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    pos = frags.pos.clone()
    vals = frags.values.clone()
    n, F = frags.shape[0], frags.n_fragments
    wsb = int(L.snapgpu_frag_sort_rows_workspace_bytes(n, F))
    ws = torch.empty(wsb, dtype=torch.uint8, device=frags.device)
    _lib.check(L.snapgpu_frag_sort_rows(frags.cell_ptr.data_ptr(), n, F, pos.data_ptr(), vals.data_ptr(),
                                        ws.data_ptr(), wsb, _lib.stream_ptr()))
    return DeviceFragments(frags.cell_ptr, pos, vals, frags.n_cols, frags.paired)


def _tile_count(L, frags, table, n_chrom, bin_size, strategy, min_len, max_len, row_nnz_ptr, order,
                status, level, stream) -> DeviceFragments:
    """snapgpu_tile_count; a row out of position order is sorted (on a copy) and
    the pass repeated. Returns the fragments the fill pass must read."""
    for attempt in range(2):
        rc = L.snapgpu_tile_count(frags.cell_ptr.data_ptr(), _lib.ptr(frags.pos),
                                  _lib.ptr(frags.values) if frags.paired else 0, frags.shape[0],
                                  table.data_ptr(), n_chrom, int(bin_size), strategy, min_len, max_len,
                                  row_nnz_ptr, _lib.ptr(order), status.data_ptr(), C.byref(level), stream)
        if rc == _lib.ERR_UNSORTED and attempt == 0:
            frags = _sorted_rows(frags)
            level.value = 0
            continue
        _lib.check(rc)
        return frags


def _as_device_fragments(m, paired: bool) -> DeviceFragments:
    if isinstance(m, DeviceFragments):
        return m
    torch = _lib.require_gpu()
    return DeviceFragments.from_scipy(m, paired=paired, device=torch.device("cuda"))


def make_tile_matrix(
    frags: DeviceFragments,
    chrom_sizes: dict[str, int],
    bin_size: int = 500,
    exclude_chroms: list[str] | None = None,
    min_frag_size: int | None = None,
    max_frag_size: int | None = None,
    counting_strategy: str = "paired-insertion",
) -> tuple[DeviceCSR, BinIndex]:
    """Cell x bin count matrix on the device (uint32 counts, int64 row pointers).

    Same output as create_tile_matrix / gen_mat_pair / gen_mat_single
    (snapatac2-core/src/feature_count/data_iter.rs:312-415) on the same input.
    """
    torch = _lib.require_gpu()
    L = _lib.lib()
    if counting_strategy not in _STRATEGY:
        # counter.rs:25-36
        raise RuntimeError("Counting strategy must be one of 'insertion', 'fragment', or "
                           "'paired-insertion'")
    strategy = _STRATEGY[counting_strategy] if frags.paired else _SINGLE
    if bin_size < 1:
        raise ValueError("bin_size must be >= 1")
    names = list(chrom_sizes.keys())
    lens = np.asarray([chrom_sizes[c] for c in names], dtype=np.uint64)
    # FragmentData::exclude keeps only names present in the index (data_iter.rs:187-194)
    excl = set(exclude_chroms or []) & set(names)
    excluded = np.asarray([c in excl for c in names], dtype=np.uint8)
    n_chrom = len(names)
    table_h = np.zeros(L.snapgpu_tile_table_bytes(n_chrom), dtype=np.uint8)
    n_bins = C.c_uint64(0)
    _lib.check(L.snapgpu_tile_table_pack(lens.ctypes.data, excluded.ctypes.data, n_chrom,
                                         int(bin_size), table_h.ctypes.data, C.byref(n_bins)))
    n_bins = int(n_bins.value)
    dev = frags.device
    table = torch.from_numpy(table_h).to(dev)
    n = frags.shape[0]
    min_len = 0 if min_frag_size is None else int(min_frag_size)
    max_len = 0xFFFFFFFF if max_frag_size is None else int(max_frag_size)
    if min_len < 0 or max_len < 0:
        raise OverflowError("can't convert negative int to unsigned")
    min_len, max_len = min(min_len, 0xFFFFFFFF), min(max_len, 0xFFFFFFFF)
    if min_frag_size is not None and min_frag_size > 0xFFFFFFFF:
        min_len = 0xFFFFFFFF  # nothing can pass; lengths are uint32
        max_len = 0
    indptr = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    status = torch.zeros(8, dtype=torch.int32, device=dev)
    stream = _lib.stream_ptr()
    lens_ptr = _lib.ptr(frags.values) if frags.paired else 0
    cell_ptr = frags.cell_ptr
    level = C.c_int32(0)
    row_nnz_ptr = indptr.data_ptr() + 8
    order = _cell_order(torch, L, cell_ptr, n, stream)
    frags = _tile_count(L, frags, table, n_chrom, bin_size, strategy, min_len, max_len, row_nnz_ptr, order,
                        status, level, stream)
    lens_ptr = _lib.ptr(frags.values) if frags.paired else 0
    ws_bytes = int(L.snapgpu_scan_workspace_bytes(n))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    _lib.check(L.snapgpu_scan_inclusive_i64(row_nnz_ptr, n, ws.data_ptr(), ws_bytes, stream))
    nnz = int(indptr[-1].item()) if n > 0 else 0
    idx_bits = 32 if n_bins < 2**31 else 64
    out_idx = torch.empty(nnz, dtype=torch.int32 if idx_bits == 32 else torch.int64, device=dev)
    out_val = torch.empty(nnz, dtype=torch.int32, device=dev)
    _lib.check(L.snapgpu_tile_fill(cell_ptr.data_ptr(), _lib.ptr(frags.pos), lens_ptr, n,
                                   table.data_ptr(), n_chrom, int(bin_size), strategy,
                                   min_len, max_len, level.value, indptr.data_ptr(),
                                   out_idx.data_ptr(), idx_bits, out_val.data_ptr(),
                                   _lib.ptr(order), status.data_ptr(), stream))
    kept = {c: chrom_sizes[c] for c in names if c not in excl}
    return DeviceCSR(indptr, out_idx, out_val, (n, n_bins)), BinIndex(kept, bin_size)


_TILE_STREAMS = 4


def _cell_order(torch, L, cell_ptr, n: int, stream):
    """Cells by decreasing fragment count (snapgpu_tile_order), or None for no cells."""
    if n == 0:
        return None
    ws_b = int(L.snapgpu_tile_order_workspace_bytes(n))
    ws = torch.empty(ws_b, dtype=torch.uint8, device=cell_ptr.device)
    order = torch.empty(n, dtype=torch.int32, device=cell_ptr.device)
    _lib.check(L.snapgpu_tile_order(cell_ptr.data_ptr(), n, order.data_ptr(), ws.data_ptr(), ws_b, stream))
    return order


def make_tile_matrix_chunked(chunks, chrom_sizes: dict[str, int], bin_size: int = 500,
                             exclude_chroms: list[str] | None = None,
                             min_frag_size: int | None = None, max_frag_size: int | None = None,
                             counting_strategy: str = "paired-insertion") -> tuple[DeviceCSR, BinIndex]:
    """Tile matrix of fragments given as consecutive chunks of whole cells
    [(first_cell, DeviceFragments), ...] written into one CSR: a count pass over
    every chunk, one scan, then a fill pass per chunk at the global offsets."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    names = list(chrom_sizes.keys())
    lens = np.asarray([chrom_sizes[c] for c in names], dtype=np.uint64)
    excl = set(exclude_chroms or []) & set(names)
    excluded = np.asarray([c in excl for c in names], dtype=np.uint8)
    n_chrom = len(names)
    table_h = np.zeros(L.snapgpu_tile_table_bytes(n_chrom), dtype=np.uint8)
    nb = C.c_uint64(0)
    _lib.check(L.snapgpu_tile_table_pack(lens.ctypes.data, excluded.ctypes.data, n_chrom,
                                         int(bin_size), table_h.ctypes.data, C.byref(nb)))
    n_bins = int(nb.value)
    dev = chunks[0][1].device
    table = torch.from_numpy(table_h).to(dev)
    n = sum(f.shape[0] for _, f in chunks)
    strategy = _STRATEGY[counting_strategy] if chunks[0][1].paired else _SINGLE
    min_len = 0 if min_frag_size is None else int(min_frag_size)
    max_len = 0xFFFFFFFF if max_frag_size is None else int(max_frag_size)
    indptr = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    main = torch.cuda.current_stream()
    # Chunks are counted and filled by a few host threads, each on its own
    # stream: every library call waits for its own launch (to read the kernel's
    # status), so one thread would leave the GPU idle behind the longest cells
    # of each chunk; several streams overlap those tails.
    n_streams = min(_TILE_STREAMS, len(chunks))
    streams = [torch.cuda.Stream(device=dev) for _ in range(n_streams)]
    statuses = [torch.zeros(8, dtype=torch.int32, device=dev) for _ in range(n_streams)]
    for st_ in streams:
        st_.wait_stream(main)
    levels = [0] * len(chunks)

    def run(fn):
        errors = []

        def worker(k):
            try:
                with torch.cuda.stream(streams[k]):
                    for ci in range(k, len(chunks), n_streams):
                        fn(ci, statuses[k], _lib.stream_ptr())
            except BaseException as exc:  # re-raised in the calling thread
                errors.append(exc)

        threads = [threading.Thread(target=worker, args=(k,)) for k in range(n_streams)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        for st_ in streams:
            main.wait_stream(st_)

    orders = [None] * len(chunks)

    frs = [f for _, f in chunks]  # a chunk with a row out of order is replaced by a sorted copy

    def count(ci, status, stream):
        c0, fr = chunks[ci]
        level = C.c_int32(0)
        orders[ci] = _cell_order(torch, L, fr.cell_ptr, fr.shape[0], stream)
        frs[ci] = _tile_count(L, fr, table, n_chrom, bin_size, strategy, min_len, max_len,
                              indptr.data_ptr() + 8 * (c0 + 1), orders[ci], status, level, stream)
        levels[ci] = level.value

    run(count)
    stream = _lib.stream_ptr()
    ws_b = int(L.snapgpu_scan_workspace_bytes(n))
    ws = torch.empty(max(ws_b, 8), dtype=torch.uint8, device=dev)
    _lib.check(L.snapgpu_scan_inclusive_i64(indptr.data_ptr() + 8, n, ws.data_ptr(), ws_b, stream))
    nnz = int(indptr[-1].item())
    idx_bits = 32 if n_bins < 2**31 else 64
    out_idx = torch.empty(nnz, dtype=torch.int32 if idx_bits == 32 else torch.int64, device=dev)
    out_val = torch.empty(nnz, dtype=torch.int32, device=dev)
    for st_ in streams:
        st_.wait_stream(main)

    def fill(ci, status, stream):
        c0 = chunks[ci][0]
        fr = frs[ci]
        _lib.check(L.snapgpu_tile_fill(fr.cell_ptr.data_ptr(), fr.pos.data_ptr(),
                                       _lib.ptr(fr.values) if fr.paired else 0, fr.shape[0],
                                       table.data_ptr(), n_chrom, int(bin_size), strategy,
                                       min_len, max_len, levels[ci], indptr.data_ptr() + 8 * c0,
                                       out_idx.data_ptr(), idx_bits, out_val.data_ptr(),
                                       _lib.ptr(orders[ci]), status.data_ptr(), stream))

    run(fill)
    kept = {c: chrom_sizes[c] for c in names if c not in excl}
    return DeviceCSR(indptr, out_idx, out_val, (n, n_bins)), BinIndex(kept, bin_size)


def add_tile_matrix(
    adata,
    *,
    bin_size: int = 500,
    inplace: bool = True,
    chunk_size: int = 500,
    exclude_chroms: list[str] | str | None = ["chrM", "chrY", "M", "Y"],
    min_frag_size: int | None = None,
    max_frag_size: int | None = None,
    counting_strategy: Literal["fragment", "insertion", "paired-insertion"] = "paired-insertion",
    value_type: Literal["target", "total", "fraction"] = "target",
    summary_type: Literal["sum", "mean"] = "sum",
    file=None,
    backend: Literal["hdf5"] = "hdf5",
    n_jobs: int = 8,
):
    """Generate cell by bin count matrix (drop-in for snapatac2.pp.add_tile_matrix).

    The counts are computed on the GPU; `adata.X` becomes a DeviceCSR when
    `adata` is a snapatac2_amd.AnnData (call `.to_scipy()` for a host copy) and
    a scipy.sparse.csr_matrix (uint32) otherwise. `chunk_size` and `n_jobs` do
    not change the result and are accepted for compatibility.
    """
    if isinstance(exclude_chroms, str):
        exclude_chroms = [exclude_chroms]

    def fun(data, out):
        # Argument checks in the order mk_tile_matrix evaluates them
        # (preprocessing.rs:303-345): strategy, value type, summary type.
        if counting_strategy not in _STRATEGY:
            raise RuntimeError("Counting strategy must be one of 'insertion', 'fragment', "
                               "or 'paired-insertion'")
        if value_type not in _VALUE_TYPES:
            raise _lib.PanicException("Value type must be one of 'target', 'total', or 'fraction'")
        if summary_type not in _SUMMARY_TYPES:
            raise _lib.PanicException("Summary type must be one of 'sum', 'mean', or 'count'")
        obsm = data.obsm
        if FRAGMENT_SINGLE in obsm:          # single-end takes precedence (mod.rs:80-86)
            frags, paired = obsm[FRAGMENT_SINGLE], False
        elif FRAGMENT_PAIRED in obsm:
            frags, paired = obsm[FRAGMENT_PAIRED], True
        elif "_values" in obsm or "__values__" in obsm:
            raise NotImplementedError("base-value (methylation) matrices are not supported")
        else:
            raise RuntimeError("one of the following keys must be present in the '.obsm': "
                               f"'{FRAGMENT_SINGLE}', '{FRAGMENT_PAIRED}'")
        frags = _as_device_fragments(frags, paired)
        chrom_sizes = read_chrom_sizes(data.uns)
        X, var_names = make_tile_matrix(frags, chrom_sizes, bin_size, exclude_chroms,
                                        min_frag_size, max_frag_size, counting_strategy)
        target = data if out is None else out
        if not isinstance(target, AnnData):
            X = X.to_scipy()
        _set_x(target, X, var_names)

    if inplace:
        if isinstance(adata, list):
            for a in adata:
                fun(a, None)
        else:
            fun(adata, None)
        return None
    if file is not None:
        raise NotImplementedError("backed (file=) output is not supported; use file=None")
    out = AnnData(obs=adata.obs.copy() if hasattr(adata.obs, "copy") else adata.obs,
                  n_obs=adata.n_obs, n_vars=0)
    fun(adata, out)
    return out


def _set_x(adata, X, var_names) -> None:
    if isinstance(adata, AnnData):
        adata.set_n_vars(X.shape[1], var_names)
        adata.X = X
        return
    import pandas as pd  # a real anndata.AnnData
    adata.X = X
    adata.var_names = pd.Index(var_names.to_numpy())


def _column_sums(X) -> np.ndarray:
    """Column sums of the count matrix as float64 (exact: integer sums)."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    if not isinstance(X, DeviceCSR):
        import scipy.sparse as sp
        Xc = sp.csr_matrix(X)
        if Xc.dtype.kind not in "ui" or Xc.data.size and Xc.data.min() < 0:
            raise TypeError("select_features expects a non-negative integer count matrix")
        dev = torch.device("cuda")
        X = DeviceCSR(torch.from_numpy(Xc.indptr.astype(np.int64)).to(dev),
                      torch.from_numpy(Xc.indices.astype(np.int32 if Xc.shape[1] < 2**31 else np.int64)).to(dev),
                      torch.from_numpy(Xc.data.astype(np.uint32).view(np.int32)).to(dev), Xc.shape)
    out = torch.empty(X.shape[1], dtype=torch.int64, device=X.device)
    ws = torch.empty(4, dtype=torch.int32, device=X.device)
    bits = 32 if X.indices.dtype == torch.int32 else 64
    _lib.check(L.snapgpu_csr_col_sum_rows_u32(X.indptr.data_ptr(), X.shape[0], X.indices.data_ptr(), bits,
                                              X.data.data_ptr(), X.shape[1], out.data_ptr(),
                                              ws.data_ptr(), _lib.stream_ptr()))
    return out.cpu().numpy().astype(np.float64)


def _find_most_accessible_features(feature_count, filter_lower_quantile, filter_upper_quantile,
                                   total_features) -> np.ndarray:
    """Same steps as _basic.py:952-967 (numpy's default argsort, so equal counts
    are ordered exactly as the reference orders them on the same host)."""
    idx = np.argsort(feature_count)
    # first position with a positive count, as the reference's loop finds it (the
    # last position when none is positive; with no features `i` stays unbound there
    # and here alike)
    if idx.size:
        pos = np.asarray(feature_count)[idx] > 0
        i = int(np.argmax(pos)) if pos.any() else idx.size - 1
    idx = idx[i:]
    n = idx.size
    n_lower = int(filter_lower_quantile * n)
    n_upper = int(filter_upper_quantile * n)
    idx = idx[n_lower:n - n_upper]
    return idx[::-1][:total_features]


def select_features(
    adata,
    n_features: int = 500000,
    filter_lower_quantile: float = 0.005,
    filter_upper_quantile: float = 0.005,
    whitelist=None,
    blacklist=None,
    max_iter: int = 1,
    inplace: bool = True,
    n_jobs: int = 8,
    verbose: bool = True,
):
    """Select the most accessible features (drop-in for snapatac2.pp.select_features).

    Column sums run on the GPU; the selection itself is the reference's numpy
    code. Whitelist and blacklist BED files are tested against var_names with
    the peak-matrix overlap kernel (intersect_bed, utils.rs:320-336), applied as
    in _basic.py:1062-1087: blacklisted features leave the selection, then
    non-zero whitelisted features join it. max_iter > 1 (iterative clustering,
    which needs Leiden) is not supported.
    """
    if isinstance(adata, list):
        res = [select_features(a, n_features, filter_lower_quantile, filter_upper_quantile,
                               whitelist, blacklist, max_iter, inplace, verbose=False) for a in adata]
        return None if inplace else res
    if max_iter > 1:
        raise NotImplementedError("iterative feature selection (max_iter > 1) is not supported")
    count = _column_sums(adata.X)
    if inplace:
        adata.var["count"] = count
    selected = _find_most_accessible_features(count, filter_lower_quantile,
                                              filter_upper_quantile, n_features)
    if blacklist is not None:
        black = intersect_bed(adata.var_names, str(blacklist))
        selected = selected[np.logical_not(black[selected])]
    result = np.zeros(adata.shape[1], dtype=bool)
    result[selected] = True
    if whitelist is not None:
        white = intersect_bed(adata.var_names, str(whitelist))
        result |= white & (count != 0)
    if verbose:
        logging.info(f"Selected {result.sum()} features.")
    if inplace:
        adata.var["selected"] = result
        return None
    return result
