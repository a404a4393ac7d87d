"""Drop-in for snapatac2.pp.make_peak_matrix (cell x peak counts), on the GPU.

Reference: snapatac2-python/python/snapatac2/preprocessing/_basic.py:577-704
(Python API), internal.mk_peak_matrix (snapatac2-python/src/preprocessing.rs:370-420),
create_peak_matrix (snapatac2-core/src/feature_count/matrix.rs:95-171),
RegionCounter (feature_count/counter.rs:73-159). The peak list is parsed and
indexed on the host (a sort of the regions); libsnapgpu counts the overlaps
of every fragment and builds the CSR.
"""
from __future__ import annotations

import ctypes as C
import gzip
from pathlib import Path

import numpy as np

from .. import _lib
from .._device import DeviceCSR
from ..anndata import AnnData
from ..genome import read_chrom_sizes
from ._basic import FRAGMENT_PAIRED, FRAGMENT_SINGLE, _STRATEGY, _SUMMARY_TYPES, _VALUE_TYPES, \
    _as_device_fragments


def parse_regions(peaks) -> tuple[list[str], list[tuple[str, int, int]]]:
    """GenomicRange::from_str of each "chrom:start-end" (bed_utils), the distinct
    regions in first-occurrence order (the regions form an index set,
    matrix.rs:122). Returns (names, [(chrom, start, end)])."""
    seen, names, regs = {}, [], []
    for p in peaks:
        s = str(p).strip()
        try:
            chrom, rng = s.rsplit(":", 1)
            a, b = rng.split("-", 1)
            r = (chrom, int(a), int(b))
        except ValueError:
            raise _lib.PanicException(f"cannot parse genomic range: {s!r}") from None
        if r[1] < 0 or r[2] < 0:
            raise _lib.PanicException(f"cannot parse genomic range: {s!r}")
        if r not in seen:
            seen[r] = len(regs)
            names.append(f"{r[0]}:{r[1]}-{r[2]}")
            regs.append(r)
    return names, regs


class PeakIndex:
    """The region index of snapgpu_peak_index on the device."""

    def __init__(self, regs, chrom_sizes: dict[str, int], torch):
        dev = torch.device("cuda")
        chroms = list(chrom_sizes)
        cid = {c: i for i, c in enumerate(chroms)}
        rc = np.fromiter((cid.get(r[0], -1) for r in regs), dtype=np.int64, count=len(regs))
        rs = np.fromiter((r[1] for r in regs), dtype=np.int64, count=len(regs))
        re_ = np.fromiter((r[2] for r in regs), dtype=np.int64, count=len(regs))
        ri = np.arange(len(regs), dtype=np.int64)
        on = rc >= 0
        rc, rs, re_, ri = rc[on], rs[on], re_[on], ri[on]
        order = np.lexsort((ri, rs, rc))
        arr = np.stack([rc[order], rs[order], re_[order], ri[order]], axis=1).reshape(-1, 4)
        starts = arr[:, 1].astype(np.uint64)
        ends = arr[:, 2].astype(np.uint64)
        pmax = np.zeros_like(ends)
        off = np.searchsorted(arr[:, 0], np.arange(len(chroms) + 1), side="left").astype(np.int64)
        for c in range(len(chroms)):
            a, b = off[c], off[c + 1]
            if b > a:
                pmax[a:b] = np.maximum.accumulate(ends[a:b])
        base = np.concatenate([[0], np.cumsum([int(chrom_sizes[c]) for c in chroms])]).astype(np.uint64)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int64 if a.dtype.itemsize == 8 else np.int32)).to(dev)
        self._keep = [t(starts if starts.size else np.zeros(1, np.uint64)),
                      t(ends if ends.size else np.zeros(1, np.uint64)),
                      t(pmax if pmax.size else np.zeros(1, np.uint64)),
                      t(arr[:, 3].astype(np.uint32) if arr.size else np.zeros(1, np.uint32)),
                      t(off), t(base)]
        self.c = _lib.PeakIndex(*[x.data_ptr() for x in self._keep], len(chroms), len(regs))


def make_peak_counts(frags, chrom_sizes: dict[str, int], peaks, counting_strategy: str = "paired-insertion",
                     min_frag_size: int | None = None, max_frag_size: int | None = None):
    """(DeviceCSR cells x regions with uint32 counts, region names)."""
    if counting_strategy not in _STRATEGY:
        raise RuntimeError("Counting strategy must be one of 'insertion', 'fragment', or "
                           "'paired-insertion'")  # counter.rs:25-36
    names, regs = parse_regions(peaks)
    return region_counts(frags, chrom_sizes, regs, counting_strategy, min_frag_size, max_frag_size), names


def region_counts(frags, chrom_sizes: dict[str, int], regs, counting_strategy: str = "paired-insertion",
                  min_frag_size: int | None = None, max_frag_size: int | None = None) -> DeviceCSR:
    """DeviceCSR cells x len(regs) with uint32 counts; regs = distinct
    (chrom, start, end) tuples, column j = regs[j]."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    st = _lib.stream_ptr()
    dev = torch.device("cuda")
    if counting_strategy not in _STRATEGY:
        raise RuntimeError("Counting strategy must be one of 'insertion', 'fragment', or "
                           "'paired-insertion'")  # counter.rs:25-36
    idx = PeakIndex(regs, chrom_sizes, torch)
    n_cells = frags.shape[0]
    n_frag = frags.n_fragments
    paired = 1 if frags.paired else 0
    strat = _STRATEGY[counting_strategy]
    lo = 0 if min_frag_size is None else int(min_frag_size)
    hi = 0xFFFFFFFF if max_frag_size is None else int(min(max_frag_size, 0xFFFFFFFF))
    indptr = torch.zeros(n_cells + 1, dtype=torch.int64, device=dev)
    if n_frag == 0 or n_cells == 0 or not regs:
        return DeviceCSR(indptr, torch.empty(0, dtype=torch.int32, device=dev),
                         torch.empty(0, dtype=torch.int32, device=dev), (n_cells, len(regs)))
    hits = torch.empty(n_frag, dtype=torch.int64, device=dev)
    wsb = int(L.snapgpu_peak_count_workspace_bytes(n_frag))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    total = C.c_int64(0)
    _lib.check(L.snapgpu_peak_count(C.byref(idx.c), frags.pos.data_ptr(), frags.values.data_ptr(), n_frag, paired,
                                    strat, lo, hi, hits.data_ptr(), ws.data_ptr(), wsb, C.byref(total), st))
    del ws
    T = total.value
    keys = torch.empty(max(T, 1), dtype=torch.int64, device=dev)
    wsb = int(L.snapgpu_peak_pairs_workspace_bytes(T))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    nnz = C.c_int64(0)
    _lib.check(L.snapgpu_peak_pairs(C.byref(idx.c), frags.cell_ptr.data_ptr(), n_cells, frags.pos.data_ptr(),
                                    frags.values.data_ptr(), n_frag, paired, strat, lo, hi, hits.data_ptr(), T,
                                    keys.data_ptr(), ws.data_ptr(), wsb, C.byref(nnz), st))
    del hits, keys
    out_idx = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=dev)
    out_val = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=dev)
    _lib.check(L.snapgpu_peak_csr(ws.data_ptr(), wsb, T, nnz.value, n_cells, indptr.data_ptr(), out_idx.data_ptr(),
                                  out_val.data_ptr(), st))
    return DeviceCSR(indptr, out_idx[:nnz.value], out_val[:nnz.value], (n_cells, len(regs)))


def make_peak_matrix(
    adata,
    *,
    use_rep: str | list[str] | None = None,
    inplace: bool = False,
    file: Path | None = None,
    backend="hdf5",
    peak_file: Path | None = None,
    chunk_size: int = 500,
    use_x: bool = False,
    min_frag_size: int | None = None,
    max_frag_size: int | None = None,
    counting_strategy: str = "paired-insertion",
    value_type: str = "target",
    summary_type: str = "sum",
):
    """Cell by peak count matrix (snapatac2.pp.make_peak_matrix, _basic.py:577-704).

    Same arguments and result for fragment data: X (cells x peaks, uint32
    counts, on the device) with var_names = the distinct peaks. `use_x` (counts
    from .X) and backed output (`file`) are not provided here.
    """
    if peak_file is not None and use_rep is not None:
        raise RuntimeError("'peak_file' and 'use_rep' cannot be both set")
    if use_rep is None and peak_file is None:
        use_rep = "peaks"
    if isinstance(use_rep, str):
        df = adata.uns[use_rep]
        peaks = df[df.columns[0]]
    else:
        peaks = use_rep
    if peak_file is not None:
        opener = gzip.open if Path(peak_file).suffix == ".gz" else open
        with opener(peak_file, "rt") as f:
            peaks = [line.strip() for line in f]
    if value_type not in _VALUE_TYPES or summary_type not in _SUMMARY_TYPES:
        raise _lib.PanicException("invalid value_type or summary_type")  # preprocessing.rs:347-363
    if use_x:
        raise NotImplementedError("use_x=True (counts from .X) is not provided")
    if file is not None:
        raise NotImplementedError("backed AnnData output needs HDF5, which is not available here")
    if FRAGMENT_SINGLE in adata.obsm:
        frags = _as_device_fragments(adata.obsm[FRAGMENT_SINGLE], False)
    elif FRAGMENT_PAIRED in adata.obsm:
        frags = _as_device_fragments(adata.obsm[FRAGMENT_PAIRED], True)
    else:
        raise RuntimeError("No fragment data found in the anndata object")  # matrix.rs:154
    X, names = make_peak_counts(frags, read_chrom_sizes(adata.uns), list(peaks), counting_strategy,
                                min_frag_size, max_frag_size)
    if inplace:
        if isinstance(adata, AnnData):
            adata.set_n_vars(len(names), names)
            adata.X = X
        else:  # a real anndata.AnnData
            import pandas as pd
            adata.X = X.to_scipy()
            adata.var_names = pd.Index(names)
        return None
    out = AnnData(X=X, obs=adata.obs.copy() if hasattr(adata, "obs") else None)
    out.set_n_vars(len(names), names)
    return out
