"""ORACLE (test infrastructure only) for the cell x peak matrix.

Restates, by brute force over each chromosome's regions:
  create_peak_matrix            snapatac2-core/src/feature_count/matrix.rs:95-171
                                (regions collected into an index set: distinct
                                regions, first-occurrence order, :122)
  RegionCounter::insert /       feature_count/counter.rs:92-139 (half-open
  insert_fragment               overlap on the same chromosome; paired-insertion
                                counts a region once per fragment)
  pair_to_fragments /           feature_count/data_iter.rs:27-123 (paired:
  single_to_fragments           [start, start + size) with the inclusive size
                                filter; single-end: one base at the stored position)
  get_position                  snapatac2-core/src/genome.rs:419-434

Pinned by the hand-computed cases of tests/test_peaks_oracle.py. The interval
index of bed_utils (GIntervalIndexSet) is not vendored: its overlap rule
(half-open, same chromosome) and de-duplication are restated from its use
here; parity unpinned for other readings of them.
"""
from __future__ import annotations

import bisect

import numpy as np


def parse_regions(peaks):
    seen, regs = {}, []
    for p in peaks:
        chrom, rng = str(p).strip().rsplit(":", 1)
        a, b = rng.split("-", 1)
        r = (chrom, int(a), int(b))
        if r not in seen:
            seen[r] = len(regs)
            regs.append(r)
    return regs


def peak_matrix(cell_ptr, pos, val, paired: bool, chrom_sizes: dict, peaks, strategy="paired-insertion",
                min_len=None, max_len=None):
    """(indptr, indices, data) of the cell x region count matrix."""
    regs = parse_regions(peaks)
    chroms = list(chrom_sizes)
    base = np.concatenate([[0], np.cumsum([int(chrom_sizes[c]) for c in chroms])]).tolist()
    by_chrom = {}
    for col, (c, s, e) in enumerate(regs):
        by_chrom.setdefault(c, []).append((s, e, col))

    def hits(c, qs, qe):
        return [col for s, e, col in by_chrom.get(c, []) if s < qe and qs < e]

    indptr, indices, data = [0], [], []
    for i in range(len(cell_ptr) - 1):
        counts = {}
        for f in range(int(cell_ptr[i]), int(cell_ptr[i + 1])):
            p = int(pos[f])
            if p >= base[-1]:
                continue
            k = bisect.bisect_right(base, p) - 1
            c, s = chroms[k], p - base[k]
            if not paired:
                hs = hits(c, s, s + 1)
            else:
                size = int(val[f])
                if (min_len is not None and size < min_len) or (max_len is not None and size > max_len):
                    continue
                a, b = s, s + size - 1
                if strategy == "fragment":
                    hs = hits(c, s, s + size)
                elif strategy == "insertion":
                    hs = hits(c, a, a + 1) + hits(c, b, b + 1)
                else:
                    hs = sorted(set(hits(c, a, a + 1)) | set(hits(c, b, b + 1)))
            for col in hs:
                counts[col] = counts.get(col, 0) + 1
        for col in sorted(counts):
            indices.append(col)
            data.append(counts[col])
        indptr.append(len(indices))
    return (np.asarray(indptr, dtype=np.int64), np.asarray(indices, dtype=np.int64),
            np.asarray(data, dtype=np.uint32))


def peak_matrix_cpp(cell_ptr, pos, val, paired: bool, chrom_sizes: dict, peaks, strategy="paired-insertion",
                    min_len=None, max_len=None):
    """The C++ restatement (oracle/peaks_oracle.cpp), rows in parallel: the timed
    CPU baseline. Same result as peak_matrix."""
    import ctypes as C

    from . import LIB_PATH, build
    if not LIB_PATH.exists():
        build()
    h = C.CDLL(str(LIB_PATH))
    h.oracle_peak_run.restype = C.c_int64
    h.oracle_peak_run.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int32,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_uint64,
                                  C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    regs = parse_regions(peaks)
    chroms = list(chrom_sizes)
    cid = {c: i for i, c in enumerate(chroms)}
    base = np.concatenate([[0], np.cumsum([int(chrom_sizes[c]) for c in chroms])]).astype(np.uint64)
    rc = np.asarray([cid.get(c, -1) for c, _, _ in regs], dtype=np.int32)
    rs = np.asarray([s for _, s, _ in regs], dtype=np.uint64)
    re_ = np.asarray([e for _, _, e in regs], dtype=np.uint64)
    col = np.arange(len(regs), dtype=np.uint32)
    cp = np.ascontiguousarray(cell_ptr, dtype=np.int64)
    p = np.ascontiguousarray(pos, dtype=np.uint32)
    v = np.ascontiguousarray(np.asarray(val).astype(np.int64).astype(np.uint32).view(np.int32))
    n = cp.size - 1
    strat = {"insertion": 0, "fragment": 1, "paired-insertion": 2}[strategy]
    lo = 0 if min_len is None else int(min_len)
    hi = 2**64 - 1 if max_len is None else int(max_len)
    row_len = np.zeros(n, dtype=np.int64)
    args = [cp.ctypes.data, n, p.ctypes.data, v.ctypes.data, 1 if paired else 0, base.ctypes.data, len(chroms),
            rc.ctypes.data, rs.ctypes.data, re_.ctypes.data, col.ctypes.data, len(regs), strat, lo, hi,
            row_len.ctypes.data]
    h.oracle_peak_run(*args, None, None, None)
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(row_len, out=indptr[1:])
    idx = np.zeros(int(indptr[-1]), dtype=np.int32)
    dat = np.zeros(int(indptr[-1]), dtype=np.uint32)
    h.oracle_peak_run(*args, indptr.ctypes.data, idx.ctypes.data if idx.size else None,
                      dat.ctypes.data if dat.size else None)
    return indptr, idx.astype(np.int64), dat


def intersect_bed(names, regs) -> np.ndarray:
    """intersect_bed (snapatac2-python/src/utils.rs:320-336) by brute force:
    names[i] ("chrom:start-end", GenomicRange::from_str) is True when some BED
    region (chrom, start, end) of `regs` shares a base with it on the same
    chromosome (half-open: r.start < end and start < r.end). Pinned by the hand
    cases of tests/test_regions_host.py (touching ends, containment, empty
    queries, other chromosomes)."""
    by_chrom: dict[str, list[tuple[int, int]]] = {}
    for c, s, e in regs:
        by_chrom.setdefault(c, []).append((s, e))
    arrs = {c: np.asarray(v, dtype=np.int64).reshape(-1, 2) for c, v in by_chrom.items()}
    out = np.zeros(len(names), dtype=bool)
    for i, nm in enumerate(names):
        chrom, rng = str(nm).strip().rsplit(":", 1)
        a, b = (int(x) for x in rng.split("-", 1))
        r = arrs.get(chrom)
        if r is not None:
            out[i] = bool(np.any((r[:, 0] < b) & (a < r[:, 1])))
    return out
