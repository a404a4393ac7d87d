"""ORACLE (test infrastructure only) for path A, the tile matrix.

tile_matrix():  ctypes front end of tile_oracle.cpp, the C++ restatement of
                gen_mat_pair / gen_mat_single (data_iter.rs:312-415) and
                GenomeBaseIndex (genome.rs:286-487). OpenMP over rows.
tile_matrix_bruteforce(): the same semantics written as plain Python dict
                loops, for tiny inputs; an independent second restatement.
genome_index_*: GenomeBaseIndex arithmetic in Python for the genome.rs test
                vectors.

Pinned by tests/test_tile_oracle.py: genome.rs:495-597 known answers, hand
derived cases per strategy, brute force vs C++, and the reference's
test_tile_matrix / test_exclude invariants.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import LIB_PATH, build

STRATEGY = {"insertion": 0, "fragment": 1, "paired-insertion": 2, "single": 3}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        h.oracle_tile_run.restype = C.c_void_p
        h.oracle_tile_run.argtypes = [C.c_void_p] * 3 + [C.c_int64, C.c_void_p, C.c_void_p,
                                                          C.c_int32, C.c_uint64, C.c_int32,
                                                          C.c_int64, C.c_int64]
        h.oracle_tile_error.restype = C.c_char_p
        h.oracle_tile_error.argtypes = [C.c_void_p]
        h.oracle_tile_nnz.restype = C.c_int64
        h.oracle_tile_nnz.argtypes = [C.c_void_p]
        h.oracle_tile_n_bins.restype = C.c_uint64
        h.oracle_tile_n_bins.argtypes = [C.c_void_p]
        h.oracle_tile_copy.argtypes = [C.c_void_p] * 4
        h.oracle_tile_free.argtypes = [C.c_void_p]
        _lib = h
    return _lib


def tile_matrix(cell_ptr, pos, length, chrom_sizes: dict, bin_size=500, exclude_chroms=None,
                min_len=None, max_len=None, strategy="paired-insertion") -> sp.csr_matrix:
    """Cell x bin uint32 CSR. Raises RuntimeError where the reference panics."""
    L = lib()
    cell_ptr = np.ascontiguousarray(cell_ptr, dtype=np.uint64)
    pos = np.ascontiguousarray(pos, dtype=np.uint32)
    single = strategy == "single"
    ln = None if single else np.ascontiguousarray(length, dtype=np.uint32)
    names = list(chrom_sizes)
    lens = np.asarray([chrom_sizes[c] for c in names], dtype=np.uint64)
    excl = set(exclude_chroms or [])
    excluded = np.asarray([c in excl for c in names], dtype=np.uint8)
    n = len(cell_ptr) - 1
    h = L.oracle_tile_run(cell_ptr.ctypes.data, pos.ctypes.data,
                          None if single else ln.ctypes.data, n, lens.ctypes.data,
                          excluded.ctypes.data, len(names), int(bin_size), STRATEGY[strategy],
                          -1 if min_len is None else int(min_len),
                          -1 if max_len is None else int(max_len))
    try:
        err = L.oracle_tile_error(h).decode()
        if err:
            raise RuntimeError(err)
        nnz = L.oracle_tile_nnz(h)
        nb = L.oracle_tile_n_bins(h)
        indptr = np.empty(n + 1, dtype=np.int64)
        idx = np.empty(nnz, dtype=np.int64)
        val = np.empty(nnz, dtype=np.uint32)
        L.oracle_tile_copy(h, indptr.ctypes.data, idx.ctypes.data, val.ctypes.data)
    finally:
        L.oracle_tile_free(h)
    return sp.csr_matrix((val, idx, indptr), shape=(n, nb))


# ---------------------------------------------------------------------------
# Pure-Python restatement (tiny inputs only)
# ---------------------------------------------------------------------------

def genome_index(chrom_sizes: dict, step: int = 1):
    lens = list(chrom_sizes.values())
    base = np.cumsum(lens).tolist()
    binned = np.cumsum([-(-l // step) for l in lens]).tolist()
    return list(chrom_sizes), base, binned


def get_position(chrom_sizes: dict, i: int):
    """genome.rs:419-434 (step 1)."""
    names, base, _ = genome_index(chrom_sizes)
    j = int(np.searchsorted(base, i, side="left"))
    if j < len(base) and base[j] == i:
        return names[j + 1], 0
    prev = 0 if j == 0 else base[j - 1]
    return names[j], i - prev


def get_position_rev(chrom_sizes: dict, step: int, chrom: str, pos: int) -> int:
    """genome.rs:387-406."""
    names, base, binned = genome_index(chrom_sizes, step)
    i = names.index(chrom)
    size = base[i] - (base[i - 1] if i else 0)
    if pos >= size:
        raise RuntimeError(f"Position {pos} is out of range for chromosome {chrom}")
    return (binned[i - 1] if i else 0) + pos // step


def get_region(chrom_sizes: dict, step: int, i: int):
    """genome.rs:437-466."""
    names, base, binned = genome_index(chrom_sizes, step)
    j = int(np.searchsorted(binned, i, side="left"))
    if j < len(binned) and binned[j] == i:
        j += 1
        size = base[j] - base[j - 1]
        return names[j], 0, min(step, size)
    size = base[j] - (base[j - 1] if j else 0)
    prev = binned[j - 1] if j else 0
    start = (i - prev) * step
    return names[j], start, min(start + step, size)


def tile_matrix_bruteforce(cell_ptr, pos, length, chrom_sizes: dict, bin_size=500,
                           exclude_chroms=None, min_len=None, max_len=None,
                           strategy="paired-insertion") -> sp.csr_matrix:
    excl = set(exclude_chroms or [])
    kept = {c: l for c, l in chrom_sizes.items() if c not in excl}
    rows = []
    for r in range(len(cell_ptr) - 1):
        count: dict[int, int] = {}
        for k in range(int(cell_ptr[r]), int(cell_ptr[r + 1])):
            chrom, start = get_position(chrom_sizes, int(pos[k]))
            if chrom in excl:
                continue
            if strategy == "single":
                b = get_position_rev(kept, bin_size, chrom, start)
                count[b] = count.get(b, 0) + 1
                continue
            L = int(length[k])
            if (min_len is not None and L < min_len) or (max_len is not None and L > max_len):
                continue
            b1 = get_position_rev(kept, bin_size, chrom, start)
            b2 = get_position_rev(kept, bin_size, chrom, start + L - 1)
            if strategy == "insertion":
                bins = [b1, b2]
            elif strategy == "fragment":
                bins = list(range(b1, b2 + 1))
            else:
                bins = [b1] if b1 == b2 else [b1, b2]
            for b in bins:
                count[b] = count.get(b, 0) + 1
        rows.append(sorted(count.items()))
    indptr = np.zeros(len(rows) + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(x) for x in rows])
    idx = np.asarray([b for row in rows for b, _ in row], dtype=np.int64)
    val = np.asarray([c for row in rows for _, c in row], dtype=np.uint32)
    nb = genome_index(kept, bin_size)[2][-1] if kept else 0
    return sp.csr_matrix((val, idx, indptr), shape=(len(rows), nb))
