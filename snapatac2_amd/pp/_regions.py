"""Region-name overlap tests of select_features' whitelist and blacklist.

Reference: internal.intersect_bed (snapatac2-python/src/utils.rs:320-336),
called by select_features (snapatac2-python/python/snapatac2/preprocessing/
_basic.py:1062-1087): the BED file's first three columns (chrom, start, end,
half-open) form an interval map, and each var_name "chrom:start-end"
(GenomicRange::from_str) is True when it shares a base with one of them on
the same chromosome.

The test runs on the GPU through the peak-matrix overlap kernel
(snapgpu_peak_count): the BED regions become the region index and every
var_name one query fragment [start, end), whose overlap count is > 0 exactly
when the reference's is_overlapped is true.
"""
from __future__ import annotations

import ctypes as C
import gzip
import re
from pathlib import Path

import numpy as np

from .. import _lib

_RANGE = re.compile(r"^(.+):(\d+)-(\d+)$")
_U32 = (1 << 32) - 1


def read_bed3(path) -> list[tuple[str, int, int]]:
    """(chrom, start, end) of every line of a BED file (plain or gzip). Like the
    reference's BED<3> reader, a line that does not parse is an error."""
    p = Path(path)
    with open(p, "rb") as f:
        gz = f.read(2) == b"\x1f\x8b"
    opener = gzip.open if gz else open
    out = []
    with opener(p, "rt") as f:
        for ln, line in enumerate(f, 1):
            line = line.rstrip("\r\n")
            if not line:
                continue
            fields = line.split("\t")
            try:
                if len(fields) < 3:
                    raise ValueError
                s, e = int(fields[1]), int(fields[2])
                if s < 0 or e < 0:
                    raise ValueError
            except ValueError:
                raise _lib.PanicException(f"{p}: line {ln} is not a BED record: {line!r}") from None
            out.append((fields[0], s, e))
    return out


def parse_range_names(names) -> tuple[list[str], np.ndarray, np.ndarray, np.ndarray]:
    """(chromosome names, chromosome code of each name, start, end) of
    "chrom:start-end" names (GenomicRange::from_str); anything else is an error.
    A tile matrix's lazy BinIndex gives its bins without formatting strings."""
    from ..genome import BinIndex
    if isinstance(names, BinIndex):
        return names.ranges()
    names = [str(x) for x in names]
    n = len(names)
    codes: dict[str, int] = {}
    code = np.empty(n, dtype=np.int32)
    start = np.empty(n, dtype=np.int64)
    end = np.empty(n, dtype=np.int64)
    for i, s in enumerate(names):
        m = _RANGE.match(s.strip())
        if m is None:
            raise _lib.PanicException(f"cannot parse genomic range: {s!r}")
        code[i] = codes.setdefault(m.group(1), len(codes))
        start[i] = int(m.group(2))
        end[i] = int(m.group(3))
    return list(codes), code, start, end


def intersect_bed(names, bed_file) -> np.ndarray:
    """Boolean array: names[i] overlaps a region of bed_file (intersect_bed)."""
    chroms, code, start, end = parse_range_names(names)
    regs = read_bed3(bed_file)
    return overlaps(chroms, code, start, end, regs)


def overlaps(chroms, code, start, end, regs) -> np.ndarray:
    """Query intervals [start[i], end[i]) on chromosome chroms[code[i]] against
    the regions `regs` ((chrom, start, end) tuples): True where they share a base."""
    from ._peaks import PeakIndex
    n = len(start)
    hit = np.zeros(n, dtype=bool)
    if n == 0 or not regs:
        return hit
    torch = _lib.require_gpu()
    L = _lib.lib()
    st = _lib.stream_ptr()
    dev = torch.device("cuda")
    # one coordinate system per batch of chromosomes: each chromosome spans the
    # largest coordinate seen on it (+1), and a batch keeps its genome below 2^32
    span: dict[str, int] = {}
    if n:
        top = np.zeros(len(chroms), dtype=np.int64)
        np.maximum.at(top, code, np.maximum(start, end) + 1)
        for c, v in zip(chroms, top.tolist()):
            span[c] = v
    for c, s, e in regs:
        span[c] = max(span.get(c, 0), s + 1, e + 1)
    if any(v > _U32 for v in span.values()):
        raise ValueError("a region ends beyond 2^32 - 1")
    batches, cur, tot = [], {}, 0
    for c in sorted(span):
        if cur and tot + span[c] > _U32:
            batches.append(cur)
            cur, tot = {}, 0
        cur[c] = span[c]
        tot += span[c]
    batches.append(cur)
    for sizes in batches:
        base, acc = {}, 0
        for c, v in sizes.items():
            base[c] = acc
            acc += v
        cbase = np.asarray([base.get(c, -1) for c in chroms], dtype=np.int64)
        qb_all = cbase[code]
        q = np.flatnonzero(qb_all >= 0)
        breg = [r for r in regs if r[0] in base]
        if q.size == 0 or not breg:
            continue
        pos = (qb_all[q] + start[q]).astype(np.uint32)
        size = np.maximum(end[q] - start[q], 0).astype(np.uint32)
        idx = PeakIndex(breg, sizes, torch)
        d_pos = torch.from_numpy(pos.view(np.int32)).to(dev)
        d_val = torch.from_numpy(size.view(np.int32)).to(dev)
        hits = torch.empty(q.size, dtype=torch.int64, device=dev)
        wsb = int(L.snapgpu_peak_count_workspace_bytes(q.size))
        ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev)
        total = C.c_int64(0)
        # paired = 1, strategy "fragment": every region overlapping [start, end)
        _lib.check(L.snapgpu_peak_count(C.byref(idx.c), d_pos.data_ptr(), d_val.data_ptr(), q.size, 1, 1, 0,
                                        _U32, hits.data_ptr(), ws.data_ptr(), wsb, C.byref(total), st))
        # hits holds the inclusive scan of the per-query counts
        h = hits.cpu().numpy()
        hit[q] = np.diff(h, prepend=0) > 0
    return hit
