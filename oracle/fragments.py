"""CPU restatement of fragment-file import. Test infrastructure only: tests
use it as the checker; the product path never imports it.

Follows
  Fragment::from_str            snapatac2-core/src/preprocessing/qc.rs:306-345
  shift_fragment                snapatac2-python/src/preprocessing.rs:173-183
  import_fragments (binding)    snapatac2-python/src/preprocessing.rs:87-171
  import_fragments (core)       snapatac2-core/src/preprocessing/import.rs:31-112
  make_arraydata                import.rs:114-154
  count_fragments               import.rs:156-204
  FragmentQCBuilder             qc.rs:427-466
  GenomeBaseIndex::get_position_rev  snapatac2-core/src/genome.rs:387-406
  qc_to_df                      import.rs:206-223

Choices where the reference leaves the behaviour to a dependency or to a
panic message (the GPU path makes the same ones):
  - lines are split at '\\n'; a trailing '\\r' is removed; empty lines are
    skipped (bed_utils' reader is not vendored: parity unpinned for them);
  - numbers are ASCII digit strings (lexical::parse of u64 / u32);
  - a record whose (shifted) end is below its start is an error (the
    reference panics in paired mode and stores a wrapped value otherwise);
  - unsorted input mixing records with and without a strand is an error
    (the reference takes the mode from the first record of its external
    sort, whose tie order is not pinned).
Every error is raised as OracleError(code, line) with the 1-based line.

Parity pinned by: hand-computed known answers (tests/test_import_oracle.py),
the round-trip property of the reference's test_import
(snapatac2-python/tests/test_tools.py:128-142), and the exclusion invariant of
test_exclude (snapatac2-python/tests/test_pipeline.py:10-31).
"""
from __future__ import annotations

import numpy as np

E_START_MISSING, E_END_MISSING, E_BARCODE_MISSING = 1, 2, 3
E_START_BAD, E_END_BAD, E_COUNT_BAD, E_STRAND_BAD = 4, 5, 6, 7
E_BARCODE_NONE, E_END_BEFORE_START, E_BARCODE_LONG = 8, 9, 10
E_REVERSE_PAIRED, E_SIZE_I32, E_RANGE, E_SIZE_U32 = 11, 12, 13, 14
E_DUPLICATE, E_MIXED = 100, 101

U64_MAX = 2**64 - 1


class OracleError(Exception):
    def __init__(self, code: int, line: int | None = None):
        super().__init__(f"code {code} at line {line}")
        self.code, self.line = code, line


def _uint(field: bytes, limit: int):
    if not field or not field.isdigit():
        return None
    v = int(field)
    return v if v <= limit else None


def _sat_add(a: int, d: int) -> int:
    return min(max(a + d, 0), U64_MAX)


def parse_lines(text: bytes):
    """Yields (line_no, record) for every record line; record =
    (chrom, start, end, barcode, count, strand) with strand None, '+' or '-'."""
    lines = text.split(b"\n")
    if lines and lines[-1] == b"":
        lines.pop()  # the text ends with a newline
    for no, line in enumerate(lines, start=1):
        if line.endswith(b"\r"):
            line = line[:-1]
        if not line or line.startswith(b"#"):
            continue
        f = line.split(b"\t")
        if len(f) < 2:
            raise OracleError(E_START_MISSING, no)
        start = _uint(f[1], U64_MAX)
        if start is None:
            raise OracleError(E_START_BAD, no)
        if len(f) < 3:
            raise OracleError(E_END_MISSING, no)
        end = _uint(f[2], U64_MAX)
        if end is None:
            raise OracleError(E_END_BAD, no)
        if len(f) < 4:
            raise OracleError(E_BARCODE_MISSING, no)
        barcode = None if f[3] == b"." else f[3]
        count = 1
        if len(f) > 4 and f[4] != b".":
            count = _uint(f[4], 2**32 - 1)
            if count is None:
                raise OracleError(E_COUNT_BAD, no)
        strand = None
        if len(f) > 5:
            if f[5] == b"+":
                strand = "+"
            elif f[5] == b"-":
                strand = "-"
            elif f[5] != b".":
                raise OracleError(E_STRAND_BAD, no)
        yield no, (f[0], start, end, barcode, count, strand)


def import_text(text: bytes, chrom_sizes: dict, mito=("chrM", "M"), min_num_fragments: int = 200,
                sorted_by_barcode: bool = True, shift_left: int = 0, shift_right: int = 0,
                whitelist=None):
    """Returns dict(paired, barcodes [bytes], indptr, pos, val, n_fragment,
    frac_dup, frac_mito). chrom_sizes in stored (name-sorted) order."""
    base, off = {}, 0
    for name in sorted(chrom_sizes):  # BTreeMap order (preprocessing.rs:95,126)
        base[name.encode()] = (off, int(chrom_sizes[name]))
        off += int(chrom_sizes[name])
    mito = {m.encode() for m in mito}
    wl = None if whitelist is None else {w.encode() if isinstance(w, str) else w for w in whitelist}

    recs = []  # (line, chrom, start, end, barcode, count, strand), shifted, end > start
    first_strand, n_rec, n_strand = None, 0, 0
    for no, (chrom, start, end, bc, count, strand) in parse_lines(text):
        n_rec += 1
        n_strand += strand is not None
        if first_strand is None:
            first_strand = [strand]
        if shift_left:
            start = _sat_add(start, shift_left)
            if strand is not None:
                end = _sat_add(end, shift_left)
        if shift_right and strand is None:
            end = _sat_add(end, shift_right)
        if end < start:
            raise OracleError(E_END_BEFORE_START, no)
        if end == start:
            continue  # x.len() > 0 (import.rs:66)
        if bc is None:
            raise OracleError(E_BARCODE_NONE, no)
        if len(bc) > 255:
            raise OracleError(E_BARCODE_LONG, no)
        recs.append((no, chrom, start, end, bc, count, strand))

    if sorted_by_barcode:
        paired = first_strand is None or first_strand[0] is None
    else:
        if 0 < n_strand < n_rec:
            raise OracleError(E_MIXED)
        paired = n_strand == 0

    # groups: runs in file order, or distinct barcodes in byte order
    groups = []
    if sorted_by_barcode:
        for r in recs:
            if groups and groups[-1][0] == r[4]:
                groups[-1][1].append(r)
            else:
                groups.append((r[4], [r]))
    else:
        by = {}
        for r in recs:
            by.setdefault(r[4], []).append(r)
        groups = [(k, by[k]) for k in sorted(by)]

    seen = set()
    rows = []
    for bc, rs in groups:
        if wl is not None and bc not in wl:
            continue
        if bc in seen:  # "Please sort fragment file by barcodes" (import.rs:139-141)
            raise OracleError(E_DUPLICATE)
        seen.add(bc)
        uniq = mito_n = total = 0
        values = []
        for no, chrom, start, end, _, count, strand in rs:
            if chrom not in base:
                continue
            total += count
            if chrom in mito:
                mito_n += 1
            else:
                uniq += 1
            cbase, clen = base[chrom]
            size = end - start
            if strand == "-":
                coord, value = end - 1, -size
            else:
                coord, value = start, size
            if coord >= clen:
                raise OracleError(E_RANGE, no)
            if size > 2**32 - 1:
                raise OracleError(E_SIZE_U32, no)
            if paired and value < 0:
                raise OracleError(E_REVERSE_PAIRED, no)
            if not paired and not (-2**31 <= value <= 2**31 - 1):
                raise OracleError(E_SIZE_I32, no)
            values.append((cbase + coord, value))
        values.sort()
        if uniq >= min_num_fragments:
            rows.append((bc, uniq, total, mito_n, values))

    indptr = np.zeros(len(rows) + 1, dtype=np.int64)
    np.cumsum([len(r[4]) for r in rows], out=indptr[1:])
    pos = np.asarray([p for r in rows for p, _ in r[4]], dtype=np.uint32)
    val = np.asarray([v for r in rows for _, v in r[4]], dtype=np.int64)
    val = val.astype(np.uint32) if paired else val.astype(np.int32)
    u = np.asarray([r[1] for r in rows], dtype=np.uint64)
    t = np.asarray([r[2] for r in rows], dtype=np.uint64)
    m = np.asarray([r[3] for r in rows], dtype=np.uint64)
    with np.errstate(divide="ignore", invalid="ignore"):
        # FragmentQCBuilder::finish (qc.rs:452-465)
        frac_dup = 1.0 - (u + m).astype(np.float64) / t.astype(np.float64)
        frac_mito = m.astype(np.float64) / (u + m).astype(np.float64)
    return dict(paired=paired, barcodes=[r[0] for r in rows], indptr=indptr, pos=pos, val=val,
                n_fragment=u, frac_dup=frac_dup, frac_mito=frac_mito)


def export_lines(barcodes, indptr, pos, val, chrom_sizes: dict, paired: bool):
    """Fragment lines as export_fragments writes them (export.rs:90-106; Display
    of Fragment, qc.rs:286-303), one row after another."""
    names = sorted(chrom_sizes)
    starts = np.cumsum([0] + [int(chrom_sizes[n]) for n in names])
    out = []
    for r, bc in enumerate(barcodes):
        for k in range(int(indptr[r]), int(indptr[r + 1])):
            p = int(pos[k])
            c = int(np.searchsorted(starts, p, side="right")) - 1
            coord = p - int(starts[c])
            v = int(val[k])
            if paired:
                out.append(f"{names[c]}\t{coord}\t{coord + v}\t{bc}\t1")
            elif v >= 0:
                out.append(f"{names[c]}\t{coord}\t{coord + v}\t{bc}\t1\t+")
            else:
                out.append(f"{names[c]}\t{max(coord + 1 + v, 0)}\t{coord + 1}\t{bc}\t1\t-")
    return out


def import_text_cpp(text: bytes, chrom_sizes: dict, mito=("chrM", "M"), min_num_fragments: int = 200,
                    sorted_by_barcode: bool = True, shift_left: int = 0, shift_right: int = 0,
                    whitelist=None):
    """The C++ restatement (oracle/import_oracle.cpp): same rules and result
    layout as import_text, one thread; the timed CPU baseline of the import."""
    import ctypes as C

    from . import LIB_PATH, build
    if not LIB_PATH.exists():
        build()
    h = C.CDLL(str(LIB_PATH))
    h.oracle_import_run.restype = C.c_void_p
    h.oracle_import_run.argtypes = [C.c_char_p, C.c_int64, C.POINTER(C.c_char_p), C.c_void_p, C.c_void_p,
                                    C.c_int32, C.POINTER(C.c_char_p), C.c_int32, C.c_uint64, C.c_int32,
                                    C.c_int64, C.c_int64]
    for name, res in [("oracle_import_error", C.c_char_p), ("oracle_import_rows", C.c_int64),
                      ("oracle_import_nnz", C.c_int64), ("oracle_import_paired", C.c_int32),
                      ("oracle_import_width", C.c_int32)]:
        getattr(h, name).restype = res
        getattr(h, name).argtypes = [C.c_void_p]
    h.oracle_import_copy.argtypes = [C.c_void_p] * 6 + [C.c_int32]
    h.oracle_import_free.argtypes = [C.c_void_p]
    names = [n.encode() for n in chrom_sizes]
    lens = np.asarray([int(chrom_sizes[n]) for n in chrom_sizes], dtype=np.uint64)
    mt = np.asarray([1 if n in set(mito) else 0 for n in chrom_sizes], dtype=np.uint8)
    wl = [] if whitelist is None else [w.encode() if isinstance(w, str) else w for w in whitelist]
    cn = (C.c_char_p * max(len(names), 1))(*names)
    cw = (C.c_char_p * max(len(wl), 1))(*wl)
    r = h.oracle_import_run(text, len(text), cn, lens.ctypes.data, mt.ctypes.data, len(names), cw,
                            -1 if whitelist is None else len(wl), max(int(min_num_fragments), 0),
                            1 if sorted_by_barcode else 0, int(shift_left), int(shift_right))
    try:
        e = h.oracle_import_error(r).decode()
        if e:
            parts = e.split()
            raise OracleError(int(parts[1]), int(parts[3]) if len(parts) > 3 else None)
        R, nnz, width = h.oracle_import_rows(r), h.oracle_import_nnz(r), h.oracle_import_width(r)
        paired = bool(h.oracle_import_paired(r))
        indptr = np.zeros(R + 1, dtype=np.int64)
        pos = np.zeros(nnz, dtype=np.uint32)
        val = np.zeros(nnz, dtype=np.uint32)
        qc = np.zeros((R, 3), dtype=np.uint64)
        nm = np.zeros((R, width), dtype=np.uint8)
        h.oracle_import_copy(r, indptr.ctypes.data, pos.ctypes.data, val.ctypes.data, qc.ctypes.data,
                             nm.ctypes.data, width)
    finally:
        h.oracle_import_free(r)
    u, t, m = qc[:, 0], qc[:, 1], qc[:, 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        frac_dup = 1.0 - (u + m).astype(np.float64) / t.astype(np.float64)
        frac_mito = m.astype(np.float64) / (u + m).astype(np.float64)
    return dict(paired=paired, barcodes=[bytes(x).rstrip(b"\0") for x in nm], indptr=indptr, pos=pos,
                val=val if paired else val.view(np.int32), n_fragment=u, frac_dup=frac_dup,
                frac_mito=frac_mito)


def random_text(genome: dict, seed: int, n_lines: int, n_bc: int = 40, sorted_bc: bool = True,
                single: bool = False, noise: bool = True) -> bytes:
    """A seeded fragment file over `genome` plus two unknown chromosome names:
    empty fragments, optional count / strand fields, comments and CRLF lines."""
    rng = np.random.default_rng(seed)
    names = ["chr1", "chr2", "chrM", "chrX", "chrUn_x", "M"]
    weights = np.asarray([40, 25, 5, 20, 5, 5], dtype=float)
    bcs = [f"{''.join(rng.choice(list('ACGT'), 12))}-{rng.integers(1, 3)}" for _ in range(n_bc)]
    rows = []
    for _ in range(n_lines):
        c = names[rng.choice(len(names), p=weights / weights.sum())]
        ln = genome.get(c, 10_000)
        size = min(int(rng.integers(0, 300)), ln // 2) if rng.random() > 0.03 else 0
        s = int(rng.integers(0, max(ln - size - 10, 1)))  # room for shifts of a few bases
        e = s + size
        bc = bcs[min(int(rng.exponential(n_bc / 4)), n_bc - 1)]
        f = [c, str(s), str(e), bc]
        r = rng.random()
        if single:
            f += [str(rng.integers(1, 4)), "+" if rng.random() < 0.5 else "-"]
        elif r < 0.3:
            f += [str(rng.integers(0, 6))]
        elif r < 0.4:
            f += ["."]
        elif r < 0.45:
            f += [".", "."]
        rows.append((bc, "\t".join(f)))
    if sorted_bc:
        rows.sort(key=lambda x: x[0])
    out = []
    for _, line in rows:
        if noise and rng.random() < 0.01:
            out.append("# a comment")
        out.append(line + ("\r" if noise and rng.random() < 0.01 else ""))
    return ("\n".join(out) + ("\n" if rng.random() < 0.5 else "")).encode()
