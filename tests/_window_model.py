"""CPU model of the window layout's deterministic parts (snapgpu_window_plan and
snapgpu_window_build, csrc/window.hip): slots per row, the rows of each
workgroup, the entry words of each (workgroup, window) bucket and their pieces,
and the piece-aligned quad chunks. The build places a bucket's segments in any
order, so buckets are compared as multisets of pieces."""
import numpy as np

WIN_BITS = 10
WIN = 1 << WIN_BITS
SLOTS = 2048
PIECE = 16
CHUNK = 64
QUADS = 256
FLAG = 1 << (WIN_BITS + 11)


def segments(ptr, idx, r):
    """(window, first entry, length) of row r's segments: maximal same-window
    stretches inside each 64-entry chunk of the row, in order."""
    e0, e1 = int(ptr[r]), int(ptr[r + 1])
    out = []
    for base in range(e0, e1, CHUNK):
        end = min(base + CHUNK, e1)
        w = idx[base:end] >> WIN_BITS
        starts = np.flatnonzero(np.r_[True, w[1:] != w[:-1]])
        ends = np.r_[starts[1:], end - base]
        out += [(int(w[s]), base + int(s), int(t - s)) for s, t in zip(starts, ends)]
    return out


def piece_numbers(segs):
    """First piece number of each segment within its (row, window) run: only a
    chunk's first segment can continue the previous segment's run."""
    first, pieces, prev_w = [], 0, None
    for w, _, n in segs:
        f = pieces if w == prev_w else 0
        first.append(f)
        pieces = f + -(-n // PIECE)
        prev_w = w
    return first


def row_slots(ptr, idx):
    n = len(ptr) - 1
    s = np.zeros(n, dtype=np.int64)
    for r in range(n):
        segs = segments(ptr, idx, r)
        firsts = piece_numbers(segs)
        s[r] = max([f + -(-m // PIECE) for (_, _, m), f in zip(segs, firsts)], default=0)
    return s


def plan(ptr, idx):
    """(block_row, row_slot): rows packed in order into workgroups of <= 2048 slots."""
    s = row_slots(ptr, idx)
    block_row, row_slot, used = [0], np.zeros(len(s), dtype=np.uint32), 0
    for r, sr in enumerate(s):
        assert sr <= SLOTS
        if used + sr > SLOTS:
            block_row.append(r)
            used = 0
        row_slot[r] = used | (int(sr) << 16)
        used += int(sr)
    block_row.append(len(s))
    return np.asarray(block_row, dtype=np.int64), row_slot


def bucket_pieces(ptr, idx, val, block_row, row_slot, dense_rows):
    """{(block, window): sorted list of pieces}, a piece being the tuple of its
    entry words in order."""
    nw = max(1, -(-dense_rows // WIN))
    out = {}
    for b in range(len(block_row) - 1):
        for r in range(int(block_row[b]), int(block_row[b + 1])):
            slot0 = int(row_slot[r]) & 0xFFFF
            segs = segments(ptr, idx, r)
            for (w, e0, n), f in zip(segs, piece_numbers(segs)):
                words = [(int(idx[e0 + i]) & (WIN - 1)) | ((slot0 + f + i // PIECE) << WIN_BITS) |
                         (FLAG if i % PIECE == 0 else 0) | (int(val[e0 + i]) << 24) for i in range(n)]
                lst = out.setdefault((b, w), [])
                for p0 in range(0, n, PIECE):
                    lst.append(tuple(words[p0:p0 + PIECE]))
    return {k: sorted(v) for k, v in out.items()}, nw


def split_pieces(words):
    """A bucket's words cut at piece starts (every piece starts with a flag)."""
    out, cur = [], None
    for wd in words:
        if wd & FLAG:
            if cur is not None:
                out.append(tuple(cur))
            cur = [wd]
        else:
            assert cur is not None, "bucket does not start at a piece start"
            cur.append(wd)
    if cur is not None:
        out.append(tuple(cur))
    return out


def quad_offsets(words):
    """Quad k's chunk start: the piece start nearest to floor(k * len / 256),
    the earlier one on a tie; the bucket's end counts as a piece start."""
    n = len(words)
    starts = [i for i, w in enumerate(words) if w & FLAG] + [n]
    out = []
    for k in range(QUADS + 1):
        t = k * n // QUADS
        up = min(p for p in starts if p >= t)
        down = max(p for p in starts if p <= t) if t > 0 else 0
        out.append(down if t - down <= up - t else up)
    return out


def apply(ptr, idx, val, block_row, row_slot, dense, scale):
    """out[row] = scale[row] * sum over the row's slots, in slot order, of each
    window's piece sums, each piece summed in entry order, all in float32 (the
    kernel's summation order; entries within a piece in column order)."""
    dense = np.asarray(dense, dtype=np.float32)
    n = len(ptr) - 1
    out = np.zeros((n, dense.shape[1]), dtype=np.float32)
    for b in range(len(block_row) - 1):
        acc = np.zeros((SLOTS, dense.shape[1]), dtype=np.float32)
        per_window = {}
        for r in range(int(block_row[b]), int(block_row[b + 1])):
            slot0 = int(row_slot[r]) & 0xFFFF
            segs = segments(ptr, idx, r)
            for (w, e0, m), f in zip(segs, piece_numbers(segs)):
                for p0 in range(0, m, PIECE):
                    a = np.zeros(dense.shape[1], dtype=np.float32)
                    for i in range(p0, min(p0 + PIECE, m)):
                        e = e0 + i  # fma: count * z is exact in float64, one rounding to float32
                        a = (np.float64(val[e]) * dense[idx[e]].astype(np.float64) + a).astype(np.float32)
                    per_window.setdefault(w, []).append((slot0 + f + p0 // PIECE, a))
        for w in sorted(per_window):
            for slot, a in per_window[w]:
                acc[slot] = (acc[slot] + a).astype(np.float32)
        for r in range(int(block_row[b]), int(block_row[b + 1])):
            rs = int(row_slot[r])
            t = np.zeros(dense.shape[1], dtype=np.float32)
            for j in range(rs >> 16):
                t = (t + acc[(rs & 0xFFFF) + j]).astype(np.float32)
            out[r] = (t * np.float32(scale[r])).astype(np.float32)
    return out
