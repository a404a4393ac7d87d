"""The window layout's CPU model (tests/_window_model.py) on hand-derived cases
(no GPU): segments cut at 64-entry chunks and windows, pieces numbered across a
run's segments, slots per row, greedy workgroup packing, and the quad chunks."""
import numpy as np

import _window_model as wm


def _csr(rows):
    ptr = np.zeros(len(rows) + 1, dtype=np.int64)
    ptr[1:] = np.cumsum([len(r) for r in rows])
    idx = np.asarray([c for r in rows for c in r], dtype=np.int64)
    return ptr, idx


def test_segments_and_slots_by_hand():
    # row 0: 70 entries of window 0 -> chunk cut after 64: segments of 64 and 6,
    #        pieces 0-3 then 4 (numbered across the run): 5 slots
    # row 1: windows 0, 0, 1, 1, 1 -> segments of 2 and 3, one piece each: 1 slot
    # row 2: empty -> 0 slots
    # row 3: 17 entries of window 2 -> pieces of 16 and 1: 2 slots
    W = wm.WIN
    rows = [list(range(70)), [3, 9, W, W + 88, W + 488], [], list(range(2 * W, 2 * W + 17))]
    ptr, idx = _csr(rows)
    assert wm.segments(ptr, idx, 0) == [(0, 0, 64), (0, 64, 6)]
    assert wm.piece_numbers(wm.segments(ptr, idx, 0)) == [0, 4]
    assert wm.segments(ptr, idx, 1) == [(0, 70, 2), (1, 72, 3)]
    assert wm.segments(ptr, idx, 2) == []
    assert wm.row_slots(ptr, idx).tolist() == [5, 1, 0, 2]


def test_run_split_at_chunk_end_with_new_window():
    # 62 entries of window 0, then 4 of window 1: the chunk ends inside window 1,
    # whose run continues into the next chunk with its piece numbering
    rows = [list(range(62)) + list(range(wm.WIN, wm.WIN + 18))]
    ptr, idx = _csr(rows)
    segs = wm.segments(ptr, idx, 0)
    assert segs == [(0, 0, 62), (1, 62, 2), (1, 64, 16)]
    assert wm.piece_numbers(segs) == [0, 0, 1]
    assert wm.row_slots(ptr, idx).tolist() == [4]


def test_greedy_packing():
    # slots 2000, 100, 48, 0, 2048: a workgroup holds at most 2048
    rows = [list(range(c * 0, 0)) for c in range(5)]
    ptr, idx = _csr(rows)
    s = np.asarray([2000, 100, 48, 0, 2048])
    orig = wm.row_slots
    try:
        wm.row_slots = lambda p, i: s
        block_row, row_slot = wm.plan(ptr, idx)
    finally:
        wm.row_slots = orig
    assert block_row.tolist() == [0, 1, 4, 5]
    assert [(int(r) & 0xFFFF, int(r) >> 16) for r in row_slot] == [(0, 2000), (0, 100), (100, 48), (148, 0),
                                                                   (0, 2048)]


def test_buckets_cover_every_entry_once_with_distinct_slots():
    rng = np.random.default_rng(5)
    rows = []
    for r in range(300):
        k = int(rng.integers(0, 200)) if r % 37 else 1500  # a few heavy rows
        rows.append(sorted(rng.choice(3000, size=min(k, 3000), replace=False).tolist()))
    ptr, idx = _csr(rows)
    val = rng.integers(1, 256, idx.size)
    block_row, row_slot = wm.plan(ptr, idx)
    buckets, nw = wm.bucket_pieces(ptr, idx, val, block_row, row_slot, 3000)
    assert nw == -(-3000 // wm.WIN)
    total = sum(len(p) for lst in buckets.values() for p in lst)
    assert total == idx.size
    for (b, w), pieces in buckets.items():
        slots = [p[0] >> wm.WIN_BITS & (wm.SLOTS - 1) for p in pieces]
        assert len(slots) == len(set(slots)), (b, w)  # one piece per slot per window
        assert all(len(p) <= wm.PIECE and p[0] & wm.FLAG for p in pieces)
        words = [x for p in pieces for x in p]
        q = wm.quad_offsets(words)
        assert q[0] == 0 and q[-1] == len(words) and all(a <= c for a, c in zip(q, q[1:]))
        assert wm.split_pieces(words) == list(pieces)
