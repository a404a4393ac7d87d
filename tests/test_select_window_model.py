"""CPU model of the windowed column selection (select_window_kernel,
csrc/spectral.hip): a block sweeps the columns in windows of 8192 table words
(262,144 columns); each row is walked from its cursor in 64-entry steps while
its columns stay below the window's end; columns below the window (rows out of
order, negative columns) use the global table and the last window takes every
remaining entry. The model follows those steps and must give the plain
selection (every entry once, in row order, columns outside [0, m) dropped) for
sorted rows, shuffled rows and invalid columns."""
import numpy as np
import pytest

WIN_WORDS = 8192
STEP = 64


def _table(m, cols):
    words = -(-m // 32)
    sel = np.zeros(2 * words, dtype=np.int64)
    cols = np.asarray(cols)
    for w in range(words):
        a, b = np.searchsorted(cols, 32 * w), np.searchsorted(cols, 32 * w + 32)
        sel[2 * w] = sum(1 << int(c - 32 * w) for c in cols[a:b])
        sel[2 * w + 1] = a
    return sel, words


def _rank(sel, words, c, w0=None, nw=None):
    """New id of column c, or -1 (the kernel's table lookup)."""
    wd = c >> 5
    if c < 0 or wd >= words:
        return -1
    mask, below = int(sel[2 * wd]), int(sel[2 * wd + 1])
    bit = c & 31
    if not (mask >> bit) & 1:
        return -1
    return below + bin(mask & ((1 << bit) - 1)).count("1")


def window_walk(indptr, idx, val, m, cols):
    sel, words = _table(m, cols)
    n = len(indptr) - 1
    nwin = -(-words // WIN_WORDS)
    cur = [int(indptr[i]) for i in range(n)]
    out = [[] for _ in range(n)]
    for w in range(nwin):
        w0 = w * WIN_WORDS
        nw = min(WIN_WORDS, words - w0)
        c_hi = 1 << 62 if w + 1 == nwin else (w0 + nw) * 32
        for i in range(n):
            e, e1 = cur[i], int(indptr[i + 1])
            while e < e1:
                take = 0
                while take < STEP and e + take < e1 and int(idx[e + take]) < c_hi:
                    take += 1
                for t in range(take):
                    nc = _rank(sel, words, int(idx[e + t]))
                    if nc >= 0:
                        out[i].append((nc, int(val[e + t])))
                e += take
                if take < STEP:
                    break
            cur[i] = e
    assert all(cur[i] == indptr[i + 1] for i in range(n)), "an entry was never visited"
    ip = np.zeros(n + 1, dtype=np.int64)
    ip[1:] = np.cumsum([len(o) for o in out])
    return ip, np.array([c for o in out for c, _ in o], dtype=np.int64), \
        np.array([v for o in out for _, v in o], dtype=np.int64)


def plain_selection(indptr, idx, val, m, cols):
    rank = np.full(m, -1, dtype=np.int64)
    rank[np.asarray(cols)] = np.arange(len(cols))
    ok = (idx >= 0) & (idx < m)
    nc = np.full(idx.size, -1, dtype=np.int64)
    nc[ok] = rank[idx[ok]]
    keep = nc >= 0
    row = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    ip = np.zeros(len(indptr), dtype=np.int64)
    np.cumsum(np.bincount(row[keep], minlength=len(indptr) - 1), out=ip[1:])
    return ip, nc[keep], val[keep].astype(np.int64)


@pytest.mark.parametrize("case", ["sorted", "shuffled", "one_window"])
def test_window_walk_equals_plain_selection(case):
    rng = np.random.default_rng(3)
    n, m = 30, (5000 if case == "one_window" else 1_100_000)
    lens = rng.integers(0, 400, n)
    lens[::7] = 0
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=indptr[1:])
    idx = np.concatenate([np.sort(rng.choice(m, k, replace=False)) for k in lens])
    val = rng.integers(1, 100, idx.size)
    if case == "shuffled":
        for r in range(0, n, 3):
            a, b = indptr[r], indptr[r + 1]
            idx[a:b] = rng.permutation(idx[a:b])
        bad = rng.random(idx.size) < 0.02
        idx[bad] = rng.choice(np.array([-1, -70000, m, m + 5, 2**31 - 1]), int(bad.sum()))
    cols = np.flatnonzero(rng.random(m) < 0.3)
    got = window_walk(indptr, idx, val, m, cols)
    ref = plain_selection(indptr, idx, val, m, cols)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
