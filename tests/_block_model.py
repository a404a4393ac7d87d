"""CPU model of the block-tiled operator layout (csrc/block.hip): the row order,
pieces, blocks, piece table, packed entries and tile starts that
snapgpu_block_plan/build produce, and the product kernel's division of each
tile between 256 four-lane groups, step by step (the 64-entry run scan of a
row emulated lane by lane). build() asserts every index it forms is in range;
product() asserts every entry is summed exactly once. Used by the block-layout
GPU tests as the expected layout and on the CPU as a check of the indexing
before any kernel runs."""
import numpy as np
import scipy.sparse as sp

kBR, kBC, kPiece = 1024, 1024, 16

def runscan_rows(idx_row):
    """emulates RunScan over one row: returns per entry (s, last)"""
    e1 = len(idx_row); out_s = np.zeros(e1, np.int64); out_last = np.zeros(e1, bool)
    run_start = 0
    for b in range(0, e1, 64):
        lanes = np.arange(64); e = b + lanes; has = e < e1
        col = np.where(has, idx_row[np.minimum(e, e1-1)], 0); pnl = np.where(has, col // kBR, -1)
        prev = np.roll(pnl, 1); prev[0] = -1 if b == 0 else idx_row[b-1] // kBR
        nxt = np.roll(pnl, -1); nxt[63] = idx_row[e[63]+1] // kBR if e[63] + 1 < e1 else -2
        nxt = np.where(e + 1 >= e1, -2, nxt)
        starts = has & (pnl != prev)
        for l in range(64):
            if not has[l]: continue
            upto = np.flatnonzero(starts[:l+1])
            s = b + upto[-1] if upto.size else run_start
            out_s[e[l]] = s; out_last[e[l]] = nxt[l] != pnl[l]
        if starts.any(): run_start = b + np.flatnonzero(starts)[-1]
    return out_s, out_last

def build(X, dense_rows, target_blocks):
    X = sp.csr_matrix(X); X.sort_indices()
    n = X.shape[0]; nnz = X.nnz
    lens = np.diff(X.indptr)
    order = np.argsort(-lens, kind='stable')
    P = -(-dense_rows // kBR)
    pieces = np.zeros(n, np.int64)
    runinfo = {}
    for q, r in enumerate(order):
        a, b = X.indptr[r], X.indptr[r+1]
        s, last = runscan_rows(X.indices[a:b])
        runinfo[r] = (s, last)
        best = 0
        for k in np.flatnonzero(last):
            best = max(best, -(-(k + 1 - s[k]) // kPiece))
        pieces[q] = best
    target = max(1, -(-max(nnz,1) // target_blocks))
    qblock = np.zeros(n, np.int64); vslot = np.zeros(n, np.int64)
    row0=[0]; nslots=[]; base=[]; cur_e=0; cur_s=0; ts=0
    for q in range(n):
        assert pieces[q] <= kBC
        if q > row0[-1] and (cur_s + pieces[q] > kBC or cur_e + lens[order[q]] > target):
            nslots.append(cur_s); base.append(ts); ts += cur_s * P; row0.append(q); cur_s = 0; cur_e = 0
        qblock[q] = len(row0) - 1; vslot[q] = cur_s; cur_s += pieces[q]; cur_e += lens[order[q]]
    nslots.append(cur_s); base.append(ts); ts += cur_s * P; row0.append(n); base.append(ts)
    nb = len(nslots)
    Lq = np.zeros(ts + 1, np.int64)
    def tpos(blk, p, s): return base[blk] + p * nslots[blk] + s
    for q, r in enumerate(order):
        a = X.indptr[r]; s_arr, last = runinfo[r]
        for k in np.flatnonzero(last):
            ln = k + 1 - s_arr[k]; pnl = X.indices[a + k] // kBR
            for j in range(-(-ln // kPiece)):
                tp = tpos(qblock[q], pnl, vslot[q] + j); assert 0 <= tp < ts
                Lq[tp] = min(kPiece, ln - j * kPiece)
    incl = np.cumsum(Lq); E = np.concatenate([[0], incl[:-1]])
    assert E[ts] == nnz, (E[ts], nnz)
    # per tile: pieces by decreasing length, then slot; row i holds entry i of
    # every piece longer than i
    rank = np.full(ts + 1, -1, np.int64)
    tile_off = np.zeros((nb * P, kPiece), np.int64)
    for blk in range(nb):
        for p in range(P):
            t0 = tpos(blk, p, 0); lens_t = Lq[t0:t0 + nslots[blk]]
            slots = np.flatnonzero(lens_t > 0)
            order_t = sorted(slots, key=lambda s_: (-lens_t[s_], s_))
            for r_, s_ in enumerate(order_t):
                rank[t0 + s_] = r_
            c = np.array([(lens_t > i).sum() for i in range(kPiece)])
            tile_off[blk * P + p] = np.cumsum(c)
    assert tile_off.max(initial=0) < 65536
    ent = np.full(nnz, -1, np.int64); val = np.zeros(nnz, np.int64)
    for q, r in enumerate(order):
        a = X.indptr[r]; s_arr, last = runinfo[r]
        for k in range(X.indptr[r+1] - a):
            col = X.indices[a+k]; pnl = col // kBR; i = k - s_arr[k]; piece = i // kPiece; row = i % kPiece
            slot = vslot[q] + piece; assert slot < nslots[qblock[q]]
            tp = tpos(qblock[q], pnl, slot)
            tile = qblock[q] * P + pnl
            dst = E[tpos(qblock[q], pnl, 0)] + (0 if row == 0 else tile_off[tile, row - 1]) + rank[tp]
            assert 0 <= dst < nnz and ent[dst] == -1
            ent[dst] = (col - pnl * kBR) | (slot << 20)
            val[dst] = X.data[a+k]
    assert (ent >= 0).all()
    tile_start = np.array([E[base[b] + p * nslots[b]] for b in range(nb) for p in range(P)] + [nnz])
    assert (np.diff(tile_start) >= 0).all()
    return dict(order=order, row0=np.array(row0), vslot=vslot, nslots=np.array(nslots), tile_start=tile_start,
                tile_off=tile_off, ent=ent, val=val, P=P, nb=nb)

def product(Lay, D, n_out):
    """Thread k of a block's workgroup owns the k-th piece of each tile: its
    entries are at tile start + start of row i + k for the rows i whose piece
    count exceeds k; each piece's sum goes to its slot."""
    out = np.zeros((n_out, D.shape[1]))
    P = Lay['P']
    for b in range(Lay['nb']):
        r0, r1 = Lay['row0'][b], Lay['row0'][b+1]; ns = Lay['nslots'][b]
        acc_l = np.zeros((max(ns, 1), D.shape[1]))
        for p in range(P):
            t = b * P + p
            st = Lay['tile_start'][t]; offs = Lay['tile_off'][t]
            seen = 0
            for k in range(kBC):
                prev = 0; acc = 0; slot = None
                for i in range(kPiece):
                    if k < offs[i] - prev:
                        at = st + prev + k
                        assert at < Lay['tile_start'][t + 1]
                        e = Lay['ent'][at]; gl = e & 0x3FF; s_ = (e >> 20) & 0x7FF
                        assert slot is None or s_ == slot
                        slot = s_
                        acc = acc + Lay['val'][at] * D[p * kBR + gl]
                        seen += 1
                    prev = offs[i]
                if slot is not None:
                    assert slot < ns
                    acc_l[slot] += acc
            assert seen == Lay['tile_start'][t + 1] - st, "every entry of the tile read once"
        for i, r in enumerate(range(r0, r1)):
            s0 = Lay['vslot'][r]; s1 = Lay['vslot'][r+1] if r + 1 < r1 else ns
            out[Lay['order'][r]] = acc_l[s0:s1].sum(0)
    return out
