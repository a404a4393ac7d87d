"""Pins the path-A oracle (oracle/tile_oracle.cpp + oracle/tile.py) without a GPU.

Checks: GenomeBaseIndex known answers (snapatac2-core/src/genome.rs:495-597),
hand-derived per-strategy cases (tests/golden/tile_cases.json), the hg38 bin
count of the add_tile_matrix docstring (_basic.py:545-546), brute force (Python
dicts) vs the C++ restatement on random inputs, and the invariants of the
reference tests test_tile_matrix (test_tools.py:144-162) and test_exclude
(test_pipeline.py:10-31).
"""
import json
import os

import numpy as np
import pytest

from oracle import tile as otile
from snapatac2_amd.genome import HG38, BinIndex, sorted_chrom_sizes

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_genome_index_known_answers():
    g = _load("genome_index.json")
    cs = dict((c, l) for c, l in g["chrom_sizes"])
    for step, pairs in g["regions"].items():
        step = int(step)
        for i, txt in pairs:
            chrom, rng = txt.split(":")
            s, e = map(int, rng.split("-"))
            assert otile.get_region(cs, step, i) == (chrom, s, e)
            assert otile.get_position_rev(cs, step, chrom, s) == i
            # the product's var_names use the same arithmetic
            assert BinIndex(cs, step)[i] == txt
    base = np.cumsum(list(cs.values()))
    for c, (lo, hi) in g["ranges_step1"].items():
        k = list(cs).index(c)
        assert (base[k - 1] if k else 0, base[k]) == (lo, hi)
    for i1, i2 in g["coarsed_step2"]:
        chrom, start, _ = otile.get_region(cs, 1, i1)
        assert otile.get_position_rev(cs, 2, chrom, start) == i2


def test_hg38_bin_count_matches_docstring():
    # 585 x 6062095 at 500 bp with the default exclusions (_basic.py:545-546)
    cs = sorted_chrom_sizes(HG38)
    kept = {c: l for c, l in cs.items() if c not in ("chrM", "chrY", "M", "Y")}
    assert len(BinIndex(kept, 500)) == 6_062_095
    assert len(BinIndex({c: l for c, l in cs.items() if c != "chrM"}, 500)) == 6_176_550


def _cells_to_csr(cells):
    cell_ptr = np.zeros(len(cells) + 1, dtype=np.uint64)
    cell_ptr[1:] = np.cumsum([len(c) for c in cells])
    pos = np.asarray([p for c in cells for p, _ in c], dtype=np.uint32)
    ln = np.asarray([l for c in cells for _, l in c], dtype=np.uint32)
    return cell_ptr, pos, ln


def _rows(m):
    return [[[int(j), int(v)] for j, v in zip(m.indices[m.indptr[r]:m.indptr[r + 1]],
                                              m.data[m.indptr[r]:m.indptr[r + 1]])]
            for r in range(m.shape[0])]


@pytest.mark.parametrize("impl", ["cpp", "python"])
def test_hand_derived_cases(impl):
    g = _load("tile_cases.json")
    cs = dict((c, l) for c, l in g["chrom_sizes"])
    cell_ptr, pos, ln = _cells_to_csr(g["cells"])
    fn = otile.tile_matrix if impl == "cpp" else otile.tile_matrix_bruteforce
    for case in g["cases"]:
        m = fn(cell_ptr, pos, ln, cs, g["bin_size"], case["exclude"], case["min"], case["max"],
               case["strategy"])
        assert m.shape == (3, case["n_bins"]), case
        assert _rows(m) == case["rows"], case


@pytest.mark.parametrize("impl", ["cpp", "python"])
def test_out_of_range_panics(impl):
    g = _load("tile_cases.json")
    cs = dict((c, l) for c, l in g["chrom_sizes"])
    cell_ptr, pos, ln = _cells_to_csr([g["panic_cell"]])
    fn = otile.tile_matrix if impl == "cpp" else otile.tile_matrix_bruteforce
    with pytest.raises(RuntimeError):
        fn(cell_ptr, pos, ln, cs, 100, [], None, None, "paired-insertion")


def random_cells(rng, chrom_sizes, n_cells, max_frags=60, max_len=400):
    names = list(chrom_sizes)
    lens = np.asarray([chrom_sizes[c] for c in names], dtype=np.int64)
    starts = np.concatenate([[0], np.cumsum(lens)])
    cells = []
    for _ in range(n_cells):
        k = int(rng.integers(0, max_frags))
        c = rng.integers(0, len(names), k)
        L = rng.integers(1, max_len, k)
        L = np.minimum(L, lens[c])
        off = (rng.random(k) * (lens[c] - L + 1)).astype(np.int64)
        pos = starts[c] + off
        frs = sorted(zip(pos.tolist(), L.tolist()))
        # duplicates are allowed in the stored encoding
        if frs and rng.random() < 0.3:
            frs.append(frs[-1])
            frs.sort()
        cells.append(frs)
    return cells


@pytest.mark.parametrize("strategy", ["insertion", "fragment", "paired-insertion", "single"])
def test_bruteforce_matches_cpp(strategy):
    rng = np.random.default_rng(7)
    cs = {"chr1": 5000, "chr10": 1234, "chr2": 3000, "chrM": 300, "chrX": 2500}
    cells = random_cells(rng, cs, 40)
    cell_ptr, pos, ln = _cells_to_csr(cells)
    for bin_size, excl, lo, hi in [(100, [], None, None), (37, ["chrM", "chrX"], 10, 300),
                                   (1000, ["chr1"], None, 200)]:
        a = otile.tile_matrix(cell_ptr, pos, ln, cs, bin_size, excl, lo, hi, strategy)
        b = otile.tile_matrix_bruteforce(cell_ptr, pos, ln, cs, bin_size, excl, lo, hi, strategy)
        assert a.shape == b.shape
        assert np.array_equal(a.indptr, b.indptr)
        assert np.array_equal(a.indices, b.indices)
        assert np.array_equal(a.data, b.data)


def test_insertion_total_independent_of_bin_size():
    # test_tools.py:144-162: insertion counts sum to the same total for any bin size
    rng = np.random.default_rng(3)
    cs = {"chr1": 20000, "chr2": 15000, "chr3": 9000}
    cell_ptr, pos, ln = _cells_to_csr(random_cells(rng, cs, 50, 80, 900))
    totals = [otile.tile_matrix(cell_ptr, pos, ln, cs, b, [], None, None, "insertion").sum()
              for b in (500, 1000, 5000, 10000)]
    assert len(set(totals)) == 1 and totals[0] == 2 * len(pos)


def test_pic_fragment_relations():
    rng = np.random.default_rng(5)
    cs = {"chr1": 20000, "chr2": 15000}
    cell_ptr, pos, ln = _cells_to_csr(random_cells(rng, cs, 30, 80, 900))
    ins = otile.tile_matrix(cell_ptr, pos, ln, cs, 300, [], None, None, "insertion")
    pic = otile.tile_matrix(cell_ptr, pos, ln, cs, 300, [], None, None, "paired-insertion")
    frag = otile.tile_matrix(cell_ptr, pos, ln, cs, 300, [], None, None, "fragment")
    starts = np.where(pos < 20000, pos, pos - 20000).astype(np.int64)
    b1, b2 = starts // 300, (starts + ln - 1) // 300
    assert pic.sum() == ins.sum() - np.sum(b1 == b2)
    assert frag.sum() == np.sum(b2 - b1 + 1)


def test_exclude_equals_omitting_chromosomes():
    # test_pipeline.py:10-31: excluding chromosomes at tile time gives the matrix of
    # an import that never saw them.
    rng = np.random.default_rng(11)
    cs = {"chr1": 5000, "chr10": 3000, "chr2": 4000, "chr3": 3500}
    cells = random_cells(rng, cs, 30)
    cell_ptr, pos, ln = _cells_to_csr(cells)
    a = otile.tile_matrix(cell_ptr, pos, ln, cs, 250, ["chr1", "chr10"], None, None,
                          "paired-insertion")
    cs2 = {"chr2": 4000, "chr3": 3500}
    shift = 5000 + 3000
    cells2 = [[(p - shift, l) for p, l in c if p >= shift] for c in cells]
    cp2, pos2, ln2 = _cells_to_csr(cells2)
    b = otile.tile_matrix(cp2, pos2, ln2, cs2, 250, [], None, None, "paired-insertion")
    assert a.shape == b.shape
    assert (a != b).nnz == 0
