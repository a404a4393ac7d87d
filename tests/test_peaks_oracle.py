"""The peak-matrix oracle against hand-computed counts (no GPU).

Genome chr1 (100 bases) then chr2 (50): positions 0-99 are chr1, 100-149 chr2.
Regions: r0 chr1:10-20, r1 chr1:15-30 (overlaps r0), r2 chr2:0-10, r3 chr1:10-20
(a duplicate of r0: dropped), r4 chr3:0-5 (unknown chromosome: never hit).
"""
import numpy as np

from oracle.peaks import parse_regions, peak_matrix

GENOME = {"chr1": 100, "chr2": 50}
PEAKS = ["chr1:10-20", "chr1:15-30", "chr2:0-10", "chr1:10-20", "chr3:0-5"]
# cell 0: fragment [12, 18) on chr1; cell 1: [19, 25) and chr2 [5, 8) (global 105)
CELL_PTR = [0, 1, 3]
POS = [12, 19, 105]
VAL = [6, 6, 3]


def test_regions_deduplicated_in_first_occurrence_order():
    assert parse_regions(PEAKS) == [("chr1", 10, 20), ("chr1", 15, 30), ("chr2", 0, 10), ("chr3", 0, 5)]


def test_paired_insertion():
    # cell 0: insertions 12 (r0) and 17 (r0, r1) -> r0 once, r1 once
    # cell 1: insertions 19 (r0, r1) and 24 (r1) -> r0 1, r1 1; chr2 5 and 7 -> r2 once
    ip, ix, d = peak_matrix(CELL_PTR, POS, VAL, True, GENOME, PEAKS, "paired-insertion")
    assert ip.tolist() == [0, 2, 5]
    assert ix.tolist() == [0, 1, 0, 1, 2] and d.tolist() == [1, 1, 1, 1, 1]


def test_insertion_and_fragment():
    ip, ix, d = peak_matrix(CELL_PTR, POS, VAL, True, GENOME, PEAKS, "insertion")
    # cell 0: 12 -> r0; 17 -> r0, r1  => r0 2, r1 1; cell 1: 19 -> r0, r1; 24 -> r1; 105, 107 -> r2 x2
    assert ix.tolist() == [0, 1, 0, 1, 2] and d.tolist() == [2, 1, 1, 2, 2]
    ip, ix, d = peak_matrix(CELL_PTR, POS, VAL, True, GENOME, PEAKS, "fragment")
    # [12,18) -> r0, r1; [19,25) -> r0, r1; [5,8) on chr2 -> r2
    assert ix.tolist() == [0, 1, 0, 1, 2] and d.tolist() == [1, 1, 1, 1, 1]


def test_size_filter_and_single_end():
    ip, ix, d = peak_matrix(CELL_PTR, POS, VAL, True, GENOME, PEAKS, "paired-insertion", min_len=4)
    assert ip.tolist() == [0, 2, 4]  # the 3-base chr2 fragment is filtered out
    ip, ix, d = peak_matrix(CELL_PTR, POS, [6, -6, 3], False, GENOME, PEAKS)
    # single-end: one base at the stored position: 12 -> r0; 19 -> r0, r1; 105 -> r2
    assert ix.tolist() == [0, 0, 1, 2] and d.tolist() == [1, 1, 1, 1]


def test_cpp_restatement_matches():
    """The C++ restatement (another search, the CPU baseline) gives the Python
    restatement's matrix on random cells, regions and all strategies."""
    from oracle.peaks import peak_matrix_cpp
    rng = np.random.default_rng(0)
    genome = {"chr1": 5000, "chr2": 3000, "chrX": 2000}
    cell_ptr = np.concatenate([[0], np.cumsum(rng.integers(0, 80, 25))])
    nf = int(cell_ptr[-1])
    pos = rng.integers(0, 10000, nf)
    val = rng.integers(1, 400, nf)
    peaks = [f"{rng.choice(['chr1', 'chr2', 'chrX', 'chrY'])}:{s}-{s + rng.integers(1, 700)}"
             for s in rng.integers(0, 5000, 120)] + ["chr1:10-20", "chr1:10-20"]
    for paired in (True, False):
        v = val if paired else val * rng.choice([-1, 1], nf)
        for strategy in ("paired-insertion", "insertion", "fragment"):
            for lo, hi in ((None, None), (50, 300)):
                a = peak_matrix(cell_ptr, pos, v, paired, genome, peaks, strategy, lo, hi)
                b = peak_matrix_cpp(cell_ptr, pos, v, paired, genome, peaks, strategy, lo, hi)
                for x, y in zip(a, b):
                    assert np.array_equal(x, y), (paired, strategy, lo, hi)
