"""The fragment-import oracle against hand-computed answers (no GPU).

Each expected value below is worked out by hand from the reference's rules:
positions are genome base indices over the name-sorted chromosomes
(GenomeBaseIndex, genome.rs:296-312,387-406); values are the fragment size
(paired) or +-size at the start / at end - 1 (single-end); rows hold only
in-genome fragments, ascending by (position, value) (import.rs:156-204);
QC: n_fragment = unique (non-mitochondrial) fragments, frac_dup = 1 -
(unique + mito) / sum(count), frac_mito = mito / (unique + mito) (qc.rs:427-466).
"""
import numpy as np
import pytest

from oracle.fragments import (E_DUPLICATE, E_END_BEFORE_START, E_MIXED, E_RANGE, E_REVERSE_PAIRED,
                              E_START_BAD, E_STRAND_BAD, OracleError, export_lines, import_text)

GENOME = {"chr2": 50, "chr1": 100, "chrM": 20}  # stored order chr1, chr2, chrM: bases 0, 100, 150

PAIRED = (b"# comment line\n"
          b"chr1\t10\t20\tA\t1\n"
          b"chr1\t5\t8\tA\t2\n"
          b"chr2\t0\t4\tA\t1\n"
          b"chrM\t1\t3\tA\t1\n"
          b"chr3\t1\t3\tA\t1\n"      # not in the genome: neither stored nor counted
          b"chr1\t7\t7\tA\t1\n"      # empty fragment: dropped before grouping
          b"chr1\t1\t2\tB\t3")       # no trailing newline


def test_paired_known_answer():
    r = import_text(PAIRED, GENOME, min_num_fragments=1)
    assert r["paired"] and r["barcodes"] == [b"A", b"B"]
    assert r["indptr"].tolist() == [0, 4, 5]
    assert r["pos"].tolist() == [5, 10, 100, 151, 1]
    assert r["val"].tolist() == [3, 10, 4, 2, 1]
    assert r["val"].dtype == np.uint32
    assert r["n_fragment"].tolist() == [3, 1]
    # A: unique 3, mito 1, total 1 + 2 + 1 + 1 = 5; B: unique 1, total 3
    assert r["frac_dup"].tolist() == [1.0 - 4.0 / 5.0, 1.0 - 1.0 / 3.0]
    assert r["frac_mito"].tolist() == [0.25, 0.0]


def test_min_fragments_and_whitelist():
    assert import_text(PAIRED, GENOME, min_num_fragments=2)["barcodes"] == [b"A"]
    assert import_text(PAIRED, GENOME, min_num_fragments=0, whitelist={"B"})["barcodes"] == [b"B"]


def test_single_end_known_answer():
    text = (b"chr1\t10\t15\tC\t1\t+\n"
            b"chr1\t20\t26\tC\t1\t-\n"   # stored at end - 1 = 25 with -6
            b"chr1\t25\t30\tC\t1\t+\n"
            b"chr2\t3\t9\tC\t1\t-\n")    # 100 + 8
    r = import_text(text, GENOME, min_num_fragments=0)
    assert not r["paired"] and r["val"].dtype == np.int32
    assert r["pos"].tolist() == [10, 25, 25, 108]
    assert r["val"].tolist() == [5, -6, 5, -6]


def test_shifts_saturate():
    # paired: start += shift_left, end += shift_right; single: both ends += shift_left
    text = b"chr1\t10\t20\tA\nchr1\t1\t9\tA\nchr1\t30\t40\tB\t1\t-\n"
    # A: (6, 15) -> 9 at 6; (0 saturated, 4) -> 4 at 0; B ('-' in a paired file) is an error
    # unless B is left out by the whitelist
    r = import_text(text, GENOME, min_num_fragments=0, shift_left=-4, shift_right=-5, whitelist={"A"})
    assert r["pos"].tolist() == [0, 6] and r["val"].tolist() == [4, 9]
    with pytest.raises(OracleError) as e:
        import_text(text, GENOME, min_num_fragments=0, shift_left=-4, shift_right=-5)
    assert e.value.code == E_REVERSE_PAIRED and e.value.line == 3


def test_unsorted_groups_in_byte_order():
    text = b"chr1\t1\t5\tb2\nchr1\t2\t5\tA\nchr1\t0\t5\tb2\nchr1\t3\t5\tB\nchr1\t9\t12\tA\n"
    r = import_text(text, GENOME, min_num_fragments=0, sorted_by_barcode=False)
    assert r["barcodes"] == [b"A", b"B", b"b2"]  # bytes: 'A' < 'B' < 'b'
    assert r["indptr"].tolist() == [0, 2, 3, 5]
    assert r["pos"].tolist() == [2, 9, 3, 0, 1]
    with pytest.raises(OracleError) as e:
        import_text(text, GENOME, min_num_fragments=0, sorted_by_barcode=True)
    assert e.value.code == E_DUPLICATE


def test_errors():
    cases = [
        (b"chr1\t1x\t5\tA\n", E_START_BAD, 1),
        (b"chr1\t1\t5\tA\nchr1\t1\t5\tA\t1\t*\n", E_STRAND_BAD, 2),
        (b"chr1\t9\t5\tA\n", E_END_BEFORE_START, 1),
        (b"chr1\t1\t5\tA\nchr2\t50\t60\tA\n", E_RANGE, 2),  # start 50 >= length of chr2
    ]
    for text, code, line in cases:
        with pytest.raises(OracleError) as e:
            import_text(text, GENOME, min_num_fragments=0)
        assert (e.value.code, e.value.line) == (code, line), text
    with pytest.raises(OracleError) as e:
        import_text(b"chr1\t1\t5\tA\nchr1\t1\t5\tA\t1\t+\n", GENOME, sorted_by_barcode=False)
    assert e.value.code == E_MIXED


def test_export_round_trip():
    """import then export gives back the records (the reference's test_import,
    snapatac2-python/tests/test_tools.py:128-142), here on in-genome records
    with count 1."""
    rng = np.random.default_rng(0)
    chroms = sorted(GENOME)
    lines = []
    for _ in range(300):
        c = chroms[rng.integers(0, 2)]  # chr1 or chr2
        s = int(rng.integers(0, GENOME[c] - 1))
        e = int(rng.integers(s + 1, GENOME[c] + 5))
        lines.append(f"{c}\t{s}\t{e}\tbc{rng.integers(0, 7)}\t1")
    text = ("\n".join(lines) + "\n").encode()
    r = import_text(text, GENOME, min_num_fragments=0, sorted_by_barcode=False)
    out = export_lines([b.decode() for b in r["barcodes"]], r["indptr"], r["pos"], r["val"], GENOME,
                       r["paired"])
    assert sorted(out) == sorted(lines)


GENOME2 = {"chr2": 3000, "chr1": 5000, "chrM": 300, "chrX": 2000}


@pytest.mark.parametrize("seed,sorted_bc,single,kw", [
    (0, True, False, dict(min_num_fragments=0)),
    (1, False, False, dict(min_num_fragments=5)),
    (2, True, True, dict(min_num_fragments=0, shift_left=-3)),
    (3, False, True, dict(min_num_fragments=2, whitelist=None)),
    (4, True, False, dict(min_num_fragments=0, shift_left=-4, shift_right=5)),
])
def test_cpp_restatement_matches(seed, sorted_bc, single, kw):
    """The C++ restatement (the timed CPU baseline) gives the Python
    restatement's result on seeded files."""
    from oracle.fragments import import_text_cpp, random_text
    text = random_text(GENOME2, seed, 5000, 60, sorted_bc, single)
    a = import_text(text, GENOME2, sorted_by_barcode=sorted_bc, **kw)
    b = import_text_cpp(text, GENOME2, sorted_by_barcode=sorted_bc, **kw)
    assert a["paired"] == b["paired"] and a["barcodes"] == b["barcodes"]
    for k in ("indptr", "pos", "val", "n_fragment"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("frac_dup", "frac_mito"):
        assert np.array_equal(a[k], b[k], equal_nan=True), k


def test_cpp_restatement_errors():
    from oracle.fragments import import_text_cpp
    for text, code, line in [(b"chr1\t1x\t5\tA\n", E_START_BAD, 1),
                             (b"chr1\t1\t5\tA\nchr2\t50\t60\tA\n", E_RANGE, 2)]:
        with pytest.raises(OracleError) as e:
            import_text_cpp(text, GENOME, min_num_fragments=0)
        assert (e.value.code, e.value.line) == (code, line)
