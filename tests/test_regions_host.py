"""Host-side region handling of the peak and gene matrices (no GPU).

The gene path builds its promoter regions as (chrom, start, end) tuples
instead of formatting "chrom:start-end" strings and parsing them back; the
tuples must equal what parse_regions reads from those strings, including the
strip of the whole string, chromosome names holding ':' and duplicates, and
the string path must still raise where the reference's parse fails.
"""
import pytest

from snapatac2_amd import _lib
from snapatac2_amd.pp._genes import _promoter_tuples, promoter_regions
from snapatac2_amd.pp._peaks import parse_regions

TX = [dict(chrom="chr1", left=101, right=200, strand="+", transcript_id="t1"),
      dict(chrom="chr1", left=151, right=400, strand="+", transcript_id="t2"),
      dict(chrom="chr1", left=501, right=600, strand="-", transcript_id="t3"),
      dict(chrom="chr1", left=101, right=200, strand="+", transcript_id="t4"),  # same region as t1
      dict(chrom=" chr2", left=5, right=9, strand="-", transcript_id="t5"),  # leading space
      dict(chrom="HLA:A", left=1, right=1, strand="+", transcript_id="t6")]  # ':' in the name


@pytest.mark.parametrize("up,down,body", [(2000, 0, True), (20, 5, False), (0, 0, True), (0, 300, False)])
def test_promoter_tuples_match_string_parse(up, down, body):
    assert _promoter_tuples(TX, up, down, body) == parse_regions(promoter_regions(TX, up, down, body))[1]


def test_promoter_tuples_known_values():
    regs = _promoter_tuples(TX, 20, 0, True)
    assert regs[:3] == [("chr1", 80, 199), ("chr1", 130, 399), ("chr1", 500, 619)]
    assert len(regs) == 5  # t4 repeats t1


def test_negative_end_falls_back_to_the_string_error():
    tx = [dict(chrom="chr1", left=0, right=0, strand="+", transcript_id="t")]
    assert _promoter_tuples(tx, 0, 0, True) is None
    with pytest.raises(_lib.PanicException):
        parse_regions(promoter_regions(tx, 0, 0, True))


def test_missing_strand_raises():
    tx = [dict(chrom="chr1", left=1, right=5, strand=".", transcript_id="t")]
    with pytest.raises(_lib.PanicException):
        _promoter_tuples(tx, 0, 0, True)


def test_parse_regions_distinct_first_occurrence():
    names, regs = parse_regions(["chr1:5-10", " chr1:5-10 ", "chr1:005-10", "chr2:1-2", "chr1:5-10"])
    assert names == ["chr1:5-10", "chr2:1-2"] and regs == [("chr1", 5, 10), ("chr2", 1, 2)]
    for bad in ["chr1:5", "chr1-5-10", "chr1:-5-10", "chr1:a-b"]:
        with pytest.raises(_lib.PanicException):
            parse_regions([bad])
