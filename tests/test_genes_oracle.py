"""The gene-matrix oracle against hand-computed values (no GPU).

Genome chr1 (1000 bases). Transcripts (1-based, inclusive):
  t1 gene A, + strand, 101-200  -> promoter [100 - 20, 199) = chr1:80-199 (upstream 20)
  t2 gene A, + strand, 151-400  -> chr1:130-399
  t3 gene B, - strand, 501-600  -> [500, 599 + 20) = chr1:500-619
One cell, single-end reads at 85, 150, 610, 700: t1 gets 85 and 150, t2 gets 150,
t3 gets 610. Gene A = max(2, 1) = 2, gene B = 1.
"""
from oracle.genes import gene_matrix, promoters

TX = [dict(chrom="chr1", left=101, right=200, strand="+", transcript_id="t1", gene_name="A"),
      dict(chrom="chr1", left=151, right=400, strand="+", transcript_id="t2", gene_name="A"),
      dict(chrom="chr1", left=501, right=600, strand="-", transcript_id="t3", gene_name="B")]


def test_promoters():
    assert promoters(TX, 20, 0, True) == ["chr1:80-199", "chr1:130-399", "chr1:500-619"]
    # without the gene body: + strand [left - up, left + down), - strand [right - down, right + up)
    assert promoters(TX, 20, 5, False) == ["chr1:80-105", "chr1:130-155", "chr1:594-619"]


def test_gene_and_transcript_counts():
    args = ([0, 4], [85, 150, 610, 700], [1, 1, 1, 1], False, {"chr1": 1000}, TX)
    ip, ix, d, names = gene_matrix(*args, id_type="transcript", upstream=20)
    assert names == ["t1", "t2", "t3"] and ix.tolist() == [0, 1, 2] and d.tolist() == [2, 1, 1]
    ip, ix, d, names = gene_matrix(*args, id_type="gene", upstream=20)
    assert names == ["A", "B"] and ix.tolist() == [0, 1] and d.tolist() == [2, 1]
