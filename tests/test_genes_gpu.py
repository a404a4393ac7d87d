"""Gene matrix on the GPU against the oracle, bit-exact: hand case, synthetic
cells with a written GTF (gene and transcript level, duplicated transcript
coordinates, both strands, with and without gene bodies), and the
make_gene_matrix API on GTF and GFF files."""
import gzip

import numpy as np
import pytest

from oracle.genes import gene_matrix

pytestmark = pytest.mark.gpu


def _frags(torch, cell_ptr, pos, val, paired):
    from snapatac2_amd._device import DeviceFragments
    dev = torch.device("cuda")
    v = np.asarray(val)
    v = v.astype(np.uint32).view(np.int32) if paired else v.astype(np.int32)
    return DeviceFragments(torch.from_numpy(np.asarray(cell_ptr, dtype=np.int64)).to(dev),
                           torch.from_numpy(np.asarray(pos, dtype=np.uint32).view(np.int32)).to(dev),
                           torch.from_numpy(v).to(dev), 0, paired)


def _check(torch, cell_ptr, pos, val, paired, genome, tx, **kw):
    from snapatac2_amd.pp._genes import gene_counts
    X, names, _ = gene_counts(_frags(torch, cell_ptr, pos, val, paired), genome, tx,
                              kw.get("id_type", "gene"), kw.get("upstream", 2000), kw.get("downstream", 0),
                              kw.get("include_gene_body", True), kw.get("strategy", "paired-insertion"))
    ip, ix, d, onames = gene_matrix(cell_ptr, pos, val, paired, genome, tx, **kw)
    assert names == onames
    assert X.shape == (len(ip) - 1, len(onames))
    assert np.array_equal(X.indptr.cpu().numpy(), ip)
    assert np.array_equal(X.indices.cpu().numpy().astype(np.int64), ix)
    assert np.array_equal(X.data.cpu().numpy().view(np.uint32), d)
    return X


TX = [dict(chrom="chr1", left=101, right=200, strand="+", transcript_id="t1", gene_name="A"),
      dict(chrom="chr1", left=151, right=400, strand="+", transcript_id="t2", gene_name="A"),
      dict(chrom="chr1", left=501, right=600, strand="-", transcript_id="t3", gene_name="B")]


@pytest.mark.parametrize("id_type", ["gene", "transcript"])
def test_known_case(torch_gpu, id_type):
    X = _check(torch_gpu, [0, 4], [85, 150, 610, 700], [1, 1, 1, 1], False, {"chr1": 1000}, TX,
               id_type=id_type, upstream=20)
    assert X.data.cpu().numpy().tolist() == ([2, 1] if id_type == "gene" else [2, 1, 1])


def _synthetic_transcripts(rng, chrom_sizes, centres, n_genes=120):
    chroms = list(chrom_sizes)
    base = np.concatenate([[0], np.cumsum([chrom_sizes[c] for c in chroms])])
    tx = []
    for g, c in enumerate(rng.choice(centres, n_genes, replace=False)):
        k = int(np.searchsorted(base, c, side="right")) - 1
        loc = int(c) - int(base[k])
        strand = "+" if rng.random() < 0.5 else "-"
        for t in range(int(rng.integers(1, 5))):
            left = max(loc + int(rng.integers(-3000, 3000)), 1)
            right = left + int(rng.integers(100, 20000))
            tx.append(dict(chrom=chroms[k], left=left, right=right, strand=strand,
                           transcript_id=f"T{g}.{t}", gene_name=f"G{g}", gene_id=f"E{g}"))
        if g % 10 == 0:  # a transcript with the coordinates of the previous one
            tx.append(dict(tx[-1], transcript_id=f"T{g}.dup"))
    return tx


@pytest.mark.parametrize("id_type,body,strategy", [("gene", True, "paired-insertion"),
                                                   ("transcript", True, "insertion"),
                                                   ("gene", False, "fragment")])
def test_synthetic_cells(torch_gpu, id_type, body, strategy):
    from snapatac2_amd import synth
    from snapatac2_amd.genome import synthetic_genome
    cs = synthetic_genome()
    tab = synth.GeneratorTables(0)
    h = synth.generate_chunk(tab, 0, 30).to_scipy()
    tx = _synthetic_transcripts(np.random.default_rng(4), cs, tab.peak_centre.astype(np.int64))
    X = _check(torch_gpu, h.indptr, h.indices, h.data, True, cs, tx, id_type=id_type, upstream=2000,
               downstream=500, include_gene_body=body, strategy=strategy)
    assert X.nnz > 100


def _write_annotation(path, tx, gff=False):
    lines = ["#!genome-build test"]
    for t in tx:
        if gff:
            attrs = f"ID={t['transcript_id']};transcript_id={t['transcript_id']};gene_name={t['gene_name']};gene_id={t['gene_id']}"
        else:
            attrs = f'gene_id "{t["gene_id"]}"; transcript_id "{t["transcript_id"]}"; gene_name "{t["gene_name"]}";'
        lines.append("\t".join([t["chrom"], "test", "gene", str(t["left"]), str(t["right"] + 50), ".", t["strand"], ".", attrs]))
        lines.append("\t".join([t["chrom"], "test", "transcript", str(t["left"]), str(t["right"]), ".", t["strand"], ".", attrs]))
        lines.append("\t".join([t["chrom"], "test", "exon", str(t["left"]), str(t["left"] + 10), ".", t["strand"], ".", attrs]))
    data = ("\n".join(lines) + "\n").encode()
    if str(path).endswith(".gz"):
        data = gzip.compress(data)
    path.write_bytes(data)


@pytest.mark.parametrize("fname", ["anno.gtf", "anno.gtf.gz", "anno.gff3"])
def test_make_gene_matrix_api(torch_gpu, tmp_path, fname):
    from snapatac2_amd import synth
    from snapatac2_amd.genome import synthetic_genome
    from snapatac2_amd.pp import make_gene_matrix
    ad = synth.synthetic_anndata(20)
    tab = synth.GeneratorTables(0)
    tx = _synthetic_transcripts(np.random.default_rng(2), synthetic_genome(), tab.peak_centre.astype(np.int64), 40)
    path = tmp_path / fname
    _write_annotation(path, tx, gff=fname.endswith("gff3"))
    out = make_gene_matrix(ad, path)
    h = ad.obsm["fragment_paired"].to_scipy()
    ip, ix, d, names = gene_matrix(h.indptr, h.indices, h.data, True, synthetic_genome(), tx)
    assert list(out.var_names) == names
    assert np.array_equal(out.X.indptr.cpu().numpy(), ip)
    assert np.array_equal(out.X.data.cpu().numpy().view(np.uint32), d)
    tr = make_gene_matrix(ad, path, id_type="transcript")
    assert list(tr.var["gene_name"]) == [t["gene_name"] for t in tx]
