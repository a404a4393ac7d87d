"""Cell x peak matrix on the GPU against the brute-force oracle, bit-exact:
row pointers, region columns (ascending per row) and counts, for the three
counting strategies, size filters, single-end reads, overlapping, duplicated
and out-of-genome regions."""
import numpy as np
import pytest

from oracle.peaks import peak_matrix

pytestmark = pytest.mark.gpu


def _frags(torch, cell_ptr, pos, val, paired):
    from snapatac2_amd._device import DeviceFragments
    dev = torch.device("cuda")
    cp = torch.from_numpy(np.asarray(cell_ptr, dtype=np.int64)).to(dev)
    p = torch.from_numpy(np.asarray(pos, dtype=np.uint32).view(np.int32)).to(dev)
    v = np.asarray(val)
    v = v.astype(np.uint32).view(np.int32) if paired else v.astype(np.int32)
    return DeviceFragments(cp, p, torch.from_numpy(v).to(dev), 0, paired)


def _check(torch, cell_ptr, pos, val, paired, genome, peaks, strategy, lo=None, hi=None):
    from snapatac2_amd.pp._peaks import make_peak_counts
    X, names = make_peak_counts(_frags(torch, cell_ptr, pos, val, paired), genome, peaks, strategy, lo, hi)
    ip, ix, d = peak_matrix(cell_ptr, pos, val, paired, genome, peaks, strategy, lo, hi)
    assert np.array_equal(X.indptr.cpu().numpy(), ip)
    assert np.array_equal(X.indices.cpu().numpy().astype(np.int64), ix)
    assert np.array_equal(X.data.cpu().numpy().view(np.uint32), d)
    return X, names


GENOME = {"chr1": 100, "chr2": 50}
PEAKS = ["chr1:10-20", "chr1:15-30", "chr2:0-10", "chr1:10-20", "chr3:0-5"]


@pytest.mark.parametrize("strategy", ["paired-insertion", "insertion", "fragment"])
def test_known_case(torch_gpu, strategy):
    X, names = _check(torch_gpu, [0, 1, 3], [12, 19, 105], [6, 6, 3], True, GENOME, PEAKS, strategy)
    assert names == ["chr1:10-20", "chr1:15-30", "chr2:0-10", "chr3:0-5"]
    _check(torch_gpu, [0, 1, 3], [12, 19, 105], [6, 6, 3], True, GENOME, PEAKS, strategy, 4, None)
    _check(torch_gpu, [0, 1, 3], [12, 19, 105], [6, -6, 3], False, GENOME, PEAKS, strategy)


def _synthetic_peaks(rng, chrom_sizes, centres, n):
    chroms = list(chrom_sizes)
    base = np.concatenate([[0], np.cumsum([chrom_sizes[c] for c in chroms])])
    out = []
    for g in rng.choice(centres, n, replace=False):
        k = int(np.searchsorted(base, g, side="right")) - 1
        c = int(g) - int(base[k])
        w = int(rng.integers(50, 1200))
        s = max(c - w // 2, 0)
        e = c + w // 2 + int(rng.integers(1, 300))  # may pass the chromosome end
        out.append(f"{chroms[k]}:{s}-{e}")
    out += out[:20]  # duplicates
    out += ["chrUn:0-1000", "chrM:0-16000"]  # chromosomes outside the genome
    rng.shuffle(out)
    return out


@pytest.mark.parametrize("strategy,lo,hi", [("paired-insertion", None, None), ("insertion", 50, 600),
                                            ("fragment", None, 300)])
def test_synthetic_cells(torch_gpu, strategy, lo, hi):
    from snapatac2_amd import synth
    from snapatac2_amd.genome import synthetic_genome
    cs = synthetic_genome()
    tab = synth.GeneratorTables(0)
    fr = synth.generate_chunk(tab, 0, 30)
    h = fr.to_scipy()
    rng = np.random.default_rng(7)
    peaks = _synthetic_peaks(rng, cs, tab.peak_centre.astype(np.int64), 400)
    X, names = _check(torch_gpu, h.indptr, h.indices, h.data, True, cs, peaks, strategy, lo, hi)
    assert X.nnz > 200


def test_single_end_random(torch_gpu):
    rng = np.random.default_rng(3)
    genome = {"chr1": 5000, "chr2": 3000}
    cell_ptr = np.concatenate([[0], np.cumsum(rng.integers(0, 300, 40))])
    nf = int(cell_ptr[-1])
    pos = rng.integers(0, 8000, nf)
    val = rng.integers(1, 200, nf) * rng.choice([-1, 1], nf)
    peaks = [f"chr{rng.integers(1, 3)}:{s}-{s + rng.integers(1, 400)}" for s in rng.integers(0, 5000, 150)]
    _check(torch_gpu, cell_ptr, pos, val, False, genome, peaks, "paired-insertion")


def test_make_peak_matrix_api(torch_gpu):
    import pandas as pd
    from snapatac2_amd import synth
    from snapatac2_amd.pp import make_peak_matrix
    ad = synth.synthetic_anndata(20)
    tab = synth.GeneratorTables(0)
    peaks = _synthetic_peaks(np.random.default_rng(1), synth.synthetic_genome(), tab.peak_centre.astype(np.int64), 100)
    ad.uns["peaks"] = pd.DataFrame({"Peaks": peaks})
    out = make_peak_matrix(ad)
    assert out.shape[0] == 20 and out.n_vars == len(set(peaks))
    h = ad.obsm["fragment_paired"].to_scipy()
    ip, ix, d = peak_matrix(h.indptr, h.indices, h.data, True, synth.synthetic_genome(), peaks)
    assert np.array_equal(out.X.indptr.cpu().numpy(), ip)
    with pytest.raises(RuntimeError):
        make_peak_matrix(ad, use_rep="peaks", peak_file="x.bed")


@pytest.mark.parametrize("strategy", ["paired-insertion", "fragment"])
def test_many_cells_against_cpp(torch_gpu, strategy):
    """2,000 synthetic cells x 50,000 fixed-width regions at the generator's peak
    centres, against the C++ restatement (another region search)."""
    from oracle.peaks import peak_matrix_cpp
    from snapatac2_amd import synth
    from snapatac2_amd.genome import synthetic_genome
    from snapatac2_amd.pp._peaks import make_peak_counts
    cs = synthetic_genome()
    chroms = list(cs)
    base = np.concatenate([[0], np.cumsum([cs[c] for c in chroms])])
    tab = synth.GeneratorTables(0)
    centres = np.sort(tab.peak_centre.astype(np.int64))
    k = np.searchsorted(base, centres, side="right") - 1
    peaks = [f"{chroms[a]}:{max(int(c - base[a]) - 250, 0)}-{int(c - base[a]) + 250}" for a, c in zip(k, centres)]
    fr = synth.generate_chunk(tab, 0, 2000)
    X, _ = make_peak_counts(fr, cs, peaks, strategy)
    h = fr.to_scipy()
    ip, ix, d = peak_matrix_cpp(h.indptr, h.indices, h.data, True, cs, peaks, strategy)
    assert np.array_equal(X.indptr.cpu().numpy(), ip)
    assert np.array_equal(X.indices.cpu().numpy().astype(np.int64), ix)
    assert np.array_equal(X.data.cpu().numpy().view(np.uint32), d)
    assert X.nnz > 100_000
