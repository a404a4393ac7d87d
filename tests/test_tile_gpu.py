"""Path A on the GPU: libsnapgpu tile kernels vs the oracle, bit-exact."""
import json
import os

import numpy as np
import pytest

from oracle import tile as otile

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _frags(torch, cells, paired=True):
    from snapatac2_amd._device import DeviceFragments
    cell_ptr = np.zeros(len(cells) + 1, dtype=np.int64)
    cell_ptr[1:] = np.cumsum([len(c) for c in cells])
    pos = np.asarray([p for c in cells for p, _ in c], dtype=np.uint32)
    ln = np.asarray([l for c in cells for _, l in c], dtype=np.uint32)
    dev = torch.device("cuda")
    t = lambda a: torch.from_numpy(a.view(np.int32)).to(dev)
    vals = t(ln) if paired else torch.from_numpy(ln.astype(np.int32)).to(dev)
    return DeviceFragments(torch.from_numpy(cell_ptr).to(dev), t(pos), vals, 0, paired), \
        cell_ptr, pos, ln


def _tile(torch, cells, cs, bin_size, excl, lo, hi, strategy):
    from snapatac2_amd.pp import make_tile_matrix
    paired = strategy != "single"
    fr, cp, pos, ln = _frags(torch, cells, paired)
    X, names = make_tile_matrix(fr, cs, bin_size, excl, lo, hi,
                                strategy if paired else "paired-insertion")
    return X.to_scipy(), names, (cp, pos, ln)


def _same(a, b):
    assert a.shape == b.shape, (a.shape, b.shape)
    assert np.array_equal(a.indptr, b.indptr)
    assert np.array_equal(a.indices.astype(np.int64), b.indices.astype(np.int64))
    assert np.array_equal(a.data.astype(np.uint32), b.data.astype(np.uint32))


def test_hand_derived_cases(torch_gpu):
    g = json.load(open(os.path.join(GOLD, "tile_cases.json")))
    cs = dict((c, l) for c, l in g["chrom_sizes"])
    cells = [[tuple(f) for f in c] for c in g["cells"]]
    for case in g["cases"]:
        m, _, _ = _tile(torch_gpu, cells, cs, g["bin_size"], case["exclude"], case["min"],
                        case["max"], case["strategy"])
        assert m.shape == (3, case["n_bins"])
        rows = [[[int(j), int(v)] for j, v in zip(m.indices[m.indptr[r]:m.indptr[r + 1]],
                                                  m.data[m.indptr[r]:m.indptr[r + 1]])]
                for r in range(3)]
        assert rows == case["rows"], case


def test_out_of_range_raises(torch_gpu):
    from snapatac2_amd._lib import PanicException
    g = json.load(open(os.path.join(GOLD, "tile_cases.json")))
    cs = dict((c, l) for c, l in g["chrom_sizes"])
    with pytest.raises(PanicException, match="out of range"):
        _tile(torch_gpu, [[tuple(f) for f in g["panic_cell"]]], cs, 100, [], None, None,
              "paired-insertion")


def test_genome_index_vectors_through_kernel(torch_gpu):
    # every base of the genome.rs test genome, one single-end read per cell
    g = json.load(open(os.path.join(GOLD, "genome_index.json")))
    cs = dict((c, l) for c, l in g["chrom_sizes"])
    total = sum(cs.values())
    cells = [[(i, 1)] for i in range(total)]
    for step in (1, 2, 3):
        m, names, _ = _tile(torch_gpu, cells, cs, step, [], None, None, "single")
        cols = m.indices.astype(np.int64)
        assert np.array_equal(m.indptr, np.arange(total + 1))
        exp = [otile.get_position_rev(cs, step, *otile.get_position(cs, i)) for i in range(total)]
        assert cols.tolist() == exp
    for i1, i2 in g["coarsed_step2"]:
        m, _, _ = _tile(torch_gpu, [[(i1, 1)]], cs, 2, [], None, None, "single")
        assert int(m.indices[0]) == i2


@pytest.mark.parametrize("strategy", ["insertion", "fragment", "paired-insertion", "single"])
def test_random_vs_oracle(torch_gpu, strategy):
    from test_tile_oracle import random_cells
    rng = np.random.default_rng(17)
    cs = {"chr1": 50000, "chr10": 12345, "chr2": 30000, "chrM": 3000, "chrX": 25000}
    cells = random_cells(rng, cs, 300, max_frags=400, max_len=1200)
    cells[5] = []
    for bin_size, excl, lo, hi in [(500, ["chrM", "chrY"], None, None), (37, ["chrX"], 10, 300),
                                   (1, [], None, None), (5000, ["chr1", "chr10"], 50, None)]:
        m, names, (cp, pos, ln) = _tile(torch_gpu, cells, cs, bin_size, excl, lo, hi, strategy)
        o = otile.tile_matrix(cp, pos, ln, cs, bin_size, excl, lo, hi, strategy)
        _same(m, o)
        assert len(names) == m.shape[1]


def test_dense_duplicates_and_long_spans(torch_gpu):
    # many fragments in one bin (pending list grows), and spans that need the
    # larger capacity levels (bin 1 with 3 kb fragments)
    cs = {"chr1": 100000}
    hot = [(1000 + (i % 7), 300 + (i % 50)) for i in range(5000)]
    longf = [(i * 10, 3000 + i % 3) for i in range(200)]
    cells = [sorted(hot), sorted(longf), [(5, 2)] * 1000]
    for strategy in ("insertion", "paired-insertion", "fragment"):
        for bin_size in (1, 100, 500):
            m, _, (cp, pos, ln) = _tile(torch_gpu, cells, cs, bin_size, [], None, None, strategy)
            o = otile.tile_matrix(cp, pos, ln, cs, bin_size, [], None, None, strategy)
            _same(m, o)


def test_synthetic_2k_cells_all_strategies(torch_gpu):
    from snapatac2_amd import synth
    from snapatac2_amd.genome import synthetic_genome
    fr = synth.generate_chunk(synth.GeneratorTables(0), 0, 2000)
    h = fr.to_scipy()
    cs = synthetic_genome()
    from snapatac2_amd.pp import make_tile_matrix
    for strategy in ("paired-insertion", "insertion", "fragment"):
        X, names = make_tile_matrix(fr, cs, 500, ["chrM", "chrY", "M", "Y"], None, None, strategy)
        o = otile.tile_matrix(h.indptr, h.indices, h.data, cs, 500, ["chrM", "chrY"], None,
                              None, strategy)
        _same(X.to_scipy(), o)
        assert X.shape[1] == 6_062_095 - 0  # chr1-22, X at 500 bp


def test_add_tile_matrix_dropin_and_determinism(torch_gpu):
    import snapatac2_amd as snap
    from snapatac2_amd import synth
    ad = synth.synthetic_anndata(300, seed=3)
    snap.pp.add_tile_matrix(ad)
    X1 = ad.X.to_scipy()
    out = snap.pp.add_tile_matrix(ad, inplace=False, bin_size=1000, exclude_chroms="chrX",
                                  counting_strategy="insertion")
    assert out.shape[0] == 300 and out.shape[1] == len(out.var_names)
    ad2 = synth.synthetic_anndata(300, seed=3)
    snap.pp.add_tile_matrix(ad2)
    _same(X1, ad2.X.to_scipy())
    assert ad.var_names[0] == "chr1:0-500"
    with pytest.raises(RuntimeError, match="Counting strategy"):
        snap.pp.add_tile_matrix(ad, counting_strategy="bogus")
    with pytest.raises(Exception, match="Value type"):
        snap.pp.add_tile_matrix(ad, value_type="bogus")
    empty = snap.AnnData(n_obs=2, n_vars=0)
    with pytest.raises(RuntimeError, match="must be present"):
        snap.pp.add_tile_matrix(empty)


def test_select_features_matches_numpy(torch_gpu):
    import snapatac2_amd as snap
    from snapatac2_amd import synth
    from snapatac2_amd.pp._basic import _find_most_accessible_features
    ad = synth.synthetic_anndata(500, seed=1)
    snap.pp.add_tile_matrix(ad)
    snap.pp.select_features(ad, n_features=20000)
    Xh = ad.X.to_scipy()
    count = np.ravel(Xh.sum(axis=0)).astype(np.float64)
    assert np.array_equal(ad.var["count"].to_numpy(), count)
    sel = np.zeros(Xh.shape[1], dtype=bool)
    sel[_find_most_accessible_features(count, 0.005, 0.005, 20000)] = True
    assert np.array_equal(ad.var["selected"].to_numpy(), sel)
    assert sel.sum() == 20000
