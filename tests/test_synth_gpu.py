"""The generator's ground-truth cell types (snapgpu_synth_cell_types) equal a
numpy restatement of the generator's counter-based draw, and the cells' peak
fragments follow their type: a cell's fragments near peaks fall mostly on the
peaks its own type favours."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(z):
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _types(seed, cells, n_types):
    with np.errstate(over="ignore"):
        inner = _mix64(M64 * np.uint64(0x2545F4914F6CDD1D) + np.uint64(3))
        h = _mix64(np.uint64(seed) ^ _mix64(cells.astype(np.uint64) * np.uint64(0x100000001B3) ^ inner))
    u = ((h >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)
    return (u * n_types).astype(np.int32)


@pytest.mark.parametrize("seed,first", [(0, 0), (7, 123_456)])
def test_cell_types_match_restatement(torch_gpu, seed, first):
    from snapatac2_amd import synth
    t = synth.cell_types(50_000, seed=seed, first_cell=first).cpu().numpy()
    assert np.array_equal(t, _types(seed, np.arange(first, first + 50_000), 10))
    counts = np.bincount(t, minlength=10)
    assert counts.min() > 4_000  # uniform over 10 types


def test_peak_fragments_follow_the_type(torch_gpu):
    from snapatac2_amd import synth
    tab = synth.GeneratorTables(0)
    n = 400
    fr = synth.generate_chunk(tab, 0, n)
    t = synth.cell_types(n).cpu().numpy()
    h = fr.to_scipy()
    centres = np.sort(tab.peak_centre.astype(np.int64))
    order = np.argsort(tab.peak_centre.astype(np.int64), kind="stable")
    acts = [np.diff(np.concatenate([[0.0], tab.type_cdf[j].astype(np.float64)])) for j in range(10)]
    hits = 0
    for r in range(n):
        pos = h.indices[h.indptr[r]:h.indptr[r + 1]].astype(np.int64)
        j = np.clip(np.searchsorted(centres, pos), 1, centres.size - 1)
        near = np.where(np.abs(centres[j] - pos) < np.abs(centres[j - 1] - pos), j, j - 1)
        close = np.abs(centres[near] - pos) < 600
        peaks = order[near[close]]
        # mean activity of the hit peaks under each type: the cell's own type wins
        score = [a[peaks].mean() for a in acts]
        hits += int(np.argmax(score) == t[r])
    assert hits >= 0.95 * n, hits
