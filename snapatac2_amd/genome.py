"""Chromosome tables and the bin index of the tile matrix.

Mirrors GenomeBaseIndex (snapatac2-core/src/genome.rs:286-487): positions are
0-based indices into the concatenation of the chromosomes in their stored
order; `with_step` gives each chromosome ceil(length / step) bins, and bin i
is named "chrom:start-end" with the last bin clipped to the chromosome end
(genome.rs:326-338, 437-466).
"""
from __future__ import annotations

import numpy as np

# hg38 lengths (snapatac2-python/python/snapatac2/genome.py:126-134).
HG38 = {
    "chr1": 248956422, "chr2": 242193529, "chr3": 198295559, "chr4": 190214555,
    "chr5": 181538259, "chr6": 170805979, "chr7": 159345973, "chr8": 145138636,
    "chr9": 138394717, "chr10": 133797422, "chr11": 135086622, "chr12": 133275309,
    "chr13": 114364328, "chr14": 107043718, "chr15": 101991189, "chr16": 90338345,
    "chr17": 83257441, "chr18": 80373285, "chr19": 58617616, "chr20": 64444167,
    "chr21": 46709983, "chr22": 50818468, "chrX": 156040895, "chrY": 57227415,
    "chrM": 16569,
}


def sorted_chrom_sizes(chrom_sizes: dict[str, int]) -> dict[str, int]:
    """Chromosomes in the order import_fragments stores them: it receives them
    as a BTreeMap, i.e. sorted by name (snapatac2-python/src/preprocessing.rs:95,126)."""
    return {k: int(chrom_sizes[k]) for k in sorted(chrom_sizes)}


def synthetic_genome() -> dict[str, int]:
    """The 24-chromosome genome of the synthetic workloads: hg38 chr1-22, X, Y
    (chrM omitted), in stored (lexicographic) order; 3,088,269,832 bases."""
    return sorted_chrom_sizes({k: v for k, v in HG38.items() if k != "chrM"})


def reference_sequences(chrom_sizes: dict[str, int]):
    """The uns['reference_sequences'] table import_fragments writes (import.rs:106-108)."""
    import pandas as pd
    return pd.DataFrame({
        "reference_seq_name": list(chrom_sizes.keys()),
        "reference_seq_length": np.asarray(list(chrom_sizes.values()), dtype=np.uint64),
    })


def read_chrom_sizes(uns) -> dict[str, int]:
    """SnapData::read_chrom_sizes (snapatac2-core/src/feature_count/mod.rs:57-71)."""
    if "reference_sequences" not in uns:
        raise KeyError("key 'reference_sequences' is not present in the '.uns'")
    df = uns["reference_sequences"]
    if isinstance(df, dict):
        names, lens = df["reference_seq_name"], df["reference_seq_length"]
    else:
        names, lens = df["reference_seq_name"].to_numpy(), df["reference_seq_length"].to_numpy()
    return {str(n): int(l) for n, l in zip(names, lens)}


class BinIndex:
    """Lazy var_names of a tile matrix: bin i -> "chrom:start-end"."""

    def __init__(self, chrom_sizes: dict[str, int], step: int):
        self.chroms = list(chrom_sizes.keys())
        self.lengths = np.asarray(list(chrom_sizes.values()), dtype=np.int64)
        self.step = int(step)
        nb = (self.lengths + self.step - 1) // self.step
        self.bin_accum = np.cumsum(nb)

    def __len__(self) -> int:
        return int(self.bin_accum[-1]) if len(self.bin_accum) else 0

    def region(self, i: int) -> tuple[str, int, int]:
        j = int(np.searchsorted(self.bin_accum, i, side="right"))
        prev = 0 if j == 0 else int(self.bin_accum[j - 1])
        start = (i - prev) * self.step
        end = min(start + self.step, int(self.lengths[j]))
        return self.chroms[j], start, end

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[k] for k in range(*i.indices(len(self)))]
        if isinstance(i, (list, np.ndarray)):
            return [self[int(k)] for k in i]
        if i < 0:
            i += len(self)
        c, s, e = self.region(int(i))
        return f"{c}:{s}-{e}"

    def ranges(self) -> tuple[list[str], np.ndarray, np.ndarray, np.ndarray]:
        """(chromosome names, chromosome code, start, end) of every bin."""
        nb = (self.lengths + self.step - 1) // self.step
        code = np.repeat(np.arange(len(self.chroms), dtype=np.int32), nb)
        first = np.concatenate([[0], self.bin_accum[:-1]]).astype(np.int64) if len(nb) else np.zeros(0, np.int64)
        start = (np.arange(len(self), dtype=np.int64) - np.repeat(first, nb)) * self.step
        end = np.minimum(start + self.step, np.repeat(self.lengths, nb))
        return list(self.chroms), code, start, end

    def to_numpy(self) -> np.ndarray:
        out = []
        for c, L in zip(self.chroms, self.lengths):
            starts = np.arange(0, int(L), self.step, dtype=np.int64)
            ends = np.minimum(starts + self.step, int(L))
            out.append(np.char.add(np.char.add(f"{c}:", starts.astype(str)),
                                   np.char.add("-", ends.astype(str))))
        return np.concatenate(out) if out else np.array([], dtype=str)

    def __iter__(self):
        return iter(self.to_numpy())
