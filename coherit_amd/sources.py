"""Genotype sources feeding the resident panel.

Mirrors the reference's ``GenoBlockSource`` protocol (geno_source.py:59-74):
``n``, ``m``, ``missing_val`` and ``read_block_variant_major(start, count)``
returning a (count, n) int8 block.  ``BedGenoSource`` reads PLINK1 BED files
directly (the reference delegates the file opening to bed_reader,
geno_source.py:81-159); its raw SNP-major rows are transcoded on the GPU.
"""
from __future__ import annotations

import os

import numpy as np

BED_MAGIC = bytes([0x6C, 0x1B, 0x01])


def _read_fam_iids(path: str) -> tuple[str, ...]:
    """Individual IDs (second column) of a PLINK .fam file."""
    ids = []
    with open(path, "r", encoding="utf-8") as fh:
        for line in fh:
            parts = line.split()
            if parts:
                ids.append(parts[1] if len(parts) > 1 else parts[0])
    return tuple(ids)


def _count_lines(path: str) -> int:
    with open(path, "rb") as fh:
        return sum(1 for line in fh if line.strip())


class BedGenoSource:
    """PLINK1 SNP-major BED source (geno_source.py:81-159).

    ``missing_val`` is -127 (bed_reader's int8 missing sentinel, probed by the
    reference at geno_source.py:131-136).  ``read_raw_rows`` returns raw BED
    rows for GPU transcoding; ``read_block_variant_major`` decodes on the
    host for callers that need int8 values.
    """

    def __init__(self, bed_prefix: str, threads: int | None = None, sample_mask=None):
        self._bed_prefix = bed_prefix
        bed_path = bed_prefix + ".bed"
        self._n_full = _count_lines(bed_prefix + ".fam")
        self.m = _count_lines(bed_prefix + ".bim")
        self._threads = max(1, int(threads or (os.cpu_count() or 1)))
        self._bps = (self._n_full + 3) // 4
        size = os.path.getsize(bed_path)
        if size < 3 + self.m * self._bps:
            raise ValueError(f"BED file too small: {size} < {3 + self.m * self._bps}")
        self._raw = np.memmap(bed_path, dtype=np.uint8, mode="r")
        if bytes(self._raw[:3]) != BED_MAGIC:
            raise ValueError("Not a valid SNP-major BED file")
        ids = _read_fam_iids(bed_prefix + ".fam")
        if sample_mask is not None:
            mask = np.asarray(sample_mask, dtype=bool)
            if mask.shape != (self._n_full,):
                raise ValueError(
                    f"sample_mask length mismatch: expected {self._n_full}, got {mask.size}."
                )
            self._sample_idx = np.flatnonzero(mask)
            self.n = int(self._sample_idx.size)
            self.sample_ids = tuple(ids[i] for i in self._sample_idx)
        else:
            self._sample_idx = None
            self.n = self._n_full
            self.sample_ids = ids
        self.missing_val = -127

    @property
    def bytes_per_snp(self) -> int:
        return self._bps

    def kept_sample_index(self) -> np.ndarray:
        if self._sample_idx is None:
            return np.arange(self._n_full, dtype=np.int32)
        return self._sample_idx.astype(np.int32)

    def read_raw_rows(self, source_ids) -> np.ndarray:
        ids = np.asarray(source_ids, dtype=np.int64)
        rows = self._raw[3 : 3 + self.m * self._bps].reshape(self.m, self._bps)
        return np.ascontiguousarray(rows[ids])

    def read_block_variant_major(self, snp_start: int, snp_count: int) -> np.ndarray:
        if snp_count <= 0:
            return np.empty((0, self.n), dtype=np.int8)
        rows = self.read_raw_rows(np.arange(snp_start, snp_start + snp_count))
        kept = self.kept_sample_index().astype(np.int64)
        g = (rows[:, kept >> 2] >> ((kept & 3) << 1).astype(np.uint8)) & 3
        lut = np.array([2, -127, 1, 0], dtype=np.int8)
        return lut[g]

    def close(self) -> None:
        self._raw = None


class ArraySource:
    """In-memory (n, m) int8 genotype matrix as a GenoBlockSource."""

    def __init__(self, block: np.ndarray, missing_val: int = -9):
        self._block = np.asarray(block, dtype=np.int8)
        self.n, self.m = self._block.shape
        self.missing_val = int(missing_val)

    def read_block_variant_major(self, snp_start: int, snp_count: int) -> np.ndarray:
        return np.ascontiguousarray(self._block[:, snp_start : snp_start + snp_count].T)

    def close(self):
        return None


def write_bed(prefix: str, codes: np.ndarray) -> None:
    """Write reference codes (n, m; 0/1/2 = A1 count, 3 = missing) as PLINK1 files."""
    codes = np.asarray(codes, dtype=np.uint8)
    n, m = codes.shape
    to_bed = np.array([3, 2, 0, 1], dtype=np.uint8)  # value -> BED code
    c = to_bed[codes].T
    bps = (n + 3) // 4
    padded = np.zeros((m, 4 * bps), dtype=np.uint8)
    padded[:, :n] = c
    p = padded.reshape(m, bps, 4)
    rows = (p[:, :, 0] | (p[:, :, 1] << 2) | (p[:, :, 2] << 4) | (p[:, :, 3] << 6)).astype(np.uint8)
    with open(prefix + ".bed", "wb") as fh:
        fh.write(BED_MAGIC)
        fh.write(rows.tobytes())
    with open(prefix + ".fam", "w", encoding="utf-8") as fh:
        for i in range(n):
            fh.write(f"f{i} i{i} 0 0 0 -9\n")
    with open(prefix + ".bim", "w", encoding="utf-8") as fh:
        for j in range(m):
            fh.write(f"1 rs{j} 0 {j + 1} A G\n")
