"""Oracle: genotype codes, 2-bit packing, BED transcode, per-SNP statistics.

Test infrastructure only (see ``oracle/__init__.py``).  Integer work here is
exact; the float32 statistics reproduce the reference's NumPy float32
expression order so that the GPU epilogue can be checked bit for bit.
"""
from __future__ import annotations

import dataclasses

import numpy as np

# BED 2-bit code -> reference value (0/1/2 = A1 count, 3 = missing).
# geno_stream.py:764 (`lut = [2, 3, 1, 0]`), PLINK spec: 00 hom-A1, 01 missing,
# 10 het, 11 hom-A2.
BED_LUT = np.array([2, 3, 1, 0], dtype=np.uint8)
MISSING = 3


def int8_pack_lut(missing_val: int) -> np.ndarray:
    """int8 source value (viewed as u8) -> code; geno_stream.py:1263-1270."""
    lut = np.zeros(256, dtype=np.uint8)
    lut[1] = 1
    lut[2] = 2
    lut[int(np.asarray([missing_val], dtype=np.int8).view(np.uint8)[0])] = MISSING
    return lut


def pack_codes(codes: np.ndarray) -> np.ndarray:
    """(n, w) codes in 0..3 -> (n, ceil(w/4)) bytes, SNP 4q+t at bits 2t.

    geno_stream.py:885-904 (`_pack_block_numba`): low bits first, padding 0.
    """
    codes = np.asarray(codes, dtype=np.uint8)
    n, w = codes.shape
    pw = (w + 3) // 4
    padded = np.zeros((n, 4 * pw), dtype=np.uint8)
    padded[:, :w] = codes & 3
    p = padded.reshape(n, pw, 4)
    return (p[:, :, 0] | (p[:, :, 1] << 2) | (p[:, :, 2] << 4) | (p[:, :, 3] << 6)).astype(np.uint8)


def unpack_codes(packed: np.ndarray, width: int | None = None) -> np.ndarray:
    """Inverse of :func:`pack_codes`; kv_impl.py:58-66 (`_unpack_u2_block`)."""
    packed = np.asarray(packed, dtype=np.uint8)
    n, pw = packed.shape
    out = np.stack([(packed >> (2 * t)) & 3 for t in range(4)], axis=2).reshape(n, 4 * pw)
    return out if width is None else out[:, :width]


def int8_stats_varmaj(block_vm: np.ndarray, missing_val: int):
    """Per-SNP int64 (cnt, s1, s2) of a variant-major int8 block.

    geno_stream.py:975-997 (`_stats_and_pack_varmaj_numba` stats loop): values
    equal to ``missing_val`` are skipped; any other value v adds (1, v, v*v).
    """
    b = np.asarray(block_vm, dtype=np.int64)
    obs = b != int(missing_val)
    cnt = obs.sum(axis=1).astype(np.int64)
    s1 = np.where(obs, b, 0).sum(axis=1).astype(np.int64)
    s2 = np.where(obs, b * b, 0).sum(axis=1).astype(np.int64)
    return cnt, s1, s2


def pack_varmaj_int8(block_vm: np.ndarray, missing_val: int) -> np.ndarray:
    """Variant-major int8 block (w, n) -> sample-major packed (n, ceil(w/4)).

    geno_stream.py:999-1005 with the LUT of geno_stream.py:1263-1270.
    """
    lut = int8_pack_lut(missing_val)
    codes = lut[np.asarray(block_vm, dtype=np.int8).view(np.uint8)].T
    return pack_codes(codes)


def bed_bytes_per_snp(n_full: int) -> int:
    return (int(n_full) + 3) // 4


def bed_codes(bed_raw: np.ndarray, n_full: int, m: int, sample_idx=None) -> np.ndarray:
    """Raw BED file bytes -> (n_kept, m) reference codes (cache = source order).

    Magic check geno_stream.py:1628-1629; sample i of SNP j at byte
    ``3 + j*ceil(n_full/4) + (i>>2)``, shift ``(i&3)*2`` (geno_stream.py:1635-1636,
    684-693); LUT geno_stream.py:764.
    """
    raw = np.asarray(bed_raw, dtype=np.uint8)
    if raw.size < 3 or raw[0] != 0x6C or raw[1] != 0x1B or raw[2] != 0x01:
        raise ValueError("Not a valid SNP-major BED file")
    bps = bed_bytes_per_snp(n_full)
    if raw.size < 3 + m * bps:
        raise ValueError(f"BED file too small: {raw.size} < {3 + m * bps}")
    kept = np.arange(n_full) if sample_idx is None else np.asarray(sample_idx, dtype=np.int64)
    rows = raw[3 : 3 + m * bps].reshape(m, bps)
    g = (rows[:, kept >> 2] >> ((kept & 3) << 1).astype(np.uint8)) & 3
    return BED_LUT[g].T.copy()


def codes_to_bed(codes: np.ndarray) -> np.ndarray:
    """(n, m) reference codes -> raw BED bytes (magic + SNP-major rows)."""
    inv = np.zeros(4, dtype=np.uint8)
    inv[BED_LUT] = np.arange(4, dtype=np.uint8)
    c = inv[np.asarray(codes, dtype=np.uint8)].T  # (m, n) BED codes
    m, n = c.shape
    bps = bed_bytes_per_snp(n)
    padded = np.zeros((m, 4 * bps), dtype=np.uint8)
    padded[:, :n] = c
    p = padded.reshape(m, bps, 4)
    rows = (p[:, :, 0] | (p[:, :, 1] << 2) | (p[:, :, 2] << 4) | (p[:, :, 3] << 6)).astype(np.uint8)
    return np.concatenate([np.array([0x6C, 0x1B, 0x01], dtype=np.uint8), rows.reshape(-1)])


def code_stats(codes_snp_major: np.ndarray):
    """int64 (cnt, s1, s2) per SNP from reference codes (rows = SNPs).

    Equivalent to geno_stream.py:783-832 (BED) for codes in 0..3.
    """
    c = np.asarray(codes_snp_major, dtype=np.int64)
    obs = c != MISSING
    cnt = obs.sum(axis=1).astype(np.int64)
    s1 = np.where(obs, c, 0).sum(axis=1).astype(np.int64)
    s2 = np.where(obs, c * c, 0).sum(axis=1).astype(np.int64)
    return cnt, s1, s2


def f32_stats(cnt, s1, s2):
    """float32 (mean, inv_sd) exactly as geno_stream.py:1780-1792."""
    cnt_f = np.asarray(cnt).astype(np.float32)
    s1_f = np.asarray(s1).astype(np.float32)
    s2_f = np.asarray(s2).astype(np.float32)
    denom = np.maximum(cnt_f, np.float32(1.0))
    mean = (s1_f / denom).astype(np.float32)
    var = np.maximum(s2_f / denom - mean * mean, np.float32(0.0))
    valid = (cnt_f > 0.0) & (var > 0.0)
    inv_sd = np.where(
        valid, np.float32(1.0) / np.sqrt(np.maximum(var, np.float32(1e-6))), np.float32(0.0)
    ).astype(np.float32)
    return mean, inv_sd


def standardized_sumsq_from_raw_stats(cnt, s1, s2, mean, inv_sd):
    """geno_stream.py:1481-1496 (FP64 trace mass under an override)."""
    cnt64 = np.asarray(cnt, dtype=np.float64)
    s1_64 = np.asarray(s1, dtype=np.float64)
    s2_64 = np.asarray(s2, dtype=np.float64)
    mean64 = np.asarray(mean, dtype=np.float64)
    inv64 = np.asarray(inv_sd, dtype=np.float64)
    centered = s2_64 - 2.0 * mean64 * s1_64 + cnt64 * mean64 * mean64
    return np.maximum(centered, 0.0) * inv64 * inv64


# --------------------------------------------------------------------------
# geometry (geno_stream.py:368-505)
# --------------------------------------------------------------------------

def normalize_component_block_sizes(m, sizes, allow_empty=False):
    """geno_stream.py:371-392."""
    if sizes is None:
        return (int(m),)
    s = np.asarray(sizes, dtype=np.int64).reshape(-1)
    if s.size == 0:
        raise ValueError("component_block_sizes must contain at least one block.")
    if allow_empty:
        if np.any(s < 0):
            raise ValueError("component_block_sizes must be non-negative.")
    elif np.any(s <= 0):
        raise ValueError("component_block_sizes must be strictly positive.")
    if int(s.sum()) != int(m):
        raise ValueError(f"component_block_sizes must sum to m={int(m)}; got {int(s.sum())}.")
    return tuple(int(x) for x in s.tolist())


def normalize_component_variant_indices(m, groups):
    """geno_stream.py:395-437: sort/unique each group, concatenate."""
    groups = list(groups or [])
    if not groups:
        raise ValueError("component_variant_indices must contain at least one component.")
    norm = []
    for k, g in enumerate(groups):
        arr = np.asarray(g, dtype=np.int64).reshape(-1)
        if arr.size == 0:
            norm.append(arr)
            continue
        if np.any((arr < 0) | (arr >= int(m))):
            raise IndexError(f"component_variant_indices[{k}] contains out-of-range SNP indices for m={int(m)}.")
        u = np.unique(arr)
        if u.size != arr.size:
            raise ValueError(f"component_variant_indices[{k}] contains duplicate SNP indices.")
        norm.append(u)
    c2s = np.concatenate(norm) if norm else np.empty((0,), np.int64)
    if c2s.size == 0:
        raise ValueError("component_variant_indices must cover at least one SNP.")
    if c2s.size > int(m):
        raise ValueError(f"component_variant_indices assign {c2s.size} SNPs, which exceeds source m={int(m)}.")
    if np.unique(c2s).size != c2s.size:
        raise ValueError("component_variant_indices must be disjoint across components.")
    if c2s.size == int(m) and not np.array_equal(np.sort(c2s), np.arange(int(m))):
        raise ValueError("component_variant_indices must cover every SNP exactly once.")
    return tuple(int(a.size) for a in norm), c2s


def call_geometry(m, call_width, component_block_sizes):
    """geno_stream.py:469-505: calls never straddle components."""
    sizes = normalize_component_block_sizes(m, component_block_sizes, allow_empty=True)
    starts, widths, comp_ids, offs = [], [], [], [0]
    s = 0
    for cid, size in enumerate(sizes):
        rem = int(size)
        while rem > 0:
            w = min(int(call_width), rem)
            starts.append(s)
            widths.append(w)
            comp_ids.append(cid)
            s += w
            rem -= w
        offs.append(len(starts))
    return (np.asarray(starts, np.int64), np.asarray(widths, np.int32),
            np.asarray(comp_ids, np.int32), np.asarray(offs, np.int32))


@dataclasses.dataclass
class Atoms:
    eff: float
    mean_diag: float
    component_eff: np.ndarray
    component_mean_diag: np.ndarray


def atoms(inv_sd, trace_mass, n, component_sizes) -> Atoms:
    """eff / mean-diagonal atoms: geno_stream.py:1409-1437."""
    inv_sd = np.asarray(inv_sd, dtype=np.float32)
    tm = np.asarray(trace_mass, dtype=np.float64)
    valid = inv_sd > 0.0
    eff = float(np.count_nonzero(valid))
    mean_diag = float(np.sum(tm[valid], dtype=np.float64) / (float(n) * eff)) if eff > 0 else 0.0
    G = len(component_sizes)
    ceff = np.zeros(G, np.float32)
    cdiag = np.zeros(G, np.float32)
    offs = np.concatenate([[0], np.cumsum(component_sizes)]).astype(np.int64)
    for g in range(G):
        a, b = offs[g], offs[g + 1]
        ceff[g] = float(np.count_nonzero(valid[a:b]))
        if ceff[g] > 0:
            cdiag[g] = float(np.sum(tm[a:b][valid[a:b]], dtype=np.float64) / (float(n) * float(ceff[g])))
    return Atoms(eff, mean_diag, ceff, cdiag)


@dataclasses.dataclass
class OraclePanel:
    """Cache-order genotype codes plus the statistics the operator uses."""

    codes: np.ndarray            # (n, m) uint8 in 0..3, cache order
    mean: np.ndarray             # (m,) f32
    inv_sd: np.ndarray           # (m,) f32
    cnt: np.ndarray              # (m,) int64
    component_sizes: tuple
    atoms: Atoms

    @property
    def n(self):
        return int(self.codes.shape[0])

    @property
    def m(self):
        return int(self.codes.shape[1])

    @property
    def component_offsets(self):
        return np.concatenate([[0], np.cumsum(self.component_sizes)]).astype(np.int64)

    @classmethod
    def from_codes(cls, codes, component_sizes=None, override=None):
        codes = np.asarray(codes, dtype=np.uint8)
        n, m = codes.shape
        sizes = normalize_component_block_sizes(m, component_sizes)
        cnt, s1, s2 = code_stats(codes.T)
        if override is None:
            mean, inv = f32_stats(cnt, s1, s2)
            tm = cnt.astype(np.float64)
        else:
            mean = np.asarray(override[0], np.float32)
            inv = np.asarray(override[1], np.float32)
            tm = standardized_sumsq_from_raw_stats(cnt, s1, s2, mean, inv)
        return cls(codes, mean, inv, cnt, sizes, atoms(inv, tm, n, sizes))

    def standardized(self, dtype=np.float64):
        """Z = (g_imp - mean) * inv_sd with missing -> mean (kv_impl.py:77-91)."""
        g = self.codes.astype(dtype)
        mu = self.mean.astype(dtype)[None, :]
        diff = np.where(self.codes == MISSING, 0.0, g - mu).astype(dtype)
        return diff * self.inv_sd.astype(dtype)[None, :]
