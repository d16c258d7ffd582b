"""Oracle: CPU twin of the counter-based synthetic genotype generator.

Test infrastructure only.  The HIP generator in ``coherit_amd/csrc/synth.hip``
must produce the same codes byte for byte; this NumPy version defines them.

Genotypes follow the workload description: per-SNP thresholds derived from
the minor-allele frequency p (``P(g=0)=(1-p)^2``, ``P(g=2)=p^2``, HWE) and an
independent missingness draw.  Randomness is Philox4x32-10 (Salmon et al.,
SC'11), keyed by the seed and addressed by (sample quad, SNP, stream).
"""
from __future__ import annotations

import numpy as np

PHILOX_M0 = np.uint64(0xD2511F53)
PHILOX_M1 = np.uint64(0xCD9E8D57)
PHILOX_W0 = np.uint32(0x9E3779B9)
PHILOX_W1 = np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32(c0, c1, c2, c3, k0, k1, rounds=10):
    """Vectorised Philox4x32 on uint32 arrays; returns four uint32 arrays."""
    c0 = np.asarray(c0, np.uint32).astype(np.uint64)
    c1 = np.asarray(c1, np.uint32).astype(np.uint64)
    c2 = np.asarray(c2, np.uint32).astype(np.uint64)
    c3 = np.asarray(c3, np.uint32).astype(np.uint64)
    k0 = np.uint64(np.uint32(k0))
    k1 = np.uint64(np.uint32(k1))
    for _ in range(rounds):
        p0 = PHILOX_M0 * c0
        p1 = PHILOX_M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK32
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & MASK32, lo1, (hi0 ^ c3 ^ k1) & MASK32, lo0
        k0 = (k0 + np.uint64(PHILOX_W0)) & MASK32
        k1 = (k1 + np.uint64(PHILOX_W1)) & MASK32
    return tuple(x.astype(np.uint32) for x in (c0, c1, c2, c3))


def thresholds(p, miss_rate):
    """Per-SNP uint32 thresholds (t0, t1) and the shared missing threshold."""
    p = np.asarray(p, dtype=np.float64)
    two32 = 4294967296.0
    t0 = np.minimum(np.floor((1.0 - p) ** 2 * two32), two32 - 1).astype(np.uint32)
    t1 = np.minimum(np.floor((1.0 - p * p) * two32), two32 - 1).astype(np.uint32)
    tm = np.uint32(min(np.floor(float(miss_rate) * two32), two32 - 1))
    return t0, t1, tm


def genotype_codes(seed, n, snp_ids, t0, t1, tmiss):
    """(n, len(snp_ids)) reference codes (0/1/2, 3 = missing).

    Sample i of SNP j uses Philox(counter=(i>>2, j, 0, 0)) word (i&3) for the
    genotype and Philox(counter=(i>>2, j, 1, 0)) word (i&3) for missingness,
    key = (seed & 0xffffffff, seed >> 32).
    """
    snp_ids = np.asarray(snp_ids, dtype=np.uint32)
    nq = (int(n) + 3) // 4
    q = np.arange(nq, dtype=np.uint32)[:, None]
    j = snp_ids[None, :]
    k0, k1 = int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF
    zero = np.zeros_like(q * j)
    g = philox4x32(q + zero, j + zero, zero, zero, k0, k1)
    mm = philox4x32(q + zero, j + zero, zero + 1, zero, k0, k1)
    t0 = np.asarray(t0, np.uint32)[None, :]
    t1 = np.asarray(t1, np.uint32)[None, :]
    out = np.empty((4 * nq, snp_ids.size), dtype=np.uint8)
    for w in range(4):
        r = g[w]
        code = np.where(r < t0, 0, np.where(r < t1, 1, 2)).astype(np.uint8)
        code = np.where(mm[w] < np.uint32(tmiss), np.uint8(3), code)
        out[w::4] = code
    return out[: int(n)]
