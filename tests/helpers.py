"""Shared helpers for the GPU parity tests (inputs built on the CPU)."""
import numpy as np

from oracle import geno


def random_codes(n, m, seed, miss=0.02):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.05, 0.5, size=m)
    g = rng.binomial(2, p[None, :], size=(n, m)).astype(np.uint8)
    g[rng.random((n, m)) < miss] = 3
    return g


def int8_block(codes, missing=-9):
    x = codes.astype(np.int8)
    x[codes == 3] = missing
    return x


def cache_codes(codes, variant_indices=None):
    if variant_indices is None:
        return codes
    _, order = geno.normalize_component_variant_indices(codes.shape[1], variant_indices)
    return codes[:, order]


def rel_err(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    den = np.linalg.norm(want)
    return float(np.linalg.norm(got - want) / (den if den > 0 else 1.0))
