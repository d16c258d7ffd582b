"""Synthetic workloads: genotypes generated straight into HBM.

Genotypes: per-SNP minor-allele frequency p ~ U(0.05, 0.5); g ~ Binomial(2, p)
under HWE via integer thresholds of a Philox4x32-10 stream (the
``coh_panel_generate`` kernel); 1% missing at random.  Phenotypes: intercept
+ N(0,1) covariates with N(0,1) effects + a genetic term on 1% causal SNPs
with N(0,1) effects on the standardized genotypes, scaled to the target
per-component h2, + Gaussian noise.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

TWO32 = 4294967296.0


def maf_thresholds(p, miss_rate):
    """uint32 thresholds: g = 0 below t0, 1 below t1, else 2; missing below tm."""
    p = np.asarray(p, dtype=np.float64)
    t0 = np.minimum(np.floor((1.0 - p) ** 2 * TWO32), TWO32 - 1).astype(np.uint32)
    t1 = np.minimum(np.floor((1.0 - p * p) * TWO32), TWO32 - 1).astype(np.uint32)
    tm = int(min(np.floor(float(miss_rate) * TWO32), TWO32 - 1))
    return t0, t1, tm


class SyntheticSource:
    """A genotype source that is generated on the GPU (no host genotype bytes).

    ``n`` samples, ``m`` SNPs with MAF ``p`` (array) and missing rate
    ``miss_rate``; SNP j is always drawn from counter stream j, so any cache
    permutation (component partitions) sees the same genotypes.
    """

    missing_val = -127

    def __init__(self, n: int, p, seed: int, miss_rate: float = 0.01):
        self.n = int(n)
        self.p = np.asarray(p, dtype=np.float64).reshape(-1)
        self.m = int(self.p.size)
        self.seed = int(seed)
        self.miss_rate = float(miss_rate)
        self.t0, self.t1, self.tmiss = maf_thresholds(self.p, self.miss_rate)

    def generate_into(self, panel, cache_to_source):
        ids = np.asarray(cache_to_source, dtype=np.int64)
        panel.generate(self.seed, ids.astype(np.uint32), self.t0[ids], self.t1[ids], self.tmiss)

    def close(self):
        return None


@dataclasses.dataclass
class Workload:
    name: str
    n: int
    m: int
    seed: int
    p: np.ndarray
    components: list            # list of source-index arrays (one per component)
    h2: np.ndarray              # per-component h2
    n_covar: int


def workload(cfg: int, *, scale: float = 1.0, seed_base: int = 20261203) -> Workload:
    """The synthetic workloads (n, m, partition, h2).  ``scale`` shrinks n and m."""
    rng = np.random.default_rng(seed_base + cfg)
    sizes = {0: (5000, 20000), 1: (100000, 500000), 2: (200000, 1000000),
             3: (300000, 600000), 4: (400000, 400000)}
    n, m = sizes[cfg]
    n = max(64, int(round(n * scale)))
    m = max(64, int(round(m * scale)))
    p = rng.uniform(0.05, 0.5, size=m)
    if cfg == 0:
        comps = [np.arange(m)]
        h2 = np.array([0.5])
        ncov = 2
    elif cfg in (1, 3):
        comps = [np.flatnonzero(p < 0.1), np.flatnonzero(p >= 0.1)]
        h2 = np.array([0.15, 0.35])
        ncov = 10
    elif cfg == 2:
        frac = rng.dirichlet(np.ones(8))
        sizes8 = np.maximum(1, np.floor(frac * m).astype(np.int64))
        sizes8[-1] = m - sizes8[:-1].sum()
        edges = np.concatenate([[0], np.cumsum(sizes8)])
        comps = [np.arange(edges[g], edges[g + 1]) for g in range(8)]
        h2 = rng.dirichlet(np.full(8, 2.0)) * 0.6
        ncov = 10
    else:
        comps = [np.arange(m)]
        h2 = np.array([0.4])
        ncov = 10
    return Workload(f"cfg{cfg}", n, m, seed_base + cfg, p, comps, h2, ncov)


def phenotype(streamer, wl: Workload, *, seed: int | None = None):
    """(y, X) on the host: X = [1, N(0,1) covariates]; y = X beta + genetic + noise.

    The genetic term of component g is Z_g b_g with 1% causal SNPs
    (b ~ N(0,1)) computed on the GPU by one Z.b pass per component, scaled
    to variance h2_g; noise has variance 1 - sum h2.
    """
    rng = np.random.default_rng(wl.seed + 7 if seed is None else seed)
    n = streamer.n
    X = np.concatenate([np.ones((n, 1)), rng.standard_normal((n, wl.n_covar))], axis=1)
    beta = rng.standard_normal(wl.n_covar + 1)
    off = streamer._component_snp_offsets
    gen = np.zeros(n)
    for g in range(streamer.n_components):
        a, b = int(off[g]), int(off[g + 1])
        size = b - a
        k = max(1, int(round(0.01 * size)))
        causal = rng.choice(size, size=k, replace=False) + a
        bvec = np.zeros((streamer.m, 1), np.float32)
        bvec[causal, 0] = rng.standard_normal(k)
        zb = zb_product(streamer, bvec)
        sd = float(np.std(zb))
        if sd > 0:
            gen += np.sqrt(wl.h2[g]) * (zb - zb.mean()) / sd
    noise = rng.standard_normal(n) * np.sqrt(max(1e-6, 1.0 - float(np.sum(wl.h2))))
    y = X @ beta + gen + noise
    return y.astype(np.float32), X.astype(np.float32)


def trait_phenotypes(streamer, wl: Workload, n_traits: int, *, seed: int | None = None,
                     max_mask: float = 0.3, h2_range=(0.1, 0.7)):
    """Multi-trait phenotypes on one panel: (Y (n, T), X (n, p+1), keep (T, n), h2 (T,)).

    Trait t has total h2_t ~ U(h2_range), split over the components in the
    proportions of ``wl.h2``; its genetic term per component is Z_g b_g with
    1% causal SNPs (one Z.B pass per component for all traits).  Each trait
    keeps samples outside a missing-phenotype mask covering a U(0, max_mask)
    fraction; covariates (intercept + ``wl.n_covar`` N(0,1)) are shared.
    """
    rng = np.random.default_rng(wl.seed + 11 if seed is None else seed)
    n, T = streamer.n, int(n_traits)
    X = np.concatenate([np.ones((n, 1)), rng.standard_normal((n, wl.n_covar))], axis=1)
    h2 = rng.uniform(h2_range[0], h2_range[1], size=T)
    share = np.asarray(wl.h2, np.float64) / float(np.sum(wl.h2))
    off = streamer._component_snp_offsets
    Y = X @ rng.standard_normal((wl.n_covar + 1, T))
    for g in range(streamer.n_components):
        a, b = int(off[g]), int(off[g + 1])
        size = b - a
        k = max(1, int(round(0.01 * size)))
        B = np.zeros((streamer.m, T), np.float32)
        for t in range(T):
            causal = rng.choice(size, size=k, replace=False) + a
            B[causal, t] = rng.standard_normal(k)
        out = streamer.panel.zb(torch.as_tensor(B), 0, streamer.n_components, precision=0)
        ZB = out.double().cpu().numpy()
        ZB = (ZB - ZB.mean(0)) / np.maximum(ZB.std(0), 1e-12)
        Y += np.sqrt(h2 * share[g])[None, :] * ZB
    Y += rng.standard_normal((n, T)) * np.sqrt(np.maximum(1e-6, 1.0 - h2))[None, :]
    frac = rng.uniform(0.0, max_mask, size=T)
    keep = rng.random((T, n)) >= frac[:, None]
    return Y.astype(np.float32), X.astype(np.float32), keep, h2


def serving_requests(streamer, n_requests: int, rate_per_hour: float, *, seed: int = 20261207,
                     n_blocks: int | None = None, median_samples: float = 50_000.0,
                     sigma: float = 1.0, min_samples: int = 10_000, max_components: int = 8,
                     covariates=(5, 20), h2_range=(0.1, 0.7)):
    """Poisson-arriving REML fit requests on a panel of contiguous blocks.

    Each request: a sample subset of size LogNormal(ln median, sigma) clipped
    to [min_samples, n]; 1..max_components components that are a random
    partition of the blocks into runs; 5..20 N(0,1) covariates plus an
    intercept; h2 ~ U(h2_range) carried by 1% causal SNPs of its components
    (N(0,1) effects, one Z.B pass for all requests).  Returns a list of
    ``serving.FitRequest`` sorted by arrival.
    """
    from .serving import FitRequest

    rng = np.random.default_rng(seed)
    n, R = streamer.n, int(n_requests)
    B = int(streamer.n_components if n_blocks is None else n_blocks)
    off = streamer._component_snp_offsets
    arrivals = np.cumsum(rng.exponential(3600.0 / float(rate_per_hour), size=R))
    arrivals -= arrivals[0]
    C = rng.standard_normal((n, int(covariates[1])))
    Bmat = np.zeros((streamer.m, R), np.float32)
    specs = []
    for r in range(R):
        size = int(np.clip(np.exp(rng.normal(np.log(median_samples), sigma)), min_samples, n))
        keep = np.zeros(n, bool)
        keep[rng.choice(n, size=size, replace=False)] = True
        G = int(rng.integers(1, min(max_components, B) + 1))
        cuts = np.sort(rng.choice(np.arange(1, B), size=G - 1, replace=False)) if G > 1 else []
        edges = np.concatenate([[0], cuts, [B]]).astype(int)
        groups = [list(range(edges[g], edges[g + 1])) for g in range(G)]
        snps = np.arange(int(off[0]), int(off[B]))
        k = max(1, int(round(0.01 * snps.size)))
        causal = rng.choice(snps, size=k, replace=False)
        Bmat[causal, r] = rng.standard_normal(k)
        p = int(rng.integers(int(covariates[0]), int(covariates[1]) + 1))
        h2 = float(rng.uniform(*h2_range))
        specs.append((keep, groups, p, h2))
    gen = streamer.panel.zb(torch.as_tensor(Bmat), 0, streamer.n_components,
                            precision=0).double().cpu().numpy()
    out = []
    for r, (keep, groups, p, h2) in enumerate(specs):
        g = gen[:, r]
        g = (g - g.mean()) / max(float(g.std()), 1e-12)
        X = np.concatenate([np.ones((n, 1)), C[:, :p]], axis=1)
        y = X @ rng.standard_normal(p + 1) + np.sqrt(h2) * g + np.sqrt(1 - h2) * rng.standard_normal(n)
        out.append(FitRequest(keep=keep, groups=groups, y=y.astype(np.float32),
                              X=X.astype(np.float32), arrival_s=float(arrivals[r])))
    return out


def initial_components(y, X, atoms, h2_init: float = 0.5) -> np.ndarray:
    """Start point (theta_g..., theta_e) scaled to the OLS residual variance.

    The fit standardizes y to unit variance and its default start spreads
    that unit total as h2_init / (1 - h2_init).  When covariates carry most
    of var(y) (ten N(0,1) effects carry ~90% here) the REML optimum has a
    total far below 1 and the first AI step from the default start
    overshoots into the bounds.  This start puts the same split on
    ``||y_s - X b||^2 / (n - p)`` of the standardized response instead,
    passed as ``var_components_init`` (reml.py:1452-1463).
    """
    y = np.asarray(y, np.float64).reshape(-1)
    ys = (y - y.mean()) / max(float(y.std()), 1e-12)
    X = np.asarray(X, np.float64)
    beta, *_ = np.linalg.lstsq(X, ys, rcond=None)
    s2 = float(np.sum((ys - X @ beta) ** 2) / max(1, ys.size - X.shape[1]))
    atoms = np.asarray(atoms, np.float64).reshape(-1)
    tg = np.where(atoms > 0, h2_init * s2 / atoms.sum(), 0.0)
    return np.concatenate([tg, [(1.0 - h2_init) * s2]])


def zb_product(streamer, b) -> np.ndarray:
    """Z b (n,) for cache-order SNP effects b (m, 1), computed on the GPU."""
    out = streamer.panel.zb(torch.as_tensor(np.asarray(b, np.float32)), 0,
                            streamer.n_components, precision=0)
    return out[:, 0].double().cpu().numpy()
