"""Batched multi-trait REML: many traits share every pass over one panel.

The reference fits traits one at a time, each on a streamer restricted to
the trait's observed samples (``sample_mask``, geno_stream.py:1161-1247):
genotype blocks keep the masked samples (geno_stream.py:2597-2598), the
per-SNP statistics, ``eff`` and trace atoms are taken over them
(geno_stream.py:1394-1448, 1780-1792), and ``fit_infinitesimal``
(reml_model.py:1836-2010) runs on that streamer.

Here the panel stays resident once and every trait keeps exactly that
semantics:

* ``SubsetStats`` computes each trait's statistics over its kept samples in
  one pass over the panel (``coh_panel_subset_stats``), then ``eff`` and the
  trace atoms per trait and component as ``geno_stream._init_stats`` does.
* ``TraitView`` presents one trait as a component-partitioned streamer on its
  ``n_t`` kept samples, so ``InfinitesimalREMLFitter`` runs unchanged on it.
* ``PassBatcher`` joins the traits' fits, each in its own host thread: every
  operator call becomes a request, and when all live fits are waiting the
  requests are packed column-wise into one wide right-hand block (rows outside
  a trait's samples zero) with a statistic set per column, and contracted in
  one ``coh_kv_sets`` call per precision (stacked per component).  Weighted
  products, single components and projected-core atoms are formed from the
  stacked result.

Each trait's fit is the fit of that trait alone: the per-column scales of
the int8-limb contraction do not depend on the other columns, and the
integer partial sums are exact, so a trait's operator results agree with the
solo fit's to FP32 rounding (tested).  The plain FP32 path takes one
statistic set per call and is run trait by trait (``fit_traits_sequential``).
"""
from __future__ import annotations

import ctypes
import dataclasses
import threading
import time
from typing import Sequence

import numpy as np
import torch

from . import _native as N
from .reml_model import FitConfig, FitResult, InfinitesimalREMLFitter

__all__ = ["SubsetStats", "TraitView", "PassBatcher", "fit_traits", "fit_traits_sequential"]


# ------------------------------------------------------------- statistics
class SubsetStats:
    """Per-subset standardization statistics on a resident panel.

    ``keep`` is (T, n) bool: row t selects trait t's samples.  Device arrays
    ``mean``/``inv`` are (T, m_slot) float32 in slot space; ``eff`` (T, G)
    counts valid SNPs per component and ``atoms`` (T, G) are the component
    mean-diagonal atoms sum_valid cnt_j / (n_t eff_g) (geno_stream.py:1409-1437).
    """

    def __init__(self, panel, keep: np.ndarray):
        keep = np.asarray(keep, dtype=bool)
        if keep.ndim != 2 or keep.shape[1] != panel.n:
            raise ValueError(f"keep must have shape (T, {panel.n}), got {keep.shape}.")
        T = int(keep.shape[0])
        if T < 1:
            raise ValueError("keep needs at least one subset.")
        self.n_subsets = T
        self.sizes = keep.sum(axis=1).astype(np.int64)
        if np.any(self.sizes < 1):
            raise ValueError("every subset needs at least one sample.")
        dev = panel.device
        self.stride = int(panel.m_slot)
        keep_dev = torch.as_tensor(keep.astype(np.uint8), device=dev)
        self.mean = torch.empty((T, self.stride), dtype=torch.float32, device=dev)
        self.inv = torch.empty((T, self.stride), dtype=torch.float32, device=dev)
        cnt = torch.empty((T, self.stride), dtype=torch.int64, device=dev)
        nb = ctypes.c_uint64()
        N.call("coh_subset_stats_workspace_bytes", panel.handle, T, ctypes.byref(nb))
        ws = torch.empty(max(int(nb.value), 256), dtype=torch.uint8, device=dev)
        N.call("coh_panel_subset_stats", panel.handle, T, N.ptr(keep_dev), N.ptr(self.mean),
               N.ptr(self.inv), N.ptr(cnt), self.stride, N.ptr(ws), int(nb.value), N.stream_ptr())
        G = int(panel.n_components)
        valid = self.inv > 0.0
        cntf = torch.where(valid, cnt, torch.zeros_like(cnt)).to(torch.float64)
        eff = torch.empty((T, G), dtype=torch.float64, device=dev)
        mass = torch.empty((T, G), dtype=torch.float64, device=dev)
        for g in range(G):
            a, b = int(panel.comp_slot_begin[g]), int(panel.comp_slot_end[g])
            eff[:, g] = valid[:, a:b].sum(dim=1).to(torch.float64)
            mass[:, g] = cntf[:, a:b].sum(dim=1)
        eff_h = eff.cpu().numpy()
        mass_h = mass.cpu().numpy()
        self.eff = eff_h.astype(np.float32)
        n_t = self.sizes.astype(np.float64)[:, None]
        with np.errstate(divide="ignore", invalid="ignore"):
            atoms = np.where(eff_h > 0, mass_h / (n_t * eff_h), 0.0)
        self.atoms = atoms.astype(np.float32)
        self.eff_total = eff_h.sum(axis=1)
        tot_mass = mass_h.sum(axis=1)
        with np.errstate(divide="ignore", invalid="ignore"):
            self.atom_total = np.where(self.eff_total > 0,
                                       tot_mass / (self.sizes * self.eff_total), 0.0)
        inv_eff = np.where(self.eff > 0, 1.0 / np.maximum(self.eff, 1e-30), 0.0).astype(np.float32)
        self.comp_scale = torch.as_tensor(inv_eff, device=dev).contiguous()
        self.index = [torch.as_tensor(np.flatnonzero(keep[t]), device=dev) for t in range(T)]

    def cache_order(self, panel, t: int):
        """Trait t's (mean, inv_sd) in cache order (m,), for a panel override."""
        idx = torch.as_tensor(panel.cache_to_slot_host.astype(np.int64), device=self.mean.device)
        return self.mean[t].index_select(0, idx).contiguous(), self.inv[t].index_select(0, idx).contiguous()


# ---------------------------------------------------------------- batcher
@dataclasses.dataclass
class _Request:
    trait: int
    V: torch.Tensor            # (n_t, k) float32, trait space
    precision: str
    result: torch.Tensor | None = None
    error: BaseException | None = None
    done: bool = False


class PassBatcher:
    """Rendezvous of concurrent trait fits over one resident panel.

    ``submit`` blocks the calling fit until its stacked product is ready.  A
    batch runs when every live fit is waiting (``leave`` retires a fit), so
    the traits' operator calls share genotype passes whatever state each fit
    is in.  Batches wider than ``max_cols`` columns are split.
    """

    def __init__(self, streamer, stats: SubsetStats, n_fits: int, *, max_cols: int = 2048):
        self.streamer = streamer
        self.panel = streamer.panel
        self.stats = stats
        self.max_cols = int(max_cols)
        self._cv = threading.Condition()
        self._live = int(n_fits)
        self._pending: list[_Request] = []
        self.n_batches = 0
        self.n_calls = 0
        self.columns = 0
        self.requests = 0
        self.busy_s = 0.0

    def submit(self, req: _Request) -> torch.Tensor:
        with self._cv:
            self._pending.append(req)
            if len(self._pending) >= self._live:
                self._run_locked()
            else:
                while not req.done:
                    self._cv.wait()
        if req.error is not None:
            raise req.error
        return req.result

    def leave(self) -> None:
        with self._cv:
            self._live -= 1
            if self._pending and len(self._pending) >= self._live:
                self._run_locked()

    def _run_locked(self) -> None:
        batch, self._pending = self._pending, []
        t0 = time.perf_counter()
        try:
            batch.sort(key=lambda r: r.trait)
            for prec in sorted({r.precision for r in batch}):
                self._contract([r for r in batch if r.precision == prec], prec)
        except BaseException as e:  # every waiting fit sees the failure
            for r in batch:
                if not r.done:
                    r.error = e
        finally:
            self.busy_s += time.perf_counter() - t0
            self.n_batches += 1
            for r in batch:
                r.done = True
            self._cv.notify_all()

    def _contract(self, reqs: Sequence[_Request], prec: str) -> None:
        # column pieces (request, c0, c1), packed greedily up to max_cols
        pieces = []
        for r in reqs:
            k = int(r.V.shape[1])
            r.result = torch.empty((self.panel.n_components, r.V.shape[0], k),
                                   dtype=torch.float32, device=r.V.device)
            for c0 in range(0, k, self.max_cols):
                pieces.append((r, c0, min(k, c0 + self.max_cols)))
        group, width = [], 0
        for piece in pieces:
            w = piece[2] - piece[1]
            if group and width + w > self.max_cols:
                self._call(group, width, prec)
                group, width = [], 0
            group.append(piece)
            width += w
        if group:
            self._call(group, width, prec)
        self.requests += len(reqs)

    def _call(self, group, width: int, prec: str) -> None:
        panel, st = self.panel, self.stats
        dev = panel.device
        n, G = panel.n, panel.n_components
        V = torch.zeros((n, width), dtype=torch.float32, device=dev)
        col_set = np.empty(width, np.int32)
        c = 0
        for r, c0, c1 in group:
            w = c1 - c0
            V[st.index[r.trait], c:c + w] = r.V[:, c0:c1]
            col_set[c:c + w] = r.trait
            c += w
        col_set_dev = torch.as_tensor(col_set, device=dev)
        out = torch.empty((G, n, width), dtype=torch.float32, device=dev)
        precision = N.PRECISIONS[prec]
        nb = panel.workspace_bytes(N.OP_KV, precision, width, 0, G)
        ws = torch.empty(max(nb, 256), dtype=torch.uint8, device=dev)
        N.call("coh_kv_sets", panel.handle, int(precision), N.ptr(V), int(width), int(V.stride(0)),
               0, int(G), int(st.n_subsets), N.ptr(st.mean), N.ptr(st.inv), int(st.stride),
               N.ptr(col_set_dev), N.ptr(st.comp_scale), 1, N.ptr(out), int(out.stride(1)),
               int(out.stride(0)), 0.0, N.ptr(ws), nb, N.stream_ptr())
        c = 0
        for r, c0, c1 in group:
            w = c1 - c0
            r.result[:, :, c0:c1] = out[:, st.index[r.trait], c:c + w]
            c += w
        self.n_calls += 1
        self.columns += width


# ------------------------------------------------------------------- view
class TraitView:
    """One trait as a component-partitioned streamer on its kept samples.

    Mirrors the ``GenoBlockStreamer`` operator methods that the REML fit
    uses (geno_stream.py:2669-2778); every product goes through the shared
    ``PassBatcher``.
    """

    def __init__(self, batcher: PassBatcher, trait: int, precision: str):
        if precision not in ("i8x2", "i8x3"):
            raise ValueError("batched traits run the int8-limb forms ('i8x2' or 'i8x3'); "
                             "use fit_traits_sequential for 'fp32'.")
        self._b = batcher
        self.trait = int(trait)
        st = batcher.stats
        self.n = int(st.sizes[trait])
        self.m = int(batcher.streamer.m)
        self.dev = batcher.panel.device
        self.precision = precision
        self._G = int(batcher.panel.n_components)
        self._eff = st.eff[trait].astype(np.float64)
        self._atoms = st.atoms[trait]

    # streamer duck type
    @property
    def n_components(self) -> int:
        return self._G

    @property
    def has_component_partition(self) -> bool:
        return self._G > 1

    @property
    def residual_precision(self) -> str:
        return "i8x3" if self.precision == "i8x2" else self.precision

    def _stacked(self, V, precision=None):
        V = torch.as_tensor(V, device=self.dev)
        squeeze = V.ndim == 1
        V = (V[:, None] if squeeze else V).to(torch.float32).contiguous()
        if V.shape[0] != self.n:
            raise ValueError(f"V must have {self.n} rows, got {V.shape[0]}.")
        out = self._b.submit(_Request(self.trait, V, precision or self.precision))
        return out, V, squeeze

    def stacked_component_kv(self, V, normalize: bool = True):
        if not normalize:
            raise NotImplementedError("batched traits apply 1/eff per component.")
        out, _, sq = self._stacked(V)
        return out[:, :, 0] if sq else out

    def component_kv(self, V, component_idx: int, normalize: bool = True):
        if component_idx < 0 or component_idx >= self._G:
            raise IndexError(f"component_idx={component_idx} out of range for {self._G} components.")
        out = self.stacked_component_kv(V, normalize)
        return out[component_idx]

    def kv(self, V, normalize: bool = True):
        """All SNPs, 1/eff of the whole panel (geno_stream.py:2669-2679)."""
        out, _, sq = self._stacked(V)
        eff_tot = float(self._eff.sum())
        acc = sum(float(self._eff[g]) * out[g] for g in range(self._G))
        acc = acc / eff_tot if (normalize and eff_tot > 0) else acc
        return acc[:, 0] if sq else acc

    def weighted_component_hv(self, theta_g, theta_e, V, precision: str | None = None):
        th = np.asarray(theta_g, np.float64).reshape(-1)
        if th.size != self._G:
            raise ValueError(f"theta_g length mismatch: expected {self._G}, got {th.size}.")
        out, Vm, sq = self._stacked(V, precision)
        acc = None
        for g in range(self._G):
            term = float(th[g]) * out[g]
            acc = term if acc is None else acc + term
        if theta_e is not None:
            acc = acc + float(np.asarray(theta_e, np.float64).reshape(-1)[0]) * Vm
        return acc[:, 0] if sq else acc

    def build_projected_core_atoms(self, U, *, subtract_identity: bool = True):
        """U^T K_g U [- I] per component (kv_impl.py:1287-1339)."""
        out, Um, _ = self._stacked(U)
        cores = torch.stack([Um.T @ out[g] for g in range(self._G)])
        cores = 0.5 * (cores + cores.transpose(-1, -2))
        if subtract_identity:
            cores = cores - torch.eye(cores.shape[-1], device=cores.device)[None]
        return cores

    def component_diag_list(self):
        return [torch.tensor(float(a), dtype=torch.float32, device=self.dev) for a in self._atoms]

    def diag(self):
        return torch.tensor(float(self._b.stats.atom_total[self.trait]), dtype=torch.float32,
                            device=self.dev)

    def close(self) -> None:
        pass


# -------------------------------------------------------------------- fit
def _slq_budget(cfg: FitConfig, n_t: int) -> float:
    """GPU budget whose SLQ share runs a trait's whole probe block as one
    batch (slq.workspace_layout's per-probe estimate)."""
    depth = min(int(cfg.slq_m), int(n_t))
    per_probe = 4.0 * n_t * (8 * (depth + 1) + 32)
    return 8.0 * per_probe * max(1, int(cfg.slq_samples))


def _trait_inputs(Y, keep, covar, t):
    idx = np.flatnonzero(keep[t])
    y = np.asarray(Y[idx, t] if Y.ndim == 2 else Y[t][idx], np.float32)
    X = None if covar is None else np.asarray(covar, np.float32)[idx]
    return y, X


def fit_traits(streamer, Y, keep, covar=None, *, cfg: FitConfig | None = None,
               var_components_init=None, max_cols: int = 2048, stats: SubsetStats | None = None,
               counters: list | None = None, keep_precond: bool = False):
    """Fit T traits concurrently on one resident panel (shared passes).

    ``Y`` is (n, T) (entries outside ``keep`` are ignored), ``keep`` (T, n)
    bool, ``covar`` (n, p) shared by all traits (restricted to each trait's
    samples).  ``var_components_init`` is None, one vector for all traits, or
    (T, G+1).  Returns ``(results, info)`` with one ``FitResult`` per trait;
    with ``keep_precond`` ``info["precond_U"][t]`` is trait t's SLQ whitening
    basis.
    """
    cfg = cfg or FitConfig()
    keep = np.asarray(keep, dtype=bool)
    Y = np.asarray(Y)
    T = int(keep.shape[0])
    if stats is None:
        stats = SubsetStats(streamer.panel, keep)
    batcher = PassBatcher(streamer, stats, T, max_cols=max_cols)
    results: list = [None] * T
    errors: list = [None] * T
    bases: list = [None] * T
    init = None if var_components_init is None else np.asarray(var_components_init, np.float64)

    def run(t):
        try:
            view = TraitView(batcher, t, cfg.precision)
            c = dataclasses.replace(cfg, gpu_budget_bytes=_slq_budget(cfg, view.n), verbose=False)
            fitter = InfinitesimalREMLFitter(c, streamers=[view])
            y, X = _trait_inputs(Y, keep, covar, t)
            vc0 = None if init is None else (init[t] if init.ndim == 2 else init)
            cnt = counters[t] if counters is not None else None
            results[t] = fitter.fit_infinitesimal(y, X, var_components_init=vc0, counters=cnt)
            if keep_precond and fitter.slq_precond_conf is not None:
                bases[t] = fitter.slq_precond_conf.U
        except BaseException as e:
            errors[t] = e
        finally:
            batcher.leave()

    t0 = time.perf_counter()
    threads = [threading.Thread(target=run, args=(t,), name=f"trait-{t}") for t in range(T)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    wall = time.perf_counter() - t0
    for t, e in enumerate(errors):
        if e is not None:
            try:
                named = type(e)(f"trait {t}: {e}")
            except Exception:
                raise e
            raise named from e
    info = dict(wall_s=wall, batches=batcher.n_batches, kv_sets_calls=batcher.n_calls,
                columns=batcher.columns, requests=batcher.requests, contract_s=batcher.busy_s)
    if keep_precond:
        info["precond_U"] = bases
    return results, info


def fit_traits_sequential(streamer, Y, keep, covar=None, *, cfg: FitConfig | None = None,
                          var_components_init=None, stats: SubsetStats | None = None,
                          traits: Sequence[int] | None = None, keep_precond: bool = False):
    """The reference's way: one fit per trait, each on the panel standardized
    with the trait's statistics and restricted to its samples.  Runs any
    precision (the plain FP32 drop-in included); restores the panel's own
    statistics afterwards.  Returns ``(results, info)`` like ``fit_traits``
    (``info["wall_s"]`` per trait in ``info["trait_s"]``)."""
    cfg = cfg or FitConfig()
    keep = np.asarray(keep, dtype=bool)
    Y = np.asarray(Y)
    panel = streamer.panel
    if stats is None:
        stats = SubsetStats(panel, keep)
    own = panel.stats()
    saved_mean = torch.as_tensor(own[0], device=panel.device)
    saved_inv = torch.as_tensor(own[1], device=panel.device)
    init = None if var_components_init is None else np.asarray(var_components_init, np.float64)
    out, bases, times = [], [], []
    try:
        for t in (range(keep.shape[0]) if traits is None else traits):
            t0 = time.perf_counter()
            mean_c, inv_c = stats.cache_order(panel, t)
            panel.set_standardization(mean_c, inv_c)
            view = _SequentialTraitView(streamer, stats, t, cfg.precision)
            c = dataclasses.replace(cfg, gpu_budget_bytes=_slq_budget(cfg, view.n), verbose=False)
            fitter = InfinitesimalREMLFitter(c, streamers=[view])
            y, X = _trait_inputs(Y, keep, covar, t)
            vc0 = None if init is None else (init[t] if init.ndim == 2 else init)
            out.append(fitter.fit_infinitesimal(y, X, var_components_init=vc0))
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            if keep_precond and fitter.slq_precond_conf is not None:
                bases.append(fitter.slq_precond_conf.U)
    finally:
        panel.set_standardization(saved_mean, saved_inv)
    info = dict(wall_s=float(sum(times)), trait_s=times)
    if keep_precond:
        info["precond_U"] = bases
    return out, info


class _SequentialTraitView(TraitView):
    """A trait on the panel's own statistics (set to the trait's), one fit
    at a time: each product is a plain ``coh_kv`` of the zero-extended block."""

    def __init__(self, streamer, stats: SubsetStats, trait: int, precision: str):
        self._st = streamer
        self._stats = stats
        self.trait = int(trait)
        self.n = int(stats.sizes[trait])
        self.m = int(streamer.m)
        self.dev = streamer.panel.device
        self.precision = precision
        self._G = int(streamer.panel.n_components)
        self._eff = stats.eff[trait].astype(np.float64)
        self._atoms = stats.atoms[trait]

    def _stacked(self, V, precision=None):
        V = torch.as_tensor(V, device=self.dev)
        squeeze = V.ndim == 1
        V = (V[:, None] if squeeze else V).to(torch.float32).contiguous()
        if V.shape[0] != self.n:
            raise ValueError(f"V must have {self.n} rows, got {V.shape[0]}.")
        panel = self._st.panel
        idx = self._stats.index[self.trait]
        full = torch.zeros((panel.n, V.shape[1]), dtype=torch.float32, device=self.dev)
        full[idx] = V
        scale = np.where(self._eff > 0, 1.0 / np.maximum(self._eff, 1e-30), 0.0).astype(np.float32)
        out = panel.kv(full, scale, 0, self._G, stacked=True,
                       precision=N.PRECISIONS[precision or self.precision])
        return out[:, idx], V, squeeze

    def diag(self):
        return torch.tensor(float(self._stats.atom_total[self.trait]), dtype=torch.float32,
                            device=self.dev)
