"""Public model API: FitConfig and InfinitesimalREMLFitter (reml_model.py).

Builds resident-panel streamers, assembles the REML operator bundle
(reml_model.py:962-1390), builds/refreshes the projected-core preconditioner
(reml_model.py:808-960, 1392-1429) and runs fit_reml (reml_model.py:1836-2010).
Dense single/partitioned/multi-streamer GRMs are supported; the sparse,
SMILE and admixture branches of the reference are out of scope and raise.
"""
from __future__ import annotations

import dataclasses
import logging
import time
from typing import Optional, Sequence

import numpy as np
import torch

from .geno_stream import GenoBlockStreamer
from .precond import ProjectedCorePrecondConf, build_lowrank_basis
from .reml import REMLProbeCache, fit_reml
from .slq import default_workspace_bytes

logger = logging.getLogger(__name__)


@dataclasses.dataclass
class FitConfig:
    """reml_model.py:34-83 (field names kept; out-of-scope fields must stay unset)."""

    bed_prefix: str | Sequence[str] = ""
    sources: Sequence | None = None
    rare_bed_prefix: str | Sequence[str] = ""
    rare_sources: Sequence | None = None
    vc_block_sizes: Sequence[int] | None = None
    component_variant_indices: Sequence[Sequence[int]] | None = None
    smile_w_files: Sequence[str] | None = None
    smile_weight_matrices: Sequence[np.ndarray] | None = None
    admix_weights: np.ndarray | None = None
    standardization_overrides: Sequence[tuple] | None = None
    sample_mask: np.ndarray | None = None
    call_width: int = 131072
    device: str | None = "auto"
    keep_host_stats: bool = True
    cpu_threads: int | None = None
    n_rand_vec: int = 100
    max_pcg_iters: int = 400
    minq_iter: int = 50
    seed: int = 0
    slq_samples: int = 4
    slq_m: int = 8
    slq_mode: str = "projected_core_residual"
    precond_rank: int = 500
    pcg_ridge: float = 1e-6
    reml_pcg_tol: float = 1e-3
    effect_pcg_tol: float = 1e-3
    response_is_standardized: bool = False
    unit_variance_components: bool = False
    n_reml_reps: int = 1
    strict_max_linesearch_trials: int = 3
    verbose: bool = True
    gpu_budget_bytes: float | None = None
    capture_reml_diagnostics: bool = False
    cache_reml_setup: bool = False
    # contraction kernels of this engine: "fp32" (plain), "i8x3", "i8x2"
    precision: str = "fp32"


@dataclasses.dataclass
class FitResult:
    var_components: np.ndarray
    history: list
    rep_var_components: Optional[np.ndarray] = None
    final_ai: Optional[np.ndarray] = None
    final_grad: Optional[np.ndarray] = None
    final_loglik: Optional[float] = None
    diagnostics: Optional[dict] = None
    heritability: Optional[float] = None
    genetic_trace_atoms: Optional[np.ndarray] = None
    residual_trace_atoms: Optional[np.ndarray] = None
    monte_carlo_se_var: Optional[np.ndarray] = None
    monte_carlo_se_h2: Optional[float] = None
    counters: Optional[dict] = None


@dataclasses.dataclass(frozen=True)
class _OperatorBundle:
    K_mvs: tuple
    diag_list: tuple
    weighted_hv: object
    stacked_kv: object
    residual_diag_list: tuple | None = None
    projected_core_atoms: object | None = None
    weighted_hv_true: object | None = None


def _as_list(x):
    if x is None or x == "":
        return []
    return list(x) if isinstance(x, (list, tuple)) else [x]


class InfinitesimalREMLFitter:
    """reml_model.py:402-2010 (dense GRM paths)."""

    def __init__(self, cfg: FitConfig, *, streamers: Sequence[GenoBlockStreamer] | None = None):
        self.cfg = cfg
        if _as_list(cfg.rare_bed_prefix) or cfg.rare_sources:
            raise NotImplementedError("Sparse rare-variant streamers are out of scope.")
        if cfg.smile_w_files or cfg.smile_weight_matrices:
            raise NotImplementedError("SMILE weighted kernels are out of scope.")
        if cfg.admix_weights is not None:
            raise NotImplementedError("Admixed covariance components are out of scope.")
        self._partitioned_streamer = None
        t0 = time.time()
        if streamers is not None:
            self.streamers = list(streamers)
        else:
            self.streamers = self._build_streamers()
        if len(self.streamers) == 1 and self.streamers[0].has_component_partition:
            self._partitioned_streamer = self.streamers[0]
        if cfg.verbose:
            logger.info("build streamers done elapsed=%.1fs", time.time() - t0)
        self.precond_conf = None
        self._reml_probe_cache = REMLProbeCache() if cfg.cache_reml_setup else None

    def _build_streamers(self):
        cfg = self.cfg
        srcs = list(cfg.sources or [])
        beds = _as_list(cfg.bed_prefix)
        if srcs and beds:
            raise ValueError("Provide either sources or bed_prefix, not both.")
        inputs = [("source", s) for s in srcs] + [("bed", b) for b in beds]
        if not inputs:
            raise ValueError("FitConfig needs sources or bed_prefix.")
        partition = cfg.vc_block_sizes is not None or cfg.component_variant_indices is not None
        if partition and len(inputs) != 1:
            raise ValueError("Component partitions require exactly one genotype source.")
        overrides = list(cfg.standardization_overrides or [None] * len(inputs))
        out = []
        for (kind, val), ov in zip(inputs, overrides):
            kw = dict(call_width=cfg.call_width, device=None if cfg.device == "auto" else cfg.device,
                      keep_host_stats=cfg.keep_host_stats, sample_mask=cfg.sample_mask,
                      standardization_override=ov, precision=cfg.precision,
                      component_block_sizes=cfg.vc_block_sizes,
                      component_variant_indices=cfg.component_variant_indices)
            if kind == "source":
                out.append(GenoBlockStreamer(val, **kw))
            else:
                out.append(GenoBlockStreamer(bed_prefix=val, **kw))
        return out

    # ---------------------------------------------------------- operators
    def _assemble_reml_operators(self) -> _OperatorBundle:
        """reml_model.py:1166-1196, 1347-1390."""
        if self._partitioned_streamer is not None:
            st = self._partitioned_streamer
            K_mvs = tuple((lambda V, g=g: st.component_kv(V, g, normalize=True))
                          for g in range(st.n_components))
            rp = st.residual_precision
            return _OperatorBundle(
                K_mvs=K_mvs, diag_list=tuple(st.component_diag_list()),
                weighted_hv=lambda tg, te, V: st.weighted_component_hv(tg, te, V),
                stacked_kv=lambda V: st.stacked_component_kv(V, normalize=True),
                projected_core_atoms=lambda U: st.build_projected_core_atoms(U, subtract_identity=False),
                weighted_hv_true=(None if rp == st.precision else
                                  (lambda tg, te, V: st.weighted_component_hv(tg, te, V, precision=rp))),
            )
        sts = tuple(self.streamers)
        K_mvs = tuple((lambda V, st=st: st.kv(V)) for st in sts)
        diag_list = tuple(st.diag() for st in sts)
        if len(sts) == 1:
            st = sts[0]
            rp = st.residual_precision
            return _OperatorBundle(K_mvs=K_mvs, diag_list=diag_list, weighted_hv=None,
                                   stacked_kv=None, projected_core_atoms=None,
                                   weighted_hv_true=(None if rp == st.precision else
                                                     (lambda tg, te, V: st.weighted_component_hv(
                                                         tg, te, V, precision=rp))))

        def weighted_hv(tg, te, V, refine=False):
            tg = np.asarray(tg, np.float64).reshape(-1)
            acc = None
            for g, st in enumerate(sts):
                # one unpartitioned streamer per GRM: component 0 spans all SNPs
                prec = st.residual_precision if refine else None
                term = st.weighted_component_hv([tg[g]], None, V, precision=prec)
                acc = term if acc is None else acc + term
            if te is not None:
                acc = acc + float(np.asarray(te).reshape(-1)[0]) * V
            return acc

        refine = any(st.residual_precision != st.precision for st in sts)

        def stacked_kv(V):
            return torch.stack([st.kv(V) for st in sts])

        def core_atoms(U):
            return torch.stack([st.build_projected_core_atom(U, subtract_identity=False) for st in sts])

        return _OperatorBundle(K_mvs=K_mvs, diag_list=diag_list, weighted_hv=weighted_hv,
                               stacked_kv=stacked_kv, projected_core_atoms=core_atoms,
                               weighted_hv_true=((lambda tg, te, V: weighted_hv(tg, te, V, True))
                                                 if refine else None))

    # ------------------------------------------------------ preconditioner
    @staticmethod
    def _theta_ref_from_var_components(n_components, var_components_init):
        """reml_model.py:808-832: theta_ref normalised to sum 1."""
        if var_components_init is not None:
            th = np.asarray(var_components_init, np.float32).reshape(-1)
            if th.size < n_components:
                raise ValueError(
                    f"var_components_init length mismatch: expected at least {n_components}, got {th.size}.")
            th = np.maximum(th[:n_components], 0.0)
        else:
            th = np.ones(n_components, np.float32)
        tot = float(np.sum(th))
        if tot <= 0.0 or not np.isfinite(tot):
            return np.full(n_components, 1.0 / max(1, n_components), np.float32)
        return (th / tot).astype(np.float32)

    @staticmethod
    def _diag_atoms(diag_list):
        return np.asarray([float(np.mean(d.detach().cpu().numpy())) if isinstance(d, torch.Tensor)
                           else float(np.mean(d)) for d in diag_list], np.float32)

    def _build_projected_core_precond(self, ops: _OperatorBundle, var_components_init):
        """reml_model.py:863-960: Nystrom basis of K_ref, cores U^T K_g U - a_g I."""
        cfg = self.cfg
        G = len(ops.K_mvs)
        diag_atoms = self._diag_atoms(ops.diag_list)
        theta_ref = self._theta_ref_from_var_components(G, var_components_init)
        if ops.weighted_hv is not None:
            K_ref = lambda V: ops.weighted_hv(theta_ref, None, V)
        else:
            def K_ref(V):
                acc = None
                for g, mv in enumerate(ops.K_mvs):
                    t = float(theta_ref[g]) * mv(V)
                    acc = t if acc is None else acc + t
                return acc
        n = self.streamers[0].n
        U, evals = build_lowrank_basis(K_ref, n, cfg.precond_rank, cfg.seed + 4321)
        r = int(U.shape[1])
        eye = torch.eye(r, dtype=torch.float32, device=U.device)
        if G == 1:
            raw = torch.diag(evals.to(U.dtype))[None]
        elif ops.projected_core_atoms is not None:
            raw = ops.projected_core_atoms(U)
        else:
            KU = torch.stack([mv(U) for mv in ops.K_mvs])
            raw = torch.einsum("nr,gns->grs", U, KU)
        cores = 0.5 * (raw + raw.transpose(-1, -2))
        cores = cores - torch.as_tensor(diag_atoms, device=U.device)[:, None, None] * eye[None]
        conf = ProjectedCorePrecondConf(
            U=U, core_atoms=cores, total_rank=r, n_grm=G, diag_mode="scalar_identity",
            diag_atoms=torch.as_tensor(diag_atoms, device=U.device),
            residual_diag_atoms=torch.ones(1, device=U.device), identity=eye)
        self.precond_conf = conf
        return conf

    def _make_pcg_precond_refresh_fn(self, ops):
        """reml_model.py:1409-1429: rebuilt after accepted G>1 steps."""
        if self.cfg.precond_rank <= 0 or len(ops.K_mvs) <= 1:
            return None
        return lambda param: self._build_projected_core_precond(ops, np.asarray(param, np.float32))

    # ------------------------------------------------------------------ fit
    def fit_infinitesimal(self, y, covar=None, h2_init: float = 0.5, var_components_init=None,
                          estimate_effects: bool = False, counters: dict | None = None) -> FitResult:
        """reml_model.py:1836-2010."""
        if estimate_effects:
            raise NotImplementedError("Effect estimation is not part of this engine yet.")
        cfg = self.cfg
        ops = self._assemble_reml_operators()
        if cfg.precond_rank > 0 and self.precond_conf is None:
            self._build_projected_core_precond(ops, var_components_init)
        # the SLQ whitening reference stays frozen for the whole fit
        self.slq_precond_conf = self.precond_conf
        G = len(ops.K_mvs)
        gen_atoms = None if cfg.unit_variance_components else self._diag_atoms(ops.diag_list).astype(np.float64)
        res_atoms = None if cfg.unit_variance_components else np.ones(1)
        E = 1
        reps = []
        diagnostics = None
        for r in range(max(1, int(cfg.n_reml_reps))):
            out = fit_reml(
                y=y, K_mvs=ops.K_mvs, diag_list=ops.diag_list, covar=covar, n_rand_vec=cfg.n_rand_vec,
                maxiter=cfg.max_pcg_iters, seed=cfg.seed + r * 9973, h2_init=h2_init,
                param_init=var_components_init, minq_iter=cfg.minq_iter,
                slq_samples=cfg.slq_samples, slq_m=cfg.slq_m, slq_mode=cfg.slq_mode,
                slq_workspace_bytes=default_workspace_bytes(cfg.gpu_budget_bytes),
                precond_conf=self.precond_conf, slq_precond_conf=self.precond_conf,
                precond_refresh_fn=self._make_pcg_precond_refresh_fn(ops),
                precond_eps=cfg.pcg_ridge, weighted_hv=ops.weighted_hv, stacked_kv=ops.stacked_kv,
                pcg_tol=cfg.reml_pcg_tol, response_is_standardized=cfg.response_is_standardized,
                unit_variance_components=cfg.unit_variance_components,
                max_linesearch_trials=cfg.strict_max_linesearch_trials,
                return_diagnostics=True, verbose=cfg.verbose,
                probe_cache=self._reml_probe_cache, counters=counters,
                weighted_hv_true=ops.weighted_hv_true,
            )
            vc, history, diagnostics = out
            reps.append((vc, history))

        def h2_of(theta):
            theta = np.asarray(theta, np.float64)
            if cfg.unit_variance_components:
                gen, res = theta[..., :G].sum(-1), theta[..., G:G + E].sum(-1)
            else:
                gen, res = theta[..., :G] @ gen_atoms, theta[..., G:G + E] @ res_atoms
            return gen / np.maximum(gen + res, 1e-8)

        vc_mean = reps[0][0]
        history = reps[0][1]
        rep_vc = se_var = se_h2 = None
        if len(reps) > 1:
            stack = np.stack([x[0] for x in reps])
            vc_mean = stack.mean(0)
            R = len(reps)
            se_var = np.sqrt(np.maximum(0.0, ((stack - vc_mean) ** 2).sum(0) / (R * (R - 1))))
            hv = h2_of(stack)
            se_h2 = float(np.sqrt(max(0.0, ((hv - hv.mean()) ** 2).sum() / (R * (R - 1)))))
            rep_vc = stack
            diagnostics = None
        return FitResult(
            var_components=np.asarray(vc_mean), history=history, rep_var_components=rep_vc,
            final_ai=None if diagnostics is None else diagnostics.get("ai"),
            final_grad=None if diagnostics is None else diagnostics.get("grad"),
            final_loglik=None if diagnostics is None else float(diagnostics.get("loglik")),
            diagnostics=diagnostics if cfg.capture_reml_diagnostics else None,
            heritability=float(h2_of(vc_mean)), genetic_trace_atoms=gen_atoms,
            residual_trace_atoms=res_atoms, monte_carlo_se_var=se_var, monte_carlo_se_h2=se_h2,
            counters=counters,
        )

    def close(self):
        for st in self.streamers:
            st.close()


__all__ = ["FitConfig", "FitResult", "InfinitesimalREMLFitter"]
