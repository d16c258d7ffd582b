"""GPU: full fits at the benchmark configurations' shapes against the FP64
oracle with identical probes.

* The single-GRM CPU-runnable configuration at its full size (n=5,000,
  m=20,000, 1% missing, 2 covariates + intercept, h2 = 0.5, 32 probes x 30
  Lanczos steps, PCG tolerance 1e-6): plain FP32 within |dh2| <= 1e-5 and
  theta within 1e-4 relative; the split form within |dh2| <= 1e-3.
* The eight-component precision-form configuration's shape at reduced size:
  eight contiguous components of Dirichlet(1,...,1) sizes carrying
  Dirichlet(2,...,2) x 0.6 of h2, projected-core preconditioner: the same
  bars for both forms.
"""
import numpy as np
import pytest

from oracle import geno
from oracle import reml as oreml
from tests.helpers import int8_block, random_codes

pytestmark = pytest.mark.gpu


def _pheno(o, h2_by_comp, n_covar, seed):
    rng = np.random.default_rng(seed)
    n = o.n
    Z = o.standardized()
    off = o.component_offsets
    X = np.concatenate([np.ones((n, 1)), rng.standard_normal((n, n_covar))], axis=1)
    y = X @ rng.standard_normal(n_covar + 1)
    for g, h2 in enumerate(h2_by_comp):
        a, b = int(off[g]), int(off[g + 1])
        k = max(1, int(round(0.01 * (b - a))))
        causal = rng.choice(b - a, size=k, replace=False) + a
        zb = Z[:, causal] @ rng.standard_normal(k)
        y = y + np.sqrt(h2) * (zb - zb.mean()) / max(zb.std(), 1e-12)
    y = y + np.sqrt(1.0 - float(np.sum(h2_by_comp))) * rng.standard_normal(n)
    return y.astype(np.float32), X.astype(np.float32)


def _h2(theta, atoms):
    G = len(atoms)
    gen = float(np.dot(theta[:G], atoms))
    return gen / (gen + float(theta[G]))


@pytest.mark.parametrize("precision,dh2,rtol", [("fp32", 1e-5, 1e-4), ("i8x2", 1e-3, 1e-2)])
def test_single_grm_fit_at_full_size_matches_oracle(precision, dh2, rtol):
    from coherit_amd.geno_stream import GenoBlockStreamer
    from coherit_amd.reml import fit_reml
    from coherit_amd.sources import ArraySource

    n, m = 5000, 20000
    codes = random_codes(n, m, seed=71, miss=0.01)
    st = GenoBlockStreamer(ArraySource(int8_block(codes), missing_val=-9), call_width=8192,
                           precision=precision)
    o = geno.OraclePanel.from_codes(codes, None)
    y, X = _pheno(o, [0.5], 2, seed=72)
    kw = dict(seed=11, slq_samples=32, slq_m=30, minq_iter=50, pcg_tol=1e-6)
    Hv_true = None
    if st.residual_precision != st.precision:
        Hv_true = lambda tg, te, V: st.weighted_component_hv(tg, te, V, precision=st.residual_precision)
    theta, _ = fit_reml(y, [lambda V: st.kv(V)], [st.diag()], X, 32, 400, verbose=False,
                        weighted_hv_true=Hv_true, **kw)
    Z = o.standardized()
    K = Z @ Z.T / float(o.atoms.eff)
    del Z
    pb = oreml.Problem([K], np.array([o.atoms.mean_diag], np.float64))
    th_o = oreml.fit_reml(pb, y, X, maxiter=400, **kw)
    atoms = np.array([o.atoms.mean_diag], np.float64)
    print(precision, "theta", theta, "oracle", th_o, "h2", _h2(theta, atoms), _h2(th_o, atoms),
          flush=True)
    assert abs(_h2(theta, atoms) - _h2(th_o, atoms)) <= dh2
    np.testing.assert_allclose(theta, th_o, rtol=rtol, atol=1e-6)
    st.close()


@pytest.mark.parametrize("precision,dh2,rtol", [("fp32", 1e-5, 1e-4), ("i8x2", 1e-3, 2e-2)])
def test_eight_component_fit_matches_oracle(precision, dh2, rtol):
    from coherit_amd.reml_model import FitConfig, InfinitesimalREMLFitter
    from coherit_amd.sources import ArraySource

    rng = np.random.default_rng(73)
    n, m = 400, 4000
    frac = rng.dirichlet(np.ones(8))
    sizes = np.maximum(20, np.floor(frac * m).astype(int))
    sizes[-1] = m - sizes[:-1].sum()
    assert sizes.min() > 0
    h2 = rng.dirichlet(np.full(8, 2.0)) * 0.6
    codes = random_codes(n, m, seed=74, miss=0.01)
    o = geno.OraclePanel.from_codes(codes, list(sizes))
    y, X = _pheno(o, h2, 3, seed=75)
    kw = dict(precond_rank=40, slq_samples=8, slq_m=16, minq_iter=20, reml_pcg_tol=1e-5, seed=9)
    cfg = FitConfig(sources=[ArraySource(int8_block(codes), missing_val=-9)],
                    vc_block_sizes=[int(x) for x in sizes], call_width=512, verbose=False,
                    precision=precision, **kw)
    fitter = InfinitesimalREMLFitter(cfg)
    res = fitter.fit_infinitesimal(y, X)
    U = fitter.slq_precond_conf.U.double().cpu().numpy()
    off = o.component_offsets
    Z = o.standardized()
    Ks = [Z[:, off[g]:off[g + 1]] @ Z[:, off[g]:off[g + 1]].T / float(o.atoms.component_eff[g])
          for g in range(8)]
    atoms = o.atoms.component_mean_diag.astype(np.float64)
    pb = oreml.Problem(Ks, atoms)
    th_o = oreml.fit_reml(pb, y, X, seed=9, slq_samples=8, slq_m=16, minq_iter=20, pcg_tol=1e-5,
                          slq_mode="projected_core_residual", precond_U=U)
    print(precision, "theta", np.round(res.var_components, 6), "oracle", np.round(th_o, 6),
          "h2", res.heritability, _h2(th_o, atoms), "iters", len(res.history), flush=True)
    assert len(res.var_components) == 9
    assert abs(res.heritability - _h2(th_o, atoms)) <= dh2
    np.testing.assert_allclose(res.var_components, th_o, rtol=rtol, atol=1e-4 if precision != "fp32" else 1e-6)
    fitter.close()
