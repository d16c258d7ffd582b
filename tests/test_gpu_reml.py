"""GPU: full REML fits match the FP64 oracle fit with identical probes.

Tolerances: |dh2| <= 1e-5 and theta within 1e-4 relative for the plain
FP32 path; SLQ log-determinants within 1e-5 relative.
"""
import numpy as np
import pytest
import torch

from oracle import geno
from oracle import reml as oreml
from tests.helpers import int8_block, random_codes, rel_err

pytestmark = pytest.mark.gpu


def _panel(n, m, sizes, seed, precision="fp32", call_width=512):
    from coherit_amd.geno_stream import GenoBlockStreamer
    from coherit_amd.sources import ArraySource

    codes = random_codes(n, m, seed=seed)
    st = GenoBlockStreamer(ArraySource(int8_block(codes), missing_val=-9), call_width=call_width,
                           component_block_sizes=sizes, precision=precision)
    o = geno.OraclePanel.from_codes(codes, sizes)
    return st, o


def _dense(o):
    Z = o.standardized()
    off = o.component_offsets
    Ks = []
    for g in range(len(o.component_sizes)):
        Zg = Z[:, off[g]:off[g + 1]]
        Ks.append(Zg @ Zg.T / float(o.atoms.component_eff[g]))
    return Ks


def _pheno(o, seed, h2=(0.3, 0.2)):
    rng = np.random.default_rng(seed)
    n = o.n
    X = np.concatenate([np.ones((n, 1)), rng.standard_normal((n, 2))], axis=1)
    Z = o.standardized()
    off = o.component_offsets
    y = X @ np.array([0.5, 1.0, -0.5])
    for g in range(len(o.component_sizes)):
        b = rng.standard_normal(off[g + 1] - off[g])
        zb = Z[:, off[g]:off[g + 1]] @ b
        y = y + np.sqrt(h2[g]) * zb / zb.std()
    y = y + np.sqrt(1 - sum(h2[: len(o.component_sizes)])) * rng.standard_normal(n)
    return y.astype(np.float32), X.astype(np.float32)


def _h2(theta, atoms):
    G = len(atoms)
    gen = float(np.dot(theta[:G], atoms))
    return gen / (gen + float(theta[G]))


def test_affine_single_grm_fit_matches_oracle():
    from coherit_amd.reml import fit_reml

    st, o = _panel(360, 1400, None, seed=31)
    y, X = _pheno(o, 5, h2=(0.5,))
    kw = dict(seed=3, slq_samples=12, slq_m=16, minq_iter=20, pcg_tol=1e-6)
    theta, hist = fit_reml(y, [lambda V: st.kv(V)], [st.diag()], X, 10, 400, verbose=False, **kw)
    pb = oreml.Problem(_dense(o), o.atoms.component_mean_diag.astype(np.float64))
    th_o = oreml.fit_reml(pb, y, X, maxiter=400, **kw)
    print("theta", theta, "oracle", th_o, flush=True)
    atoms = o.atoms.component_mean_diag.astype(np.float64)
    assert abs(_h2(theta, atoms) - _h2(th_o, atoms)) <= 1e-5
    np.testing.assert_allclose(theta, th_o, rtol=1e-4)
    st.close()


@pytest.mark.parametrize("precision", ["fp32", "i8x3"])
def test_two_component_multi_slq_fit_matches_oracle(precision):
    from coherit_amd.reml import fit_reml

    st, o = _panel(300, 1500, [600, 900], seed=32, precision=precision)
    y, X = _pheno(o, 6)
    # pcg_tol 1e-5: the tolerance the reference forces for multi-GRM fits (reml.py:1400-1408)
    kw = dict(seed=5, slq_samples=8, slq_m=12, minq_iter=12, pcg_tol=1e-5)
    K_mvs = [lambda V, g=g: st.component_kv(V, g) for g in range(2)]
    counters = {}
    theta, hist = fit_reml(y, K_mvs, st.component_diag_list(), X, 10, 400, verbose=False,
                           weighted_hv=st.weighted_component_hv, stacked_kv=st.stacked_component_kv,
                           counters=counters, **kw)
    pb = oreml.Problem(_dense(o), o.atoms.component_mean_diag.astype(np.float64))
    th_o, hist_o, info = oreml.fit_reml(pb, y, X, maxiter=400, return_history=True, **kw)
    print("theta", theta, "oracle", th_o, counters, flush=True)
    for h in hist:
        print("gpu   ", h["iter"], h["status"], f"{h['loglik']:.10f}", h["params"],
              [(t["alpha"], t["dll"]) for t in h["line_search_trace"]], flush=True)
    for h in hist_o:
        print("oracle", h["iter"], h["status"], h.get("loglik"), h["params"], flush=True)
    atoms = o.atoms.component_mean_diag.astype(np.float64)
    assert abs(_h2(theta, atoms) - _h2(th_o, atoms)) <= 1e-5
    np.testing.assert_allclose(theta, th_o, rtol=1e-4, atol=1e-7)
    st.close()


def test_fitter_with_projected_core_whitening_matches_oracle():
    from coherit_amd.reml_model import FitConfig, InfinitesimalREMLFitter
    from coherit_amd.sources import ArraySource

    codes = random_codes(320, 1200, seed=33)
    sizes = [500, 700]
    cfg = FitConfig(sources=[ArraySource(int8_block(codes), missing_val=-9)], vc_block_sizes=sizes,
                    call_width=400, precond_rank=40, slq_samples=8, slq_m=12, minq_iter=12,
                    reml_pcg_tol=1e-5, seed=7, verbose=False)
    fitter = InfinitesimalREMLFitter(cfg)
    o = geno.OraclePanel.from_codes(codes, sizes)
    y, X = _pheno(o, 7)
    res = fitter.fit_infinitesimal(y, X)
    U = fitter.slq_precond_conf.U.double().cpu().numpy()
    pb = oreml.Problem(_dense(o), o.atoms.component_mean_diag.astype(np.float64))
    th_o = oreml.fit_reml(pb, y, X, seed=7, slq_samples=8, slq_m=12, minq_iter=12, pcg_tol=1e-5,
                          slq_mode="projected_core_residual", precond_U=U)
    print("theta", res.var_components, "oracle", th_o, "h2", res.heritability, flush=True)
    atoms = o.atoms.component_mean_diag.astype(np.float64)
    assert abs(res.heritability - _h2(th_o, atoms)) <= 1e-5
    np.testing.assert_allclose(res.var_components, th_o, rtol=1e-4, atol=1e-7)
    fitter.close()


def test_split_form_fit_refines_pcg_and_stays_within_tolerance():
    """2-limb contraction drives PCG/SLQ; 3-limb true residuals refine PCG to
    1e-5.  The chosen form's bar: |dh2| <= 1e-3 against the FP64 oracle."""
    from coherit_amd.reml_model import FitConfig, InfinitesimalREMLFitter
    from coherit_amd.sources import ArraySource

    codes = random_codes(320, 1200, seed=33)
    sizes = [500, 700]
    cfg = FitConfig(sources=[ArraySource(int8_block(codes), missing_val=-9)], vc_block_sizes=sizes,
                    call_width=400, precond_rank=40, slq_samples=8, slq_m=12, minq_iter=12,
                    reml_pcg_tol=1e-5, seed=7, verbose=False, precision="i8x2")
    fitter = InfinitesimalREMLFitter(cfg)
    o = geno.OraclePanel.from_codes(codes, sizes)
    y, X = _pheno(o, 7)
    counters = {}
    res = fitter.fit_infinitesimal(y, X, counters=counters)
    U = fitter.slq_precond_conf.U.double().cpu().numpy()
    pb = oreml.Problem(_dense(o), o.atoms.component_mean_diag.astype(np.float64))
    th_o = oreml.fit_reml(pb, y, X, seed=7, slq_samples=8, slq_m=12, minq_iter=12, pcg_tol=1e-5,
                          slq_mode="projected_core_residual", precond_U=U)
    atoms = o.atoms.component_mean_diag.astype(np.float64)
    print("theta", res.var_components, "oracle", th_o, "h2", res.heritability, counters, flush=True)
    assert counters.get("hv_true", 0) > 0
    assert abs(res.heritability - _h2(th_o, atoms)) <= 1e-3
    np.testing.assert_allclose(res.var_components, th_o, rtol=1e-2, atol=1e-4)
    fitter.close()


def test_slq_value_and_gradient_do_not_depend_on_probe_batching():
    """workspace_bytes only sets how many probes share a batch (slq.py:24-41;
    tests/test_multi_slq_score_consistency.py:100-107)."""
    from coherit_amd import slq

    rng = np.random.default_rng(12)
    n = 200
    Ks = []
    for w in (30, 55):
        Z = rng.standard_normal((n, w))
        Ks.append(torch.as_tensor(Z @ Z.T / w, dtype=torch.float32, device="cuda"))
    th = [0.4, 0.3, 0.5]
    H = th[0] * Ks[0] + th[1] * Ks[1] + th[2] * torch.eye(n, device="cuda")

    def mv(V):
        return H @ V

    def pull(left, right):
        prods = [Ks[0] @ left, Ks[1] @ left, left]
        return H @ left, torch.stack([(right * p).sum() for p in prods]).double().cpu().numpy()

    out = []
    for ws in (10**12, slq.workspace_layout(n, 16, 12, 1)["peak_bytes"] // 2 + 1, 1):
        print("workspace", ws, slq.workspace_layout(n, 16, 12, ws)["batch_width"], flush=True)
        out.append(slq.logdet_value_and_grad(mv, pull, n, 5, nsamples=16, m=12, workspace_bytes=ws))
    for v, g in out[1:]:
        assert float(v) == pytest.approx(float(out[0][0]), rel=1e-5)
        np.testing.assert_allclose(np.asarray(g), np.asarray(out[0][1]), rtol=1e-4)


def test_reorth_primitives_match_fp64_products():
    """coh_probe_dots / _project_out / _rank_update against FP64 torch on the
    device, on a probe-major basis view with depth padding (slq.py:49-74)."""
    from coherit_amd import slq

    torch.manual_seed(0)
    S, D, r, n = 5, 9, 6, 5000 + 37
    basis = torch.randn(S, D, n, device="cuda")
    P = basis[:, :r, :]
    x = torch.randn(S, n, device="cuda")
    c = slq._probe_dots(P, x)
    ref = torch.einsum("sjn,sn->sj", P.double(), x.double())
    torch.testing.assert_close(c.double(), ref, rtol=1e-6, atol=1e-4)
    y = slq._project_out(P, c, x)
    ref_y = x.double() - torch.einsum("sjn,sj->sn", P.double(), c.double())
    torch.testing.assert_close(y.double(), ref_y, rtol=1e-5, atol=1e-4)
    bb = torch.zeros(S, D, n, device="cuda")
    coeff = torch.randn(S, r, 4, device="cuda")
    vecs = torch.randn(S, 4, n, device="cuda")
    from coherit_amd import _native as N
    view = bb[:, :r, :]
    N.call("coh_probe_rank_update", S, r, n, 4, N.ptr(coeff), coeff.stride(0), coeff.stride(1),
           N.ptr(vecs), vecs.stride(0), vecs.stride(1), N.ptr(view), view.stride(0),
           view.stride(1), N.stream_ptr())
    ref_b = torch.einsum("sjq,sqn->sjn", coeff.double(), vecs.double())
    torch.testing.assert_close(bb[:, :r, :].double(), ref_b, rtol=1e-5, atol=1e-5)
    assert float(bb[:, r:, :].abs().max()) == 0.0
    # the two-pass step equals the torch formulation on the same inputs
    c1, w2, c2, w3 = slq._reorth(P, x)
    Pd, xd = P.double(), x.double()
    c1r = torch.bmm(Pd, xd[:, :, None])[:, :, 0]
    w2r = xd - torch.bmm(Pd.transpose(1, 2), c1r[:, :, None])[:, :, 0]
    torch.testing.assert_close(w2.double(), w2r, rtol=1e-5, atol=1e-3)


@pytest.mark.parametrize("precision", ["fp32", "i8x2"])
def test_effects_and_prediction_match_oracle(precision):
    """estimate_effects / predict (reml_model.py:1544-1816) against the FP64
    oracle at fixed variance components; test panel standardized with the
    training statistics."""
    from coherit_amd.reml_model import FitConfig, InfinitesimalREMLFitter
    from coherit_amd.sources import ArraySource

    sizes = [500, 700]
    codes = random_codes(320, 1200, seed=41)
    codes_test = random_codes(150, 1200, seed=42)
    kw = dict(vc_block_sizes=sizes, call_width=400, precond_rank=40, seed=3, verbose=False,
              effect_pcg_tol=1e-6, precision=precision)
    fitter = InfinitesimalREMLFitter(FitConfig(
        sources=[ArraySource(int8_block(codes), missing_val=-9)], **kw))
    test_fitter = InfinitesimalREMLFitter(FitConfig(
        sources=[ArraySource(int8_block(codes_test), missing_val=-9)], **kw))
    o = geno.OraclePanel.from_codes(codes, sizes)
    y, X = _pheno(o, 9)
    rng = np.random.default_rng(5)
    X_test = np.concatenate([np.ones((150, 1)), rng.standard_normal((150, 2))], 1).astype(np.float32)
    theta = np.array([0.25, 0.2, 0.55], np.float32)
    eff = fitter.estimate_effects(y, var_components=theta, covar=X)
    pred = fitter.predict(eff, test_fitter=test_fitter, test_covar=X_test)

    off = o.component_offsets
    Z = o.standardized()
    Zs = [Z[:, off[g]:off[g + 1]] for g in range(2)]
    pb = oreml.Problem(_dense(o), o.atoms.component_mean_diag.astype(np.float64))
    ref = oreml.estimate_effects(pb, Zs, o.atoms.component_eff, theta, y, X)
    ot = geno.OraclePanel.from_codes(codes_test, sizes, override=(o.mean, o.inv_sd))
    Zt = ot.standardized()
    ref_p = oreml.predict(ref, [Zt[:, off[g]:off[g + 1]] for g in range(2)], X_test)

    tol = 1e-4 if precision == "fp32" else 2e-3
    print(precision, "pcg", eff.pcg_iters, eff.pcg_rel_res, flush=True)
    np.testing.assert_allclose(eff.fixed_effects.cpu().numpy(), ref["beta"], rtol=tol, atol=tol)
    assert rel_err(eff.random_effect.cpu().numpy(), ref["random"]) < tol
    for g in range(2):
        assert rel_err(eff.random_effect_components[g].cpu().numpy(), ref["comps"][g]) < tol
        assert rel_err(eff.snp_effects[g].cpu().numpy(), ref["snp"][g]) < tol
    assert eff.y_mean == pytest.approx(ref["y_mean"], rel=1e-6)
    assert eff.y_scale == pytest.approx(ref["y_scale"], rel=1e-6)
    assert rel_err(pred.y_pred_std.cpu().numpy(), ref_p["y_pred_std"]) < tol
    assert rel_err(pred.y_pred.cpu().numpy(), ref_p["y_pred"]) < tol
    for g in range(2):
        assert rel_err(pred.random_effect_components[g].cpu().numpy(), ref_p["comps"][g]) < tol
    fitter.close()
    test_fitter.close()
