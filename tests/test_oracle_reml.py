"""CPU checks that pin the REML/SLQ oracle (finite differences, closed forms,
the exact dense REML likelihood)."""
import numpy as np
import pytest

from oracle import reml as oreml
from oracle import slq as oslq


def small_problem(G=2, n=48, seed=3):
    rng = np.random.default_rng(seed)
    Ks, atoms = [], []
    for g in range(G):
        Z = rng.standard_normal((n, 30 + 17 * g))
        Z = (Z - Z.mean(0)) / Z.std(0)
        K = Z @ Z.T / Z.shape[1]
        Ks.append(K)
        atoms.append(np.trace(K) / n)
    X = np.concatenate([np.ones((n, 1)), rng.standard_normal((n, 2))], axis=1)
    L = np.linalg.cholesky(sum(0.3 * K for K in Ks) + 0.5 * np.eye(n))
    y = X @ np.array([1.0, 0.5, -0.3]) + L @ rng.standard_normal(n)
    return oreml.Problem(Ks, atoms), y, X


def _objective(pb, p, y, X, kw):
    return oreml.eval_once(pb, p, y, X, None, **kw)


@pytest.mark.parametrize("G", [1, 2])
def test_oracle_score_is_derivative_of_its_objective(G):
    pb, y, X = small_problem(G)
    ys, _, _ = oreml.standardize_response(y)
    ys = ys.astype(np.float64)
    p = np.concatenate([np.full(G, 0.25), [0.6]])
    affine = oreml.affine_cache(pb.Ks[0], 11, 8, 20) if G == 1 else None
    kw = dict(key=11, tol=1e-10, maxiter=500, slq_samples=8, slq_m=20, slq_mode="raw",
              core=None, ref=None, eps=1e-6, affine=affine)
    ll, grad, AI, *_ = _objective(pb, p, ys, X, kw)
    h = 1e-5
    for i in range(p.size):
        e = np.eye(p.size)[i] * h
        fd = (_objective(pb, p + e, ys, X, kw)[0] - _objective(pb, p - e, ys, X, kw)[0]) / (2 * h)
        assert grad[i] == pytest.approx(fd, rel=1e-4, abs=1e-7)
    # the AI direction is an ascent direction (tests/test_affine_slq_score_consistency.py:128-146)
    step = np.linalg.solve(AI + 1e-4 * np.eye(p.size), grad)
    assert float(step @ grad) > 0


def test_affine_logdet_derivative_in_theta_e_is_n_at_zero_theta_g():
    # tests/test_affine_slq_score_consistency.py:149-170
    pb, _, _ = small_problem(1)
    n = pb.Ks[0].shape[0]
    A, B = oreml.affine_cache(pb.Ks[0], 5, 6, 12)
    ld, d = oreml.affine_logdet(A, B, n, 6, 0.0, 1.0)
    assert ld == pytest.approx(0.0, abs=1e-9)
    assert d[1] == pytest.approx(n, rel=1e-9)
    assert np.isfinite(d[0])


def test_affine_equals_rerunning_lanczos_on_h():
    # Krylov shift/scale invariance: T_H = theta_g T_K + theta_e I (reml.py:421-428)
    pb, _, _ = small_problem(1)
    n = pb.Ks[0].shape[0]
    K = pb.Ks[0]
    A, B = oreml.affine_cache(K, 9, 5, 10)
    Ah, Bh = oreml.affine_cache(0.7 * K + 0.4 * np.eye(n), 9, 5, 10)
    np.testing.assert_allclose(Ah, 0.7 * A + 0.4, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(Bh, 0.7 * B, rtol=1e-9, atol=1e-12)


def test_oracle_slq_gradient_matches_finite_differences():
    pb, _, _ = small_problem(2)
    n = pb.Ks[0].shape[0]
    D = pb.derivs()
    p = np.array([0.3, 0.2, 0.5])

    def run(pp, pull=True):
        H = sum(pp[i] * D[i] for i in range(3))
        mv = lambda v: H @ v
        pb_ = (lambda l, r: (H @ l, np.array([np.sum(r * (Dk @ l)) for Dk in D]))) if pull else None
        return oslq.logdet_value_and_grad(mv, pb_, n, 21, nsamples=4, m=15)

    v, g = run(p)
    for i in range(3):
        e = np.eye(3)[i] * 1e-6
        fd = (run(p + e, False)[0] - run(p - e, False)[0]) / 2e-6
        assert g[i] == pytest.approx(fd, rel=1e-4, abs=1e-6)


def test_oracle_fit_reaches_exact_reml_optimum_neighbourhood():
    # With many probes and full depth the fixed-probe objective approaches the
    # exact REML likelihood; the fitted theta must be near-stationary for it.
    pb, y, X = small_problem(1, n=48)
    ys = oreml.standardize_response(y)[0].astype(np.float64)
    theta = oreml.fit_reml(pb, y, X, slq_samples=64, slq_m=48, minq_iter=30, rel_dll_tol=1e-9,
                           pcg_tol=1e-8)
    f = lambda t: oreml.exact_reml_loglik(pb, t, ys, X)
    best = f(theta)
    for d in (np.array([0.05, 0]), np.array([0, 0.05]), np.array([-0.05, 0]), np.array([0, -0.05])):
        cand = np.maximum(theta + d * theta, 1e-3)
        assert f(cand) <= best + 5e-3


def test_oracle_effects_satisfy_gls_and_blup_identities():
    # At the GLS/BLUP solution: X^T alpha = 0 and
    # sum_g theta_g K_g alpha + theta_e alpha = y_s - X beta (H alpha = y_s - X beta).
    pb, y, X = small_problem(2)
    rng = np.random.default_rng(8)
    Zs = [rng.standard_normal((pb.Ks[0].shape[0], 7)) for _ in range(2)]
    theta = np.array([0.3, 0.2, 0.5])
    eff = oreml.estimate_effects(pb, Zs, [7.0, 7.0], theta, y, X)
    np.testing.assert_allclose(X.T @ eff["alpha"], 0.0, atol=1e-9)
    th = theta.astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(eff["random"] + th[2] * eff["alpha"], eff["ys"] - X @ eff["beta"],
                               rtol=1e-9, atol=1e-9)
    # SNP effects reproduce the random effect of a kernel built from the same Z
    for g in range(2):
        np.testing.assert_allclose(eff["snp"][g], th[g] / 7.0 * Zs[g].T @ eff["alpha"], rtol=1e-12)
    pred = oreml.predict(eff, Zs, X)
    np.testing.assert_allclose(pred["y_pred"], eff["y_mean"] + eff["y_scale"] * pred["y_pred_std"])
