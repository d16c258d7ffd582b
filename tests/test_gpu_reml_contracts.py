"""GPU: the strict optimiser's work contracts, as the reference's
lazy-evaluation tests state them (tests/test_reml_lazy_evaluation.py:34-54,
77-104): strict REML solves only the [X | y] and AI right-hand sides, the
Hutchinson probe count is inert, the probe cache keeps its signature, and
the SLQ accept gate never repeats forward products and skips the reverse
sweep of a rejected value.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _problem(groups):
    rng = np.random.default_rng(684)
    n = 24
    Ks = []
    for k in range(groups):
        x = rng.normal(size=(n, 12 + 5 * k)).astype(np.float32)
        Ks.append(torch.as_tensor(x @ x.T / x.shape[1] + 0.1 * np.eye(n, dtype=np.float32),
                                  device="cuda"))
    y = rng.normal(size=n).astype(np.float32)
    return dict(
        y=y, K_mvs=tuple((lambda b, K=K: K @ b) for K in Ks),
        diag_list=tuple(torch.diagonal(K).clone() for K in Ks),
        covar=np.ones((n, 1), np.float32), n_rand_vec=7, maxiter=100, minq_iter=1,
        slq_samples=4, slq_m=8, pcg_tol=1e-5, verbose=False, response_is_standardized=True,
    )


@pytest.mark.parametrize("groups", [1, 2])
def test_strict_solves_only_xy_and_ai_columns_and_probe_count_is_inert(monkeypatch, groups):
    from coherit_amd import reml as R

    solve = R.pcg_solve
    widths = []

    def observed(hv, rhs, *a, **kw):
        widths.append(int(rhs.shape[1]))
        return solve(hv, rhs, *a, **kw)

    monkeypatch.setattr(R, "pcg_solve", observed)
    cache = R.REMLProbeCache()
    args = _problem(groups)
    first = R.fit_reml(**args, probe_cache=cache)
    signature = cache.signature
    second = R.fit_reml(**{**args, "n_rand_vec": 17}, probe_cache=cache)
    print("widths", sorted(set(widths)), "theta", first[0], flush=True)
    # [X | y] has p + 1 = 2 columns; the AI solve G + 1 (one per component and residual)
    assert set(widths) <= {2, groups + 1}
    assert cache.signature == signature
    assert cache.kvrand_stack is None
    np.testing.assert_array_equal(first[0], second[0])


@pytest.mark.parametrize("budget", [1, 50000, 1 << 20])
@pytest.mark.parametrize("accepted", [False, True])
def test_slq_gate_preserves_value_and_never_repeats_forward_products(budget, accepted):
    from coherit_amd import slq

    calls = {"forward": 0, "reverse": 0}
    matrix = torch.diag(torch.linspace(0.3, 2.0, 24, device="cuda"))

    def mv(b):
        calls["forward"] += 1
        return matrix @ b

    def pullback(left, right):
        calls["reverse"] += 1
        product = matrix @ left
        return product, np.array([float((product * right).sum())])

    args = dict(nsamples=6, m=8, workspace_bytes=budget)
    expected = slq.logdet_value_and_grad(mv, pullback, 24, 54, **args)
    eager = dict(calls)
    calls.update(forward=0, reverse=0)
    actual = slq.logdet_value_and_grad(mv, pullback, 24, 54, **args,
                                       accept_value=lambda value: accepted)
    assert actual[0] == expected[0]
    assert calls["forward"] == eager["forward"]
    if accepted:
        np.testing.assert_array_equal(actual[1], expected[1])
        assert calls["reverse"] == eager["reverse"]
    else:
        assert actual[1] is None
        # a budget too small to retain the basis runs the reverse sweep eagerly
        assert calls["reverse"] == (eager["reverse"] if budget == 1 else 0)
