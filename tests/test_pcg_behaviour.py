"""The reference's PCG behaviour tests (tests/test_pcg_stability.py), run on
the NumPy oracle (CPU) and on the HIP-backed product pcg_solve (GPU)."""
import math

import numpy as np
import pytest

from oracle import pcg as opcg


class OracleBackend:
    name = "oracle"

    def arr(self, x):
        return np.asarray(x, dtype=np.float32)

    def np(self, x):
        return np.asarray(x)

    def solve(self, *a, **k):
        return opcg.pcg_solve(*a, **k)

    def matmul(self, h):
        return lambda v: (h @ v).astype(np.float32)

    def zeros_like(self, x):
        return np.zeros_like(x)


class GpuBackend:
    name = "gpu"

    def __init__(self):
        import torch

        self.torch = torch

    def arr(self, x):
        return self.torch.as_tensor(np.asarray(x, dtype=np.float32), device="cuda")

    def np(self, x):
        return x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)

    def solve(self, *a, **k):
        from coherit_amd.pcg import pcg_solve

        return pcg_solve(*a, **k)

    def matmul(self, h):
        return lambda v: h @ v

    def zeros_like(self, x):
        return self.torch.zeros_like(x)


@pytest.fixture(params=["oracle", pytest.param("gpu", marks=pytest.mark.gpu)])
def be(request):
    return OracleBackend() if request.param == "oracle" else GpuBackend()


def _true_relative(be, hv, x, b):
    residual = np.asarray(be.np(b), np.float64) - np.asarray(be.np(hv(x)), np.float64)
    rhs = np.asarray(be.np(b), np.float64)
    den = np.linalg.norm(rhs, axis=0)
    return float(np.max(np.linalg.norm(residual, axis=0) / np.where(den > 0, den, 1)))


@pytest.mark.parametrize("check_every", [1, 2, 5])
@pytest.mark.parametrize("preconditioned", [False, True])
def test_diagonal_solve_handles_zero_and_extreme_rhs_scales(be, check_every, preconditioned):
    scales = np.asarray([0, 1e-30, 1e-8, 1, 1e30], dtype=np.float32)
    b = be.arr(np.arange(1, 17, dtype=np.float32)[:, None] * scales)
    calls = []

    def hv(value):
        calls.append(None)
        return 2 * value

    pre = (lambda v: v / 2) if preconditioned else None
    x, rel, it = be.solve(hv, b, M=pre, tol=1e-6, maxiter=20, check_every=check_every)
    assert len(calls) == it + 2
    xn = be.np(x)
    np.testing.assert_allclose(xn[:, 1:] / scales[1:],
                               np.broadcast_to(np.arange(1, 17)[:, None] / 2, (16, 4)), rtol=1e-6)
    np.testing.assert_array_equal(xn[:, 0], 0)
    assert float(rel) == pytest.approx(_true_relative(be, lambda v: 2 * v, x, b), abs=1e-8)
    assert float(rel) <= 1e-6
    assert it <= check_every


@pytest.mark.parametrize("preconditioned", [False, True])
def test_spd_solution_with_scaled_rhs_and_warm_start_matches_dense(be, preconditioned):
    rng = np.random.default_rng(237)
    root = rng.normal(size=(24, 24))
    h = (root @ root.T / 24 + np.eye(24)).astype(np.float32)
    rhs = rng.normal(size=(24, 3)).astype(np.float32)
    scales = np.asarray([1e-20, 1, 1e20], dtype=np.float32)
    b = be.arr(rhs * scales)
    x0 = 0.1 * b
    H = be.arr(h)
    hv = be.matmul(H)
    dg = be.arr(np.diag(h)[:, None])
    pre = (lambda v: v / dg) if preconditioned else None
    x, rel, _ = be.solve(hv, b, M=pre, X0=x0, tol=1e-6, maxiter=100)
    expected = np.linalg.solve(h.astype(np.float64), rhs)
    np.testing.assert_allclose(be.np(x) / scales, expected, atol=2e-6, rtol=2e-5)
    assert float(rel) <= 1e-6
    assert float(rel) == pytest.approx(_true_relative(be, hv, x, b), rel=2e-6, abs=1e-12)


def test_vector_input_and_converged_warm_start_are_preserved(be):
    diagonal = be.arr(np.arange(1, 9))
    hv = lambda v: diagonal * v
    b = be.arr(np.arange(8) * 1e-20)
    x0 = b / diagonal
    x, rel, it = be.solve(hv, b, X0=x0, tol=1e-6)
    assert tuple(x.shape) == tuple(b.shape)
    np.testing.assert_array_equal(be.np(x), be.np(x0))
    assert it == 0
    assert float(rel) == pytest.approx(_true_relative(be, hv, x, b), abs=1e-8)


def test_zero_rhs_ignores_nonzero_warm_start_and_identity_may_alias(be):
    b = be.arr([[0, 1], [0, -2], [0, 3]])
    x, rel, _ = be.solve(lambda v: v, b, X0=be.arr(np.ones((3, 2))), tol=1e-6)
    np.testing.assert_array_equal(be.np(x), be.np(b))
    assert float(rel) == 0


@pytest.mark.parametrize("maxiter", [1, 314])
def test_ill_conditioned_return_reports_true_residual_even_at_budget(be, maxiter):
    rng = np.random.default_rng(12)
    basis = np.linalg.qr(rng.normal(size=(60, 60)))[0]
    h = ((basis * np.geomspace(1, 1e4, 60)) @ basis.T).astype(np.float32)
    b = be.arr(rng.normal(size=(60, 2)))
    hv = be.matmul(be.arr(h))
    x, rel, it = be.solve(hv, b, tol=1e-5, maxiter=maxiter)
    assert it <= maxiter
    assert float(rel) == pytest.approx(_true_relative(be, hv, x, b), rel=2e-6, abs=1e-10)


@pytest.mark.parametrize("maxiter", [1, 100])
def test_optimistic_recurrence_is_verified_and_restarted_within_budget(be, monkeypatch, maxiter):
    if be.name == "oracle":
        mod, name = opcg, "pcg_state_update"
    else:
        import coherit_amd.pcg as mod

        name = "_pcg_state_update"
    original = getattr(mod, name)
    calls = {"updates": 0, "operator": 0}

    def lose_residual_once(*args):
        x, residual, bd = original(*args)
        calls["updates"] += 1
        if calls["updates"] == 1:
            residual = be.zeros_like(residual)
        return x, residual, bd

    diagonal = be.arr(np.linspace(1, 4, 20)[:, None])
    b = be.arr(np.ones((20, 1)))

    def hv(v):
        calls["operator"] += 1
        return diagonal * v

    monkeypatch.setattr(mod, name, lose_residual_once)
    x, rel, it = be.solve(hv, b, tol=1e-6, maxiter=maxiter, check_every=1)
    ops = calls["operator"]
    assert float(rel) == pytest.approx(_true_relative(be, lambda v: diagonal * v, x, b), rel=2e-6,
                                       abs=1e-10)
    assert it <= maxiter
    if maxiter == 1:
        assert float(rel) > 1e-6
    else:
        assert ops > it + 2
        assert float(rel) <= 1e-6
        np.testing.assert_allclose(be.np(x), be.np(b) / be.np(diagonal), rtol=3e-6)


def test_zero_iteration_budget_returns_verified_initial_residual(be):
    b = be.arr(np.ones((5, 2)) * 1e-20)
    x, rel, it = be.solve(lambda v: 2 * v, b, maxiter=0)
    np.testing.assert_array_equal(be.np(x), 0)
    assert it == 0
    assert float(rel) == pytest.approx(1)


@pytest.mark.parametrize("factor", [0.0, -1.0, math.nan])
def test_invalid_operator_cannot_be_reported_as_converged(be, factor):
    with pytest.raises(FloatingPointError):
        be.solve(lambda v: factor * v, be.arr(np.ones((8, 1))), maxiter=4)


def _quantized(be, h, bits):
    """H.V with V rounded per column to ``bits`` bits of its max magnitude: a
    deterministic, slightly nonlinear operator like the limb contraction."""
    h_np = np.asarray(h, np.float64)

    def apply(v):
        vv = np.asarray(be.np(v), np.float64)
        vv = vv[:, None] if vv.ndim == 1 else vv
        step = np.max(np.abs(vv), axis=0, keepdims=True) / 2.0**bits
        step = np.where(step > 0, step, 1.0)
        q = np.round(vv / step) * step
        return be.arr(h_np @ q)

    return apply


def test_refinement_reaches_tolerance_below_operator_floor(be):
    rng = np.random.default_rng(3)
    n = 96
    A = rng.standard_normal((n, n))
    h = A @ A.T / n + np.eye(n)
    b = be.arr(rng.standard_normal((n, 3)))
    exact = be.matmul(be.arr(h))
    coarse = _quantized(be, h, 10)
    # the coarse operator alone cannot verify 1e-6
    _, rel_c, _ = be.solve(coarse, b, tol=1e-6, maxiter=300)
    assert float(rel_c) > 1e-6
    x, rel, it = be.solve(coarse, b, tol=1e-6, maxiter=300, Hv_true=exact)
    assert float(rel) <= 1e-6
    assert float(rel) == pytest.approx(_true_relative(be, exact, x, b), rel=1e-3, abs=1e-9)
    np.testing.assert_allclose(be.np(x), np.linalg.solve(h, be.np(b)), rtol=1e-4, atol=1e-5)
