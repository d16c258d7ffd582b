"""Oracle: block PCG restated in NumPy (pcg.py:31-167).  Test infrastructure only.

Works in the dtype of B (FP32 reproduces the reference's arithmetic width,
FP64 gives the exact-arithmetic limit).  Pinned by the reference's eight
behaviour tests (tests/test_pcg_stability.py), restated in
tests/test_oracle_pcg.py.
"""
from __future__ import annotations

import math

import numpy as np


def _identity(v):
    return v


def pcg_state_update(X, R, P, rs, AP, active):
    """pcg.py:31-48."""
    denom = np.sum(P * AP, axis=0)
    valid = (rs > 0) & (denom > 0) & np.isfinite(rs) & np.isfinite(denom)
    moving = active & valid
    alpha = np.where(moving, rs / np.where(moving, denom, 1), 0).astype(X.dtype)
    return X + P * alpha, R - AP * alpha, bool(np.any(active & ~valid))


def pcg_direction_update(P, R_new, Z_new, rs_old, threshold):
    """pcg.py:51-65."""
    rs_new = np.sum(R_new * Z_new, axis=0)
    rnorm = np.linalg.norm(R_new, axis=0)
    active = rnorm > threshold
    beta = np.where(active, rs_new / np.where(rs_old > 0, rs_old, 1), 0).astype(P.dtype)
    P_new = np.where(active, Z_new + P * beta, 0).astype(P.dtype)
    return P_new, np.where(active, rs_new, 0), rnorm, active


def pcg_solve(Hv, B, M=None, tol=1e-2, maxiter=200, X0=None, check_every=2, counter=None,
              Hv_true=None):
    """pcg.py:68-167; returns (X, max_rel_res, iters).

    ``Hv_true`` (an extension; default ``Hv``) is the operator of the true
    residuals, which turns the verify-and-restart step into iterative
    refinement when ``Hv`` is a cheaper, less accurate form.
    """
    if M is None:
        M = _identity
    if Hv_true is None:
        Hv_true = Hv
    check_every = max(1, int(check_every))
    B = np.asarray(B)
    dt = B.dtype
    magnitude = np.max(np.abs(B), axis=0)
    nonzero = magnitude > 0
    _, exponent = np.frexp(magnitude)
    rhs_scale = np.where(nonzero, np.ldexp(np.ones_like(magnitude), exponent - 1), 1).astype(dt)
    b_norm = np.where(nonzero, np.linalg.norm(B / rhs_scale, axis=0), 1)
    threshold = tol * b_norm
    solution = np.zeros_like(B) if X0 is None else np.where(nonzero, np.asarray(X0, dt), 0).astype(dt)

    def op(f, v):
        if counter is not None:
            counter[0] += 1
        return np.asarray(f(v), dtype=dt)

    def true_residual(value):
        residual = ((B - op(Hv_true, value)) / rhs_scale).astype(dt)
        norms = np.linalg.norm(residual, axis=0)
        rel = float(np.max(norms / b_norm))
        if not math.isfinite(rel):
            raise FloatingPointError("Non-finite true residual in PCG.")
        return residual, norms, rel

    R, rnorm, rel = true_residual(solution)
    if rel <= tol or maxiter <= 0:
        return solution, rel, 0
    X = (solution / rhs_scale).astype(dt)
    active = rnorm > threshold
    Z = np.asarray(M(R), dt)
    P = np.where(active, Z, 0).astype(dt)
    rs = np.where(active, np.sum(R * Z, axis=0), 0)
    k = 0
    breakdown = False
    while k < maxiter:
        AP = op(Hv, P)
        X, R, bad = pcg_state_update(X, R, P, rs, AP, active)
        breakdown = breakdown or bad
        Z = np.asarray(M(R), dt)
        P, rs, rnorm, active = pcg_direction_update(P, R, Z, rs, threshold)
        k += 1
        if k % check_every == 0 or k == maxiter:
            recursive = math.inf if breakdown else float(np.max(rnorm / b_norm))
            if not math.isfinite(recursive):
                raise FloatingPointError(
                    "PCG encountered nonpositive curvature or a non-finite recurrence.")
            if recursive <= tol or k == maxiter:
                solution = (X * rhs_scale).astype(dt)
                R, rnorm, rel = true_residual(solution)
                if rel <= tol or k == maxiter:
                    return solution, rel, k
                X = (solution / rhs_scale).astype(dt)
                active = rnorm > threshold
                Z = np.asarray(M(R), dt)
                P = np.where(active, Z, 0).astype(dt)
                rs = np.where(active, np.sum(R * Z, axis=0), 0)
    return (X * rhs_scale).astype(dt), rel, k
