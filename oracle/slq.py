"""Oracle: fixed-probe SLQ value + gradient in NumPy (slq.py:16-213).

Test infrastructure only.  Forward Lanczos with two re-orthogonalisation
passes (slq.py:49-74), FP64 quadrature (slq.py:84-121) and the reverse sweep
(slq.py:124-140) with the adjoint of the step.  Pinned by central finite
differences of its own value and by the dense probe oracle at full depth
(tests/test_oracle_slq_reml.py).
"""
from __future__ import annotations

import numpy as np

EPS32 = float(np.finfo(np.float32).eps)


def workspace_layout(n, nsamples, m, workspace_bytes, *, itemsize=4):
    """slq.py:24-41."""
    depth = min(int(m), int(n))
    per_probe = itemsize * n * (8 * (depth + 1) + 32)
    capacity = max(1, int(workspace_bytes) // per_probe)
    width = min(int(nsamples), 1 << (capacity.bit_length() - 1))
    retained = itemsize * n * (nsamples * (2 * depth + 1) + width * (6 * (depth + 1) + 32))
    return dict(batch_width=width, retained_bytes=retained)


def rademacher_probes(seed, n, nsamples):
    """The project's probe stream: (nsamples, n) +-1 from PCG64(seed)."""
    rng = np.random.default_rng(int(seed))
    return (rng.integers(0, 2, size=(int(nsamples), int(n)), dtype=np.int8) * 2 - 1).astype(np.float32)


def _parts(basis, product, prev_beta, index):
    q = basis[:, index]
    prev = basis[:, index - 1] if index > 0 else np.zeros_like(q)
    active = (q * q).sum(1) > 0
    w0 = product - prev * prev_beta[:, None]
    alpha = (q * w0).sum(1)
    w1 = w0 - q * alpha[:, None]
    P = basis[:, : index + 1]
    c1 = np.einsum("srn,sn->sr", P, w1)
    w2 = w1 - np.einsum("srn,sr->sn", P, c1)
    c2 = np.einsum("srn,sn->sr", P, w2)
    w3 = w2 - np.einsum("srn,sr->sn", P, c2)
    sq = (w3 * w3).sum(1)
    norm = np.sqrt(np.where(sq > 0, sq, 1))
    scale = np.sqrt((product * product).sum(1))
    thr = 16 * EPS32 * np.maximum(scale, np.abs(alpha))
    keep = active & (sq > thr * thr)
    return locals()


def step(basis, product, prev_beta, index):
    s = _parts(basis, product, prev_beta, index)
    keep = s["keep"]
    beta = np.where(keep, s["norm"], 0)
    nq = np.where(keep[:, None], s["w3"] / np.where(keep, s["norm"], 1)[:, None], 0)
    return nq, np.where(s["active"], s["alpha"], 1), beta


def step_pullback(basis, product, prev_beta, index, nq_bar, a_bar, b_bar, basis_bar):
    s = _parts(basis, product, prev_beta, index)
    keep = s["keep"][:, None]
    norm = s["norm"][:, None]
    nxt = s["w3"] / norm
    w3b = np.where(keep, (nq_bar - nxt * (nxt * nq_bar).sum(1, keepdims=True)) / norm
                   + b_bar[:, None] * nxt, 0)
    P = s["P"]
    c2b = -np.einsum("srn,sn->sr", P, w3b)
    w2b = w3b + np.einsum("srn,sr->sn", P, c2b)
    c1b = -np.einsum("srn,sn->sr", P, w2b)
    w1b = w2b + np.einsum("srn,sr->sn", P, c1b)
    basis_bar[:, : index + 1] += (
        -s["c2"][:, :, None] * w3b[:, None] + c2b[:, :, None] * s["w2"][:, None]
        - s["c1"][:, :, None] * w2b[:, None] + c1b[:, :, None] * s["w1"][:, None])
    q = s["q"]
    atot = np.where(s["active"], a_bar, 0) - (q * w1b).sum(1)
    w0b = w1b + q * atot[:, None]
    basis_bar[:, index] += -s["alpha"][:, None] * w1b + atot[:, None] * s["w0"]
    if index > 0:
        basis_bar[:, index - 1] += -prev_beta[:, None] * w0b
    return w0b, -(s["prev"] * w0b).sum(1)


def quadrature(alphas, betas, weight):
    """slq.py:84-121 in FP64."""
    alphas = np.asarray(alphas, np.float64)
    betas = np.asarray(betas, np.float64)
    depth, S = alphas.shape
    T = np.zeros((S, depth, depth))
    i = np.arange(depth)
    T[:, i, i] = alphas.T
    o = np.arange(depth - 1)
    T[:, o + 1, o] = betas.T
    T[:, o, o + 1] = betas.T
    vals, vecs = np.linalg.eigh(T)
    floor = max(EPS32 * depth, 1e-6)
    cl = np.maximum(vals, floor)
    lv = np.log(cl)
    lead = vecs[:, 0, :]
    value = weight * np.sum(lead * lead * lv)
    diff = vals[:, :, None] - vals[:, None, :]
    scale = np.maximum(np.abs(vals[:, :, None]), np.abs(vals[:, None, :]))
    close = np.abs(diff) <= 8 * EPS32 * np.maximum(scale, floor)
    rel = (cl[:, :, None] - cl[:, None, :]) / cl[:, None, :]
    near = np.abs(rel) < 0.5
    ld = np.where(near, np.log1p(np.where(near, rel, 0)), lv[:, :, None] - lv[:, None, :])
    mid = 0.5 * (vals[:, :, None] + vals[:, None, :])
    der = np.where(mid > floor, 1 / np.maximum(mid, floor), 0)
    div = np.where(close, der, ld / np.where(close, 1, diff))
    g = weight * (vecs @ (div * lead[:, :, None] * lead[:, None, :]) @ np.swapaxes(vecs, 1, 2))
    return value, np.diagonal(g, axis1=1, axis2=2).T.copy(), 2 * g[:, o, o + 1].T.copy()


def logdet_value_and_grad(matvec, pullback, n, key, *, nsamples, m,
                          workspace_bytes=256 * 1024**2, accept_value=None):
    """slq.py:143-213 (FP64).  matvec/pullback take and return (n, k) arrays."""
    depth = min(int(m), int(n))
    width = workspace_layout(n, nsamples, m, workspace_bytes)["batch_width"]
    probes = rademacher_probes(key, n, nsamples).astype(np.float64)
    total, grad = 0.0, None
    batches = []
    for start in range(0, nsamples, width):
        cnt = min(width, nsamples - start)
        basis = np.zeros((cnt, depth + 1, n))
        basis[:, 0] = probes[start : start + cnt] / np.sqrt(n)
        prods = np.zeros((cnt, depth, n))
        al, be = [], []
        pb = np.zeros(cnt)
        for i in range(depth):
            pr = np.asarray(matvec(basis[:, i].T), np.float64).T
            nq, a, b = step(basis, pr, pb, i)
            prods[:, i] = pr
            basis[:, i + 1] = nq
            al.append(a)
            be.append(b)
            pb = b
        A = np.stack(al)
        B = np.stack(be[:-1]) if depth > 1 else np.empty((0, cnt))
        v, abar, bsm = quadrature(A, B, n / nsamples)
        total += v
        batches.append((basis, prods, np.stack(be), abar, np.concatenate([bsm, np.zeros((1, cnt))])))
    if not np.isfinite(total):
        raise FloatingPointError("Nonfinite SLQ value.")
    if pullback is None:
        return total, None
    if accept_value is not None and not accept_value(total):
        return total, None
    for basis, prods, betas, abar, bbar in batches:
        S = basis.shape[0]
        qb = np.zeros_like(basis)
        bbar = bbar.copy()
        for i in range(depth - 1, -1, -1):
            pb = betas[i - 1] if i else np.zeros(S)
            dw, db = step_pullback(basis, prods[:, i], pb, i, qb[:, i + 1], abar[i], bbar[i], qb)
            inp, part = pullback(dw.T, basis[:, i].T)
            qb[:, i] += np.asarray(inp).T
            grad = np.asarray(part, np.float64) if grad is None else grad + part
            if i:
                bbar[i - 1] += db
    return total, grad
