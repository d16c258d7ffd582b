"""GPU: the fused Lanczos step (coh_probe_step_prep / coh_probe_step_finish
around the two re-orthogonalisation passes) against the same step written
as element-wise torch operations (slq.py:49-74), including an inactive probe
(q = 0) and a breakdown probe (product along q, so w3 vanishes)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _inputs(S, n, index, seed):
    rng = np.random.default_rng(seed)
    depth = index + 2
    basis = np.zeros((S, depth + 1, n), np.float32)
    for s in range(S):
        Q, _ = np.linalg.qr(rng.standard_normal((n, index + 1)))
        basis[s, : index + 1] = Q.T.astype(np.float32)
    product = rng.standard_normal((S, n)).astype(np.float32)
    beta = rng.uniform(0.5, 2.0, S).astype(np.float32)
    basis[0, index] = 0.0                               # inactive probe
    product[1] = 3.0 * basis[1, index]                   # breakdown probe
    if index > 0:
        product[1] += beta[1] * basis[1, index - 1]
    dev = "cuda"
    return (torch.as_tensor(basis, device=dev), torch.as_tensor(product, device=dev),
            torch.as_tensor(beta, device=dev))


@pytest.mark.parametrize("n,index", [(3000, 0), (4096 + 17, 1), (2048 * 3, 4), (700, 6)])
def test_fused_step_matches_elementwise_step(n, index):
    from coherit_amd import slq

    S = 5
    basis, product, beta = _inputs(S, n, index, seed=n + index)
    _compare(slq, basis, product, beta, index)


def test_fused_step_with_one_probe_and_a_transposed_product():
    """A one-probe batch: the product row comes from a one-column transpose
    whose row stride torch leaves at 1."""
    from coherit_amd import slq

    n, index = 2500, 2
    basis, product, beta = _inputs(3, n, index, seed=5)
    basis, beta = basis[2:].contiguous(), beta[2:].contiguous()
    col = torch.empty((n, 1), device=product.device)     # a matvec result, strides (1, 1)
    col.copy_(product[2:].T)
    product1 = col.T.contiguous()                        # (1, n), row stride 1
    assert product1.stride(0) == 1
    ref = slq._step_parts_torch(basis, product1.clone(), beta, index)
    got = slq._step_parts_device(basis, product1, beta, index)
    torch.testing.assert_close(got["w3"], ref["w3"], rtol=1e-4, atol=2e-5)
    torch.testing.assert_close(got["norm"], ref["norm"], rtol=2e-5, atol=1e-6)


def _compare(slq, basis, product, beta, index):
    ref = slq._step_parts_torch(basis, product, beta, index)
    got = slq._step_parts_device(basis, product, beta, index)
    for key in ("active", "keep"):
        assert torch.equal(got[key], ref[key]), key
    assert not bool(got["active"][0]) and not bool(got["keep"][1]) and bool(got["keep"][2])
    torch.testing.assert_close(got["w0"], ref["w0"], rtol=0, atol=0)
    for key in ("alpha", "norm"):
        torch.testing.assert_close(got[key], ref[key], rtol=2e-5, atol=1e-6)
    for key in ("w1", "c1", "w2", "c2", "w3"):
        torch.testing.assert_close(got[key], ref[key], rtol=1e-4, atol=2e-5)
    # the forward step writes the next vector into the basis
    nxt_ref, a_ref, b_ref = slq._step(basis.cpu(), product.cpu(), beta.cpu(), index)
    work = basis.clone()
    nxt, a, b = slq._step(work, product, beta, index)
    assert nxt.data_ptr() == work[:, index + 1, :].data_ptr()
    torch.testing.assert_close(a.cpu(), a_ref, rtol=2e-5, atol=1e-6)
    torch.testing.assert_close(b.cpu(), b_ref, rtol=2e-5, atol=1e-6)
    torch.testing.assert_close(nxt.cpu(), nxt_ref, rtol=1e-4, atol=2e-5)
    assert float(b[1]) == 0.0 and float(nxt[1].abs().max()) == 0.0
    assert float(a[0]) == 1.0


@pytest.mark.parametrize("n,index,S", [(3000, 0, 5), (4096 + 17, 1, 5), (2048 * 3, 4, 5),
                                       (700, 6, 5), (2500, 3, 1)])
def test_fused_step_adjoint_matches_elementwise_adjoint(n, index, S):
    """coh_probe_step_adjoint_head/tail with the re-orthogonalisation adjoint
    against the element-wise torch adjoint of the same step (run on the CPU
    in FP32), including the inactive and breakdown probes."""
    from coherit_amd import slq

    basis, product, beta = _inputs(max(S, 3), n, index, seed=7 * n + index)
    if S == 1:
        basis, product, beta = basis[2:].contiguous(), product[2:].contiguous(), beta[2:]
    rng = np.random.default_rng(index)
    depth1 = basis.shape[1]
    q_bar = torch.as_tensor(rng.standard_normal((S, depth1, n)).astype(np.float32), device="cuda")
    a_bar = torch.as_tensor(rng.standard_normal(S).astype(np.float32), device="cuda")
    b_bar = torch.as_tensor(rng.standard_normal(S).astype(np.float32), device="cuda")
    qb_ref = q_bar.cpu().clone()
    dw_ref, db_ref = slq._step_pullback(basis.cpu(), product.cpu(), beta.cpu(), index,
                                        qb_ref[:, index + 1, :], a_bar.cpu(), b_bar.cpu(), qb_ref)
    dw, db = slq._step_pullback(basis, product, beta, index, q_bar[:, index + 1, :], a_bar,
                                b_bar, q_bar)
    scale = float(dw_ref.abs().max())
    torch.testing.assert_close(dw.cpu(), dw_ref, rtol=1e-4, atol=2e-5 * scale)
    if index > 0:
        torch.testing.assert_close(db.cpu(), db_ref, rtol=1e-4, atol=1e-4)
    scale = float(qb_ref.abs().max())
    torch.testing.assert_close(q_bar.cpu(), qb_ref, rtol=1e-4, atol=2e-5 * scale)
