"""GPU: the preconditioner applications' element-wise tails
(coh_core_correction, coh_scale_add) give exactly the values of the
reference's torch expressions (precond.py:42-58, 61-84) on the same
device tensors."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rand(shape, dtype, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.randn(shape, dtype=dtype, device="cuda", generator=g)


@pytest.mark.parametrize("r,k", [(508, 16), (40, 1), (3, 7)])
def test_tail_kernels_equal_the_elementwise_expressions(r, k):
    from coherit_amd import precond

    d = 0.3719346781234567
    z = _rand((k, r), torch.float64, 1).T              # column-major, as cholesky_solve returns
    rhs = _rand((r, k), torch.float64, 2)
    assert k == 1 or not z.is_contiguous()
    assert torch.equal(precond._core_correction32(z, rhs, d), (z - d * rhs).to(torch.float32))
    n = 70001
    V = _rand((n, k), torch.float32, 3)
    G = _rand((n, k), torch.float32, 4)
    assert torch.equal(precond._scaled_sum(d, V, G), d * V + G)
    v = V[:, 0].contiguous()
    assert torch.equal(precond._scaled_sum(d, v[:, None], G[:, :1].contiguous()),
                       d * v[:, None] + G[:, :1])


def test_preconditioner_applications_are_unchanged():
    from coherit_amd import precond

    n, r, k = 30000, 120, 12
    U, _ = torch.linalg.qr(_rand((n, r), torch.float64, 5))
    U = U.to(torch.float32).contiguous()
    core = _rand((r, r), torch.float64, 6)
    middle = core @ core.T / r + 0.8 * torch.eye(r, dtype=torch.float64, device="cuda")
    chol = torch.linalg.cholesky(middle)
    ev, vecs = torch.linalg.eigh(middle)
    ev = torch.clamp(ev, min=1e-8)
    d = 0.8
    V = _rand((n, k), torch.float32, 7)
    # M^{-1} V as the reference's expressions
    rhs64 = precond.tall_t_matmul(U, V).double()
    z = torch.cholesky_solve(rhs64, chol, upper=False)
    ref = (1.0 / d) * V + U @ (z - (1.0 / d) * rhs64).to(torch.float32)
    got = precond._apply_scalar_diag(V, 1.0 / d, chol, U)
    assert torch.equal(got, ref)
    # M^{-1/2} V
    ds = float(d ** -0.5)
    rhs = precond.tall_t_matmul(U, V).double()
    zz = vecs @ ((vecs.T @ rhs) / torch.sqrt(torch.clamp(ev, min=1e-8))[:, None])
    ref = ds * V + U @ (zz - ds * rhs).to(torch.float32)
    got = precond._apply_scalar_diag_invsqrt(V, ds, ev, vecs, U)
    assert torch.equal(got, ref)
    got_vec = precond._apply_scalar_diag_invsqrt(V[:, 3].contiguous(), ds, ev, vecs, U)
    rhs1 = precond.tall_t_matmul(U, V[:, 3:4].contiguous()).double()
    z1 = vecs @ ((vecs.T @ rhs1) / torch.sqrt(torch.clamp(ev, min=1e-8))[:, None])
    ref1 = (ds * V[:, 3:4] + U @ (z1 - ds * rhs1).to(torch.float32))[:, 0]
    assert torch.equal(got_vec, ref1)
