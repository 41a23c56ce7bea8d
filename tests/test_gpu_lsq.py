"""Least squares with iterative refinement on the tile QR engine (no
reference counterpart).  A = U diag(sigma) V^T (U, V orthonormal from seeded
Gaussian matrices, sigma log-spaced to 1/kappa), b = A x_true + 1e-8 noise.

Pinned by LAPACK's least-squares solution (numpy.linalg.lstsq, FP64): the
FP64 path and the mixed path (FP32 trailing updates) after refinement both
agree with it to FP64 accuracy scaled by kappa, and satisfy the normal
equations as well as LAPACK's solution does (||A^T r|| / (||A||_F ||r||),
which for a small residual is limited by the rounding of r itself)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lsq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import hybrid3_amd
    hybrid3_amd.lib()
    from hybrid3_amd import lsq
    return lsq


@pytest.mark.parametrize("kappa", [1e2, 1e6])
def test_refined_lstsq_matches_lapack(lsq, kappa):
    import torch
    m, n, nb, ib = 1536, 384, 128, 32
    print(f"lstsq kappa={kappa:g}", flush=True)
    A = lsq.svd_matrix(m, n, kappa, seed=5, nb=nb, ib=ib)
    g = torch.Generator(device="cuda").manual_seed(6)
    xt = torch.randn(n, 1, generator=g, dtype=torch.float64, device="cuda")
    b = A @ xt + 1e-8 * torch.randn(m, 1, generator=g, dtype=torch.float64, device="cuda")
    x_ref = np.linalg.lstsq(A.cpu().numpy(), b.cpu().numpy(), rcond=None)[0]
    scale = np.linalg.norm(x_ref)
    # the normal-equations indicator of LAPACK's own solution, same formula
    eta_ref = lsq.backward_indicator(A, b, torch.from_numpy(x_ref).cuda())
    for mixed in (False, True):
        x, info = lsq.lstsq(A, b, nb=nb, ib=ib, mixed=mixed, max_iter=20)
        err = np.linalg.norm(x.cpu().numpy() - x_ref) / scale
        eta = lsq.backward_indicator(A, b, x)
        print(mixed, info["steps"], err, eta, eta_ref, flush=True)
        assert err < 1e-14 * kappa * 10, (mixed, err)
        assert eta < 10 * eta_ref, (mixed, eta, eta_ref)
