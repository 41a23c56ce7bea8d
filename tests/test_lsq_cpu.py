"""The refinement arithmetic of hybrid3_amd.lsq on the CPU (no GPU): the
R-preconditioned CGLS, fed with the R of a deliberately inaccurate
factorisation (the QR of A computed in FP32, then promoted), reaches the
FP64 least-squares solution that LAPACK (numpy.linalg.lstsq) gives, on
several right-hand sides at once; with the exact R it converges within a few steps (the rest is the
stagnation test at the rounding floor).
This pins the iteration independently of the tile QR kernels."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def problem(m, n, kappa, k, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.logspace(0, -np.log10(kappa), n)
    A = (U * s) @ V.T
    X = rng.standard_normal((n, k))
    B = A @ X + 1e-8 * rng.standard_normal((m, k))
    return A, B


@pytest.mark.parametrize("kappa", [1e2, 1e6])
def test_cgls_with_fp32_factor_reaches_lapack(kappa):
    from hybrid3_amd.lsq import _cgls
    A, B = problem(300, 60, kappa, 4, 7)
    R32 = np.linalg.qr(A.astype(np.float32), mode="r").astype(np.float64)
    At, Bt, Rt = (torch.from_numpy(np.ascontiguousarray(a)) for a in (A, B, R32))

    def rs(c):
        return torch.linalg.solve_triangular(Rt, c, upper=True)

    def rts(c):
        return torch.linalg.solve_triangular(Rt.t(), c, upper=False)

    nA = torch.linalg.norm(At)
    x, steps, hist, conv = _cgls(lambda v: At @ v, lambda v: At.t() @ v, rs, rts, Bt, 200, nA)
    assert conv, hist[-3:]
    xl = np.linalg.lstsq(A, B, rcond=None)[0]
    err = np.abs(x.numpy() - xl).max() / np.abs(xl).max()
    assert err < 1e-13 * kappa, err
    # a factor accurate to FP32 needs more than one step; an exact one does not
    assert steps > 1
    Rq = np.linalg.qr(A, mode="r")
    Rq_t = torch.from_numpy(Rq)
    x2, steps2, _, conv2 = _cgls(lambda v: At @ v, lambda v: At.t() @ v,
                                 lambda c: torch.linalg.solve_triangular(Rq_t, c, upper=True),
                                 lambda c: torch.linalg.solve_triangular(Rq_t.t(), c, upper=False),
                                 Bt, 200, nA)
    assert conv2 and steps2 <= 8
    assert np.abs(x2.numpy() - xl).max() / np.abs(xl).max() < 1e-13 * kappa
