"""Least squares min ||A x - b|| on one GPU with the tile QR engine, and
iterative refinement to FP64 accuracy (no reference counterpart).

The factorisation is either the plain FP64 tile QR or the mixed one (FP64
panels, FP32 LARFB / SSRFB updates).  The solution is refined on the
augmented system [I A; A^T 0] [r; x] = [b; 0] with FP64 residuals
(Bjorck's refinement): per step

    f = b - r - A x,  g = -A^T r                      (FP64, original A)
    d = Q^T f,  h = R^-T g,  dx = R^-1 (d[:n] - h)    (the computed factors)
    dr = f - A dx;  x += dx;  r += dr

The fixed point satisfies the FP64 normal equations A^T (b - A x) = 0 up to
FP64 rounding whatever the accuracy of Q and R; the factors only set the
convergence rate, about kappa * u of the factorisation, so the mixed
factorisation converges while kappa * u_fp32 < 1.

A is given column-major (m x n, m >= n, FP64, on the device).  Q^T is applied
with the library's tile ormqr (the reflectors and T of the factorisation);
the triangular solves with R and the products with A use torch (rocBLAS):
they are the refinement's residual arithmetic, not the factorisation.
"""
from __future__ import annotations

import time

from . import cm_to_tile_dev, sync, tile_ormqr_dev, tile_qr_dev, tile_qr_mixed_dev, tile_to_cm_dev

__all__ = ["TileQR", "lstsq", "backward_indicator", "gaussian_orthonormal", "svd_matrix"]


class TileQR:
    """Tile QR of a column-major m x n device matrix (zero-padded to tiles)."""

    def __init__(self, A, nb=256, ib=32, mixed=False):
        import torch
        m, n = A.shape
        self.m, self.n, self.nb, self.ib, self.mixed = m, n, nb, ib, mixed
        self.MT, self.NT = (m + nb - 1) // nb, (n + nb - 1) // nb
        MT, NT = self.MT, self.NT
        self.RY = torch.empty(MT * NT * nb * nb, dtype=torch.float64, device=A.device)
        self.TAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device=A.device)
        self.T = torch.zeros(MT * NT * ib * nb, dtype=torch.float64, device=A.device)
        Ac = A if _is_colmajor(A) else A.t().contiguous().t()
        cm_to_tile_dev(nb, [m], [n], [_addr(Ac)], [m], [self.RY])
        t0 = time.perf_counter()
        (tile_qr_mixed_dev if mixed else tile_qr_dev)(MT, NT, nb, ib, self.RY, self.TAU, self.T)
        sync()
        torch.cuda.synchronize()
        self.factor_seconds = time.perf_counter() - t0
        # R (n x n upper) column-major for the triangular solves
        full = torch.empty(MT * nb * NT * nb, dtype=torch.float64, device=A.device)
        tile_to_cm_dev(nb, [MT * nb], [NT * nb], [self.RY], [full], [MT * nb])
        sync()
        self.R = torch.triu(full.view(NT * nb, MT * nb).t()[:n, :n]).contiguous()

    def qt(self, Y):
        """Q^T Y for a column-major m x k device block (k <= nb)."""
        import torch
        m, k = Y.shape
        X = torch.zeros(self.MT * self.nb * self.nb, dtype=torch.float64, device=Y.device)
        Yc = Y.t().contiguous()  # column-major storage of Y (k x m row-major)
        cm_to_tile_dev(self.nb, [m], [k], [_addr(Yc)], [m], [X])
        tile_ormqr_dev(self.MT, self.NT, self.nb, self.ib, self.RY, self.TAU, X, 1, T=self.T)
        out = torch.empty(k * m, dtype=torch.float64, device=Y.device)
        tile_to_cm_dev(self.nb, [m], [k], [X], [out], [m])
        sync()
        return out.view(k, m).t()


def _is_colmajor(A):
    return A.dim() == 2 and A.stride(0) == 1 and A.stride(1) == A.shape[0]


def _addr(t):
    return t.data_ptr()


def lstsq(A, b, nb=256, ib=32, mixed=False, max_iter=12, tol=None, factor=None):
    """x minimising ||A x - b||_2 (A m x n, b m x k, FP64 torch device tensors).

    Returns (x, info): info holds the factorisation and total times, the
    refinement steps taken, and per step ||dx|| / ||x|| and the normwise
    backward-error indicator ||A^T r|| / (||A||_F ||r||)."""
    import torch
    m, n = A.shape
    if b.dim() == 1:
        b = b[:, None]
    t0 = time.perf_counter()
    F = factor if factor is not None else TileQR(A, nb, ib, mixed)
    R = F.R

    def rsolve(c):
        return torch.linalg.solve_triangular(R, c, upper=True)

    def rtsolve(c):
        return torch.linalg.solve_triangular(R.t(), c, upper=False)

    x = rsolve(F.qt(b)[:n])
    r = b - A @ x
    steps, hist = 0, []
    tol = 2.0 ** -52 if tol is None else tol
    for it in range(max_iter):
        f = b - r - A @ x
        g = -(A.t() @ r)
        dx = rsolve(F.qt(f)[:n] - rtsolve(g))
        dr = f - A @ dx
        x = x + dx
        r = r + dr
        steps += 1
        rel = (torch.linalg.norm(dx) / torch.linalg.norm(x)).item()
        hist.append({"dx_rel": rel, "backward": backward_indicator(A, b, x)})
        # converged, or stagnated at the rounding level of the problem
        if rel <= tol or (it > 0 and rel > 0.5 * hist[-2]["dx_rel"]):
            break
    torch.cuda.synchronize()
    info = {"factor_seconds": F.factor_seconds, "total_seconds": time.perf_counter() - t0,
            "steps": steps, "history": hist, "mixed": mixed}
    return x, info


def backward_indicator(A, b, x):
    """||A^T r|| / (||A||_F ||r||), r = b - A x, in FP64: zero at the exact
    least-squares solution; for a small residual it is limited by the
    rounding of r, so it is compared with the value of a reference solver."""
    import torch
    if b.dim() == 1:
        b = b[:, None]
    if x.dim() == 1:
        x = x[:, None]
    r = b - A @ x
    return (torch.linalg.norm(A.t() @ r) / (torch.linalg.norm(A) * torch.linalg.norm(r))).item()


def gaussian_orthonormal(m, n, seed, nb=256, ib=32):
    """m x n matrix with orthonormal columns spanning a seeded Gaussian
    matrix G: Q = G R^-1 with R from the tile QR of G, orthonormalised once
    more the same way (two passes leave ||Q^T Q - I|| at rounding level for
    a Gaussian G).  Column-major, on the device."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    Q = torch.randn(n, m, generator=g, dtype=torch.float64, device="cuda").t()  # column-major m x n
    for _ in range(2):
        R = TileQR(Q, nb, ib).R
        Q = torch.linalg.solve_triangular(R, Q, upper=True, left=False)
        Q = Q.t().contiguous().t()
    return Q


def svd_matrix(m, n, kappa, seed, nb=256, ib=32):
    """A = U diag(sigma) V^T with U (m x n), V (n x n) orthonormal from
    seeded Gaussian QR and sigma log-spaced from 1 to 1/kappa."""
    import torch
    U = gaussian_orthonormal(m, n, seed, nb, ib)
    V = gaussian_orthonormal(n, n, seed + 1, nb, ib)
    sigma = torch.logspace(0, -float(torch.log10(torch.tensor(kappa))), n, dtype=torch.float64, device="cuda")
    A = (U * sigma[None, :]) @ V.t().contiguous()
    return A.t().contiguous().t()  # column-major
