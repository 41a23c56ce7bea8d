"""Least squares min ||A x - b|| on one GPU with the tile QR engine, refined
to FP64 accuracy (no reference counterpart).

The factorisation A = QR is the plain FP64 tile QR, or a mixed one: FP64
panels with the trailing updates in FP32 ("fp32") or in FP16 hi + lo splits
("split").  Two refinements with FP64 residuals are offered:

* ``method="cgls"`` (default): CGLS on min ||A R^-1 y - b||, x = R^-1 y,
  started from the semi-normal solution x0 = R^-1 R^-T A^T b.  Every step
  takes two products with A (the residual is recomputed from x, not
  updated), one with A^T and two with R^-1 (formed once by a triangular
  solve); Q is not needed.  With R from a factorisation of relative accuracy u_f,
  A R^-1 has condition number about 1 + kappa(A) u_f, so the iteration
  converges in a few steps while kappa u_f << 1, and still converges (more
  slowly) beyond it, where classical refinement stalls.
* ``method="refine"``: Bjorck's refinement of the augmented system
  [I A; A^T 0][r; x] = [b; 0]: f = b - r - A x, g = -A^T r,
  dx = R^-1 (Q^T f - R^-T g), dr = f - A dx.  Needs Q^T (the library's tile
  ormqr) per step; converges while kappa u_f < 1.

Both stop when the iteration makes no further progress or after
``max_iter`` steps.  ``converged`` says whether the result reached the
rounding floor of its FP64 residual: a normwise backward-error indicator
||A^T r|| / (||A||_F ||r||) within eps (||b|| + ||A||_F ||x||) / ||r|| (for
the refinement, also a relative update at FP64 rounding).  A result that did
not converge carries the error it reports.
A is column-major (m x n, m >= n, FP64, on the device).  The triangular
solves and the products with A are torch (rocBLAS): they are the
refinement's residual arithmetic, not the factorisation.
"""
from __future__ import annotations

import time

from . import (batched_prepare, batched_tile_qr_ex_dev, cm_to_tile_dev, sync, tile_ormqr_cols_dev, tile_qr_dev,
               tile_qr_mixed_dev, tile_qr_prepare, tile_qr_split_dev, tile_to_cm_dev)

__all__ = ["TileQR", "lstsq", "batched_lstsq", "backward_indicator", "gaussian_orthonormal", "svd_matrix",
           "PRECISIONS"]

PRECISIONS = {"fp64": 0, "fp32": 1, "split": 2}
EPS = 2.0 ** -52
PROGRESS = 0.8        # a CGLS step whose eta and update shrink by less than this made no progress
SMALL_UPDATE = 1e-6   # relative update below which such a step counts toward the stop at once
TRANSIENT = 8         # steps after which every step without progress counts


def _prec(prec, mixed):
    if prec is None:
        prec = "fp32" if mixed else "fp64"
    if prec not in PRECISIONS:
        raise ValueError(f"prec must be one of {sorted(PRECISIONS)}")
    return prec


class TileQR:
    """Tile QR of a column-major m x n device matrix (zero-padded to tiles)."""

    def __init__(self, A, nb=256, ib=32, mixed=False, prec=None):
        import torch
        m, n = A.shape
        self.prec = _prec(prec, mixed)
        self.m, self.n, self.nb, self.ib = m, n, nb, ib
        self.MT, self.NT = (m + nb - 1) // nb, (n + nb - 1) // nb
        MT, NT = self.MT, self.NT
        self.RY = torch.empty(MT * NT * nb * nb, dtype=torch.float64, device=A.device)
        self.TAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device=A.device)
        self.T = torch.zeros(MT * NT * ib * nb, dtype=torch.float64, device=A.device)
        Ac = A if _is_colmajor(A) else A.t().contiguous().t()
        cm_to_tile_dev(nb, [m], [n], [_addr(Ac)], [m], [self.RY])
        fn = {"fp64": tile_qr_dev, "fp32": tile_qr_mixed_dev, "split": tile_qr_split_dev}[self.prec]
        tile_qr_prepare(MT, NT, nb, ib, PRECISIONS[self.prec])   # schedule and scratch (cached)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn(MT, NT, nb, ib, self.RY, self.TAU, self.T)
        sync()
        torch.cuda.synchronize()
        self.factor_seconds = time.perf_counter() - t0
        # R (n x n upper) from the top NT tile rows: tile (i, j) holds
        # entries [i nb + row, j nb + col] at [j][i][col][row]
        top = self.RY.view(NT, MT, nb, nb)[:, :NT]
        self.R = torch.triu(top.permute(1, 3, 0, 2).reshape(NT * nb, NT * nb)[:n, :n]).contiguous()

    def qt(self, Y):
        """Q^T Y for a column-major m x k device block (k <= nb): the k
        columns in one zero-padded tile column, applied by the narrow-C
        entry (tasks only on the slabs that hold columns; 16-column tasks
        below 256 columns)."""
        import torch
        m, k = Y.shape
        X = torch.zeros(self.MT * self.nb * self.nb, dtype=torch.float64, device=Y.device)
        Yc = Y.t().contiguous()  # column-major storage of Y (k x m row-major)
        cm_to_tile_dev(self.nb, [m], [k], [_addr(Yc)], [m], [X])
        tile_ormqr_cols_dev(self.MT, self.NT, self.nb, self.ib, self.RY, self.TAU, X, 1, k, T=self.T)
        out = torch.empty(k * m, dtype=torch.float64, device=Y.device)
        tile_to_cm_dev(self.nb, [m], [k], [X], [out], [m])
        sync()
        return out.view(k, m).t()


def _is_colmajor(A):
    return A.dim() == 2 and A.stride(0) == 1 and A.stride(1) == A.shape[0]


def _addr(t):
    return t.data_ptr()


class _Cgls:
    """R-preconditioned CGLS on every column of b (device tensors with the
    right-hand sides in the last dimension; normA: ||A||_F, a scalar or one
    per problem broadcasting against (.., 1, k)).

    One pass over A per step forms both A x (the true residual of the
    current iterate, so rounding in the recursive residual cannot
    accumulate over steps) and q = A t (the step direction), a second pass
    A^T r.  Each step measures every column's normwise backward-error
    indicator eta = ||A^T r|| / (||A||_F ||r||) and keeps the iterate with
    the smallest eta.  A step makes progress if some column's eta falls
    below PROGRESS times its smallest value so far, or its relative update
    ||dx|| / ||x|| below PROGRESS times the smallest so far (eta need not
    fall monotonically while x converges); the iteration stops after
    `patience` steps without progress (as LAPACK's refinement stops once the
    backward error stops falling by a fixed factor) -- counted once the
    update is below SMALL_UPDATE or after TRANSIENT steps -- or after max_iter
    steps.  Converged means every column's best eta is within the rounding
    floor of its FP64 residual, eps (||b|| + ||A||_F ||x||) / ||r||.

    step() enqueues one step and returns a 4-vector on the device (progress
    anywhere, every update small, largest eta, largest relative update);
    decide() takes it on the host and says whether to stop, so several
    problems can step together with one host synchronisation per step."""

    def __init__(self, Amul, ATmul, rsolve, rtsolve, b, max_iter, normA, patience=2):
        self.Amul, self.ATmul, self.rsolve, self.rtsolve = Amul, ATmul, rsolve, rtsolve
        self.b, self.normA, self.max_iter, self.patience = b, normA, max_iter, patience
        self.k = b.shape[-1]
        x = rsolve(rtsolve(ATmul(b)))              # semi-normal start
        r = b - Amul(x)
        atr = ATmul(r)
        s = rtsolve(atr)
        self.x, self.p = x, s
        self.gam = (s * s).sum(dim=-2, keepdim=True)
        eta = self._norms(atr) / (normA * self._norms(r))
        self.best_eta, self.best_x, self.best_dx = eta, x, None
        self.hist, self.steps, self.since, self.done = [], 0, 0, False

    @staticmethod
    def _norms(v):
        import torch
        return torch.linalg.vector_norm(v, dim=-2, keepdim=True)

    def step(self):
        import torch
        k, norms = self.k, self._norms
        t = self.rsolve(self.p)
        AXT = self.Amul(torch.cat([self.x, t], dim=-1))
        r = self.b - AXT[..., :k]
        q = AXT[..., k:]
        qq = (q * q).sum(dim=-2, keepdim=True)
        alpha = torch.where(qq > 0, self.gam / torch.where(qq > 0, qq, torch.ones_like(qq)), torch.zeros_like(qq))
        dx = alpha * t
        self.x = self.x + dx
        r = r - alpha * q
        atr = self.ATmul(r)
        s = self.rtsolve(atr)
        gn = (s * s).sum(dim=-2, keepdim=True)
        gam = self.gam
        beta = torch.where(gam > 0, gn / torch.where(gam > 0, gam, torch.ones_like(gam)), torch.zeros_like(gam))
        self.p = s + beta * self.p
        self.gam = gn
        self.steps += 1
        eta = norms(atr) / (self.normA * norms(r))
        nx = norms(self.x)
        rel = norms(dx) / torch.where(nx > 0, nx, torch.ones_like(eta))
        progress = eta < PROGRESS * self.best_eta
        if self.best_dx is not None:
            progress = progress | (rel < PROGRESS * self.best_dx)
        self.best_dx = rel if self.best_dx is None else torch.minimum(rel, self.best_dx)
        better = eta < self.best_eta
        self.best_x = torch.where(better, self.x, self.best_x)
        self.best_eta = torch.where(better, eta, self.best_eta)
        return torch.stack([progress.any().double(), (rel < SMALL_UPDATE).all().double(), eta.max(), rel.max()])

    def decide(self, flags):
        """flags: the host values of step()'s vector; True when finished."""
        progress, small, eta_max, rel_max = flags
        self.hist.append({"eta": float(eta_max), "dx_rel": float(rel_max)})
        # a step without progress counts toward the stop once the update is
        # small (at the rounding floor) or the start-up transient is over: with
        # factors far less accurate than 1/kappa, eta oscillates over the first
        # steps while x still moves by O(1)
        if progress:
            self.since = 0
        elif small or self.steps > TRANSIENT:
            self.since += 1
        self.done = self.since >= self.patience or self.steps >= self.max_iter
        return self.done

    def finish(self):
        """(best x, steps, history, converged flag on the device)."""
        norms = self._norms
        r = self.b - self.Amul(self.best_x)
        floor = EPS * (norms(self.b) + self.normA * norms(self.best_x)) / norms(r)
        return self.best_x, self.steps, self.hist, (self.best_eta <= floor).all()


def _cgls_many(solvers):
    """Step every solver together until each has stopped: one host
    synchronisation per step for all of them."""
    import torch
    live = [c for c in solvers if c.max_iter > 0]
    while live:
        flags = torch.stack([c.step() for c in live]).cpu().numpy()
        live = [c for c, f in zip(live, flags) if not c.decide(f)]
    out = [c.finish() for c in solvers]
    conv = torch.stack([o[3] for o in out]).cpu().numpy() if out else []
    return [(x, st, h, bool(cv)) for (x, st, h, _), cv in zip(out, conv)]


def _cgls(Amul, ATmul, rsolve, rtsolve, b, max_iter, normA, patience=2):
    """One problem (see _Cgls).  Returns (x, steps, history, converged)."""
    return _cgls_many([_Cgls(Amul, ATmul, rsolve, rtsolve, b, max_iter, normA, patience)])[0]


def _tri_inv(R, leaf=1024):
    """Inverse of an upper triangular R by recursive halving:
    [A B; 0 C]^-1 = [A^-1, -A^-1 B C^-1; 0, C^-1], the off-diagonal blocks
    as two GEMMs; leaves by a triangular solve.  (One triangular solve
    against the identity is a few times slower on rocBLAS.)  Works on
    batches (..., n, n)."""
    import torch
    n = R.shape[-1]
    if n <= leaf:
        eye = torch.eye(n, dtype=R.dtype, device=R.device).expand(R.shape)
        return torch.linalg.solve_triangular(R, eye, upper=True)
    h = (n // 2 + 255) // 256 * 256 if n > 512 else n // 2
    Ai, Ci = _tri_inv(R[..., :h, :h], leaf), _tri_inv(R[..., h:, h:], leaf)
    out = torch.zeros_like(R)
    out[..., :h, :h] = Ai
    out[..., h:, h:] = Ci
    out[..., :h, h:] = -(Ai @ R[..., :h, h:]) @ Ci
    return out


def lstsq(A, b, nb=256, ib=32, mixed=False, max_iter=None, tol=None, factor=None, prec=None, method="cgls",
          fallback=True):
    """x minimising ||A x - b||_2 (A m x n, b m x k, FP64 torch device tensors).

    Returns (x, info): info holds the factorisation and total times, the
    refinement steps taken, per step ||dx|| / ||x||, whether the iteration
    reached FP64 rounding (``converged``), and the final normwise
    backward-error indicator ||A^T r|| / (||A||_F ||r||).  With
    ``fallback``, a reduced-precision solve that does not converge (kappa
    beyond about 1/u of its factorisation) is repeated on the FP64
    factorisation (info["fallback"]; the times include both)."""
    import torch
    m, n = A.shape
    if b.dim() == 1:
        b = b[:, None]
    tol = 4.0 * EPS if tol is None else tol
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    F = factor if factor is not None else TileQR(A, nb, ib, mixed, prec)
    R = F.R

    def rsolve(c):
        return torch.linalg.solve_triangular(R, c, upper=True)

    def rtsolve(c):
        return torch.linalg.solve_triangular(R.t(), c, upper=False)

    if method == "cgls":
        # the preconditioner as an explicit inverse (one BLAS-3 triangular
        # solve): every step is then matrix products only.  It need not be
        # exact -- x is updated directly and its residual formed with A.
        Ri = _tri_inv(R)
        RiT = Ri.t().contiguous()
        x, steps, hist, conv = _cgls(lambda v: A @ v, lambda v: A.t() @ v, lambda c: Ri @ c, lambda c: RiT @ c, b,
                                     100 if max_iter is None else max_iter, torch.linalg.norm(A))
    elif method == "refine":
        max_iter = 30 if max_iter is None else max_iter
        x = rsolve(F.qt(b)[:n])
        r = b - A @ x
        steps, hist, conv = 0, [], False
        stall = 0
        for it in range(max_iter):
            f = b - r - A @ x
            g = -(A.t() @ r)
            dx = rsolve(F.qt(f)[:n] - rtsolve(g))
            dr = f - A @ dx
            x = x + dx
            r = r + dr
            steps += 1
            rel = (torch.linalg.norm(dx) / torch.linalg.norm(x)).item()
            hist.append({"dx_rel": rel})
            if rel <= tol:
                conv = True
                break
            # no progress for three steps: at the rounding floor, or stalled
            # (kappa u_f >= 1), which `converged` tells apart
            stall = stall + 1 if it > 0 and rel >= 0.9 * hist[-2]["dx_rel"] else 0
            if stall >= 3:
                # stalled: at the rounding floor if the backward indicator
                # is within the floor of the FP64 residual
                rr = torch.linalg.norm(b - A @ x)
                floor = EPS * (torch.linalg.norm(b) + torch.linalg.norm(A) * torch.linalg.norm(x)) / rr
                conv = backward_indicator(A, b, x) <= floor.item()
                break
    else:
        raise ValueError("method must be 'cgls' or 'refine'")
    torch.cuda.synchronize()
    info = {"factor_seconds": F.factor_seconds, "total_seconds": time.perf_counter() - t0,
            "steps": steps, "history": hist, "prec": F.prec, "mixed": F.prec != "fp64", "method": method,
            "converged": conv, "fallback": False}
    if fallback and not conv and F.prec != "fp64" and factor is None:
        del F, R
        x, i2 = lstsq(A, b, nb=nb, ib=ib, max_iter=max_iter, tol=tol, prec="fp64", method=method, fallback=False)
        info.update({"total_seconds": info["total_seconds"] + i2["total_seconds"], "converged": i2["converged"],
                     "fallback": True, "fallback_steps": i2["steps"],
                     "factor_seconds": info["factor_seconds"] + i2["factor_seconds"]})
    info["backward"] = backward_indicator(A, b, x)
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


# ---------------------------------------------------------------- batches
def batched_lstsq(groups, nb=256, ib=32, prec="fp64", max_iter=100):
    """Many least-squares problems: groups maps a shape (m, n) to (A, b) with
    A of shape (B, n, m) (each A[i] the column-major storage of an m x n
    matrix) and b of shape (B, m, k).  Every problem of every group is
    factored in ONE batched launch, then refined by R-preconditioned CGLS
    group by group (batched products and triangular solves).

    Returns ({shape: x (B, n, k)}, info)."""
    import torch
    pv = PRECISIONS[prec]
    shapes = list(groups)
    MT, NT, NC, RY, TAU, cols, owners = [], [], [], [], [], [], []
    for sh in shapes:
        m, n = sh
        A, _ = groups[sh]
        B = A.shape[0]
        mt, nt = (m + nb - 1) // nb, (n + nb - 1) // nb
        ry = torch.empty(B, mt * nt * nb * nb, dtype=torch.float64, device="cuda")
        ta = torch.zeros(B, mt * nt * nb, dtype=torch.float64, device="cuda")
        cm_to_tile_dev(nb, [m] * B, [n] * B, [A[i].data_ptr() for i in range(B)], [m] * B,
                       [ry[i].data_ptr() for i in range(B)])
        RY.append(ry)
        TAU.append(ta)
        for i in range(B):
            MT.append(mt)
            NT.append(nt)
            NC.append(n)
            cols.append(ry[i].data_ptr())
            owners.append(ta[i].data_ptr())
    batched_prepare(pv, nb, ib, MT, NT)   # schedule (cached)
    sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batched_tile_qr_ex_dev(pv, 0, nb, ib, MT, NT, cols, owners, ncol=NC)   # padding columns skipped
    sync()
    torch.cuda.synchronize()
    t_fact = time.perf_counter() - t0
    # every shape group's CGLS steps together (one host synchronisation per
    # step for all groups)
    solvers = []
    for g, sh in enumerate(shapes):
        m, n = sh
        A, b = groups[sh]
        B = A.shape[0]
        mt, nt = (m + nb - 1) // nb, (n + nb - 1) // nb
        top = RY[g].view(B, nt, mt, nb, nb)[:, :, :nt]     # [b][tj][ti][col][row]
        R = torch.triu(top.permute(0, 2, 4, 1, 3).reshape(B, nt * nb, nt * nb)[:, :n, :n]).contiguous()
        At = A.transpose(1, 2)      # (B, m, n) view of the column-major storage
        Ri = _tri_inv(R)            # preconditioner, explicit
        RiT = Ri.transpose(1, 2).contiguous()
        normA = torch.linalg.matrix_norm(A).view(B, 1, 1)
        solvers.append(_Cgls(lambda v, At=At: torch.bmm(At, v), lambda v, A=A: torch.bmm(A, v),
                             lambda c, Ri=Ri: torch.bmm(Ri, c), lambda c, RiT=RiT: torch.bmm(RiT, c), b, max_iter,
                             normA))
    out, steps, conv = {}, {}, True
    for sh, (x, st, _, cv) in zip(shapes, _cgls_many(solvers)):
        out[sh] = x
        steps[sh] = st
        conv = conv and cv
    torch.cuda.synchronize()
    info = {"factor_seconds": t_fact, "total_seconds": time.perf_counter() - t0, "steps": steps,
            "converged": conv, "prec": prec}
    return out, info
