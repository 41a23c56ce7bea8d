"""Pins the oracle's restatement of the homegrown tile_kernel.h API
(tile_kernel.c:19-716: beta >= 0, v = [1; x/(alpha-beta)], TAU = t*y_head^2,
T stored lower = T_LAPACK^T).  tile_kernel.c cannot be compiled here (the
image has no cblas.h), so the checks are closed forms and invariants built
from explicit reflector matrices, never the oracle's own formulas."""
import numpy as np
import pytest


def refl(v, tau):
    return np.eye(len(v)) - tau * np.outer(v, v)


def test_geqr2_closed_form_2x2(orc):
    a, b, c, d = 0.3, -0.7, 0.4, 0.9
    A = np.array([[a, b], [c, d]])
    Ao, TAU, Y, T = orc.tile_dgeqrt(A, block=2)
    beta = np.hypot(a, c)
    assert abs(Ao[0, 0] - beta) < 1e-15              # beta >= 0
    assert abs(Ao[0, 1] - (a * b + c * d) / beta) < 1e-15
    assert abs(Ao[1, 1] - (-(a * d - b * c) / beta)) < 1e-15  # det(H) = -1
    yh = a - beta
    assert abs(Ao[1, 0] - c / yh) < 1e-15 and Y[0, 0] == 1.0 and Y[1, 1] == 0.0
    assert abs(TAU[0] - 2 * yh * yh / (yh * yh + c * c)) < 1e-15


@pytest.mark.parametrize("n,block", [(8, 8), (12, 12), (12, 4), (10, 3), (16, 5)])
def test_tile_dgeqrt_reflectors_reproduce_r(orc, n, block):
    rng = np.random.default_rng(n * 31 + block)
    A = rng.uniform(-1, 1, (n, n))
    Ao, TAU, Y, T = orc.tile_dgeqrt(A, block)
    R = np.triu(Ao)
    # apply H_i (explicit) in order; the last column of a square tile is not reflected
    X = A.copy()
    for i in range(n - 1):
        X = refl(Y[:, i], TAU[i]) @ X
    assert np.abs(X - R).max() < 1e-13
    assert np.all(np.diag(R)[:-1] >= 0)
    # |R| equals the LAPACK R up to row signs
    assert np.abs(np.abs(R) - np.abs(np.linalg.qr(A, mode="r"))).max() < 1e-13
    # Y: unit diagonal, zero above, last diagonal forced to 0
    assert np.all(np.diag(Y)[:-1] == 1.0) and Y[-1, -1] == 0.0
    assert np.all(np.triu(Y, 1) == 0.0)
    # each block-diagonal piece of T is the lower T^T of its own block
    for s in range(0, n, block):
        e = min(n, s + block) if s + block < n else n
        if s + block >= n:
            e = n
        Yb = Y[:, s:e]
        Tb = T[s:e, s:e]
        Q = np.eye(n)
        for i in range(s, e):
            Q = Q @ refl(Y[:, i], TAU[i] if i < n - 1 else 0.0)
        assert np.abs(Q - (np.eye(n) - Yb @ Tb.T @ Yb.T)).max() < 1e-13
        assert np.all(np.triu(Tb, 1) == 0.0)


def test_tile_dlarft_block1_gives_transposed_t(orc):
    rng = np.random.default_rng(3)
    n = 12
    A = rng.uniform(-1, 1, (n, n))
    Ao, TAU, Y, _ = orc.tile_dgeqrt(A, n)
    T = orc.tile_dlarft(Y, TAU, 1)
    Q = np.eye(n)
    for i in range(n):
        Q = Q @ refl(Y[:, i], TAU[i])
    assert np.abs(Q - (np.eye(n) - Y @ T.T @ Y.T)).max() < 1e-13
    assert np.all(np.triu(T, 1) == 0.0)


def test_tile_dlarfb_applies_qt(orc):
    rng = np.random.default_rng(4)
    n = 16
    A = rng.uniform(-1, 1, (n, n))
    _, TAU, Y, _ = orc.tile_dgeqrt(A, n)
    T = orc.tile_dlarft(Y, TAU, 1)
    Cm = rng.uniform(-1, 1, (n, n))
    out = orc.tile_dlarfb(Cm, Y, T)
    X = Cm.copy()
    for i in range(n):
        X = refl(Y[:, i], TAU[i]) @ X
    assert np.abs(out - X).max() < 1e-13


@pytest.mark.parametrize("n,block", [(8, 8), (12, 4), (10, 3), (9, 2)])
def test_tile_dtsqrt_and_dssrfb(orc, n, block):
    rng = np.random.default_rng(100 + n + block)
    A1 = np.triu(rng.uniform(-1, 1, (n, n)))
    A2 = rng.uniform(-1, 1, (n, n))
    a1, v2, TAU, T = orc.tile_dtsqrt(A1, A2, block)
    S = np.vstack([A1, A2])
    for i in range(n):
        v = np.concatenate([np.eye(n)[i], v2[:, i]])
        S = refl(v, TAU[i]) @ S
    assert np.abs(S[:n] - np.triu(a1)).max() < 1e-13
    assert np.abs(S[n:]).max() < 1e-13
    assert np.all(np.diag(a1) >= 0)
    # full-width T via tile_dlarft(block=1) on V2, then tile_dssrfb == explicit Q^T
    Tf = orc.tile_dlarft(v2, TAU, 1)
    B1 = rng.uniform(-1, 1, (n, n))
    B2 = rng.uniform(-1, 1, (n, n))
    o1, o2 = orc.tile_dssrfb(B1, B2, v2, Tf)
    X = np.vstack([B1, B2])
    for i in range(n):
        v = np.concatenate([np.eye(n)[i], v2[:, i]])
        X = refl(v, TAU[i]) @ X
    assert np.abs(o1 - X[:n]).max() < 1e-13 and np.abs(o2 - X[n:]).max() < 1e-13


def test_tile_dlarft_blocked_keeps_reference_loop_bound(orc):
    """tile_kernel.c:401-436: 2 <= block < col combines diagonal blocks; the
    loop i < col-block computes the off-diagonal rows of the middle blocks
    only.  With col = 3*block, rows block..2*block-1 are produced (and equal
    the true T^T rows), rows 2*block.. keep what was there."""
    rng = np.random.default_rng(5)
    n, b = 12, 4
    A = rng.uniform(-1, 1, (n, n))
    Ao, TAU, Y, Tdiag = orc.tile_dgeqrt(A, b)   # diagonal blocks of T only
    Tfull = orc.tile_dlarft(Y, TAU, 1)
    T = orc.tile_dlarft(Y, TAU, b, T=Tdiag.copy())
    assert np.abs(T[b:2 * b, :b] - Tfull[b:2 * b, :b]).max() < 1e-13
    assert np.all(T[2 * b:, :2 * b] == Tdiag[2 * b:, :2 * b])
    # block >= col computes nothing
    T0 = orc.tile_dlarft(Y, TAU, n, T=Tdiag.copy())
    assert np.all(T0 == Tdiag)


def test_tile_dssrfb_rows_past_col_are_zeroed(orc):
    """tile_kernel.c:121-136,313-363 with row > col: the over-block copy of A1
    has `row` rows, but only its first col rows receive Y^T A2 and the T
    product, so after A1 -= that copy, A1's rows col..row-1 are zero; the
    first col rows and A2 follow the block reflector (closed form)."""
    rng = np.random.default_rng(17)
    row, col = 12, 8
    Y = rng.uniform(-1, 1, (row, col))
    T = np.tril(rng.uniform(-1, 1, (col, col)))
    A1 = rng.uniform(-1, 1, (row, col))
    A2 = rng.uniform(-1, 1, (row, col))
    o1, o2 = orc.tile_dssrfb(A1, A2, Y, T)
    W = T @ (A1[:col] + Y.T @ A2)
    assert np.abs(o1[:col] - (A1[:col] - W)).max() < 1e-13
    assert np.all(o1[col:] == 0.0)
    assert np.abs(o2 - (A2 - Y @ W)).max() < 1e-13
