"""GPU parity of the homegrown tile_kernel.h API (tile_kernel.c conventions:
beta >= 0, explicit unit-lower Y, lower T) against the CPU oracle, through
the host-pointer C-ABI; plus Hybrid_make_I / Hybrid_dgemm (main.c's
verification chain).  FP64, 1e-12 relative."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def hq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import hybrid3_amd
    hybrid3_amd.lib()
    return hybrid3_amd


def P(a):
    assert a.flags.f_contiguous or a.ndim == 1
    return a.ctypes.data_as(C.POINTER(C.c_double))


def F(a):
    return np.array(a, dtype=np.float64, order="F", copy=True)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("n,block", [(16, 16), (64, 64), (32, 8), (48, 5), (128, 128), (256, 256)])
def test_tile_dgeqrt(hq, orc, n, block):
    rng = np.random.default_rng(n + block)
    A = F(rng.uniform(-1, 1, (n, n)))
    Ao, TAUo, Yo, To = orc.tile_dgeqrt(A, block)
    A2, Y, T, TAU = F(A), F(np.zeros((n, n))), F(np.zeros((n, n))), np.zeros(n)
    hq.lib().tile_dgeqrt(n, n, block, P(A2), n, P(TAU), P(Y), P(T), n, None)
    assert hq.last_error()[0] == 0
    assert rel(A2, Ao) < TOL and rel(Y, Yo) < TOL and rel(T, To) < TOL
    assert np.abs(TAU - TAUo).max() < TOL


@pytest.mark.parametrize("n,block", [(16, 16), (64, 64), (32, 8), (30, 7), (128, 128)])
def test_tile_dtsqrt(hq, orc, n, block):
    rng = np.random.default_rng(2 * n + block)
    A1 = F(np.triu(rng.uniform(-1, 1, (n, n))))
    A2 = F(rng.uniform(-1, 1, (n, n)))
    a1o, a2o, TAUo, To = orc.tile_dtsqrt(A1, A2, block)
    g1, g2, T, TAU = F(A1), F(A2), F(np.zeros((n, n))), np.zeros(n)
    hq.lib().tile_dtsqrt(n, n, block, P(g1), n, P(g2), n, P(TAU), P(T), n, None)
    assert hq.last_error()[0] == 0
    assert rel(g1, a1o) < TOL and rel(g2, a2o) < TOL and rel(T, To) < TOL
    assert np.abs(TAU - TAUo).max() < TOL


@pytest.mark.parametrize("n", [16, 64, 128, 256])
def test_tile_dlarft_dlarfb_dssrfb(hq, orc, n):
    rng = np.random.default_rng(3 * n)
    A = F(rng.uniform(-1, 1, (n, n)))
    _, TAU, Y, _ = orc.tile_dgeqrt(A, n)
    # dlarft block = 1 (the callers' form) and a blocked form
    To = orc.tile_dlarft(Y, TAU, 1)
    T = F(np.zeros((n, n)))
    hq.lib().tile_dlarft(n, n, 1, P(Y), n, P(TAU), P(T), n, None)
    assert rel(T, To) < TOL
    _, TAUb, Yb, Tdiag = orc.tile_dgeqrt(A, 4)   # diagonal blocks of T only
    Tb_o = orc.tile_dlarft(Yb, TAUb, 4, T=F(Tdiag))
    Tb = F(Tdiag)
    hq.lib().tile_dlarft(n, n, 4, P(Yb), n, P(TAUb), P(Tb), n, None)
    assert rel(Tb, Tb_o) < TOL
    # dlarfb
    Cm = F(rng.uniform(-1, 1, (n, n)))
    Co = orc.tile_dlarfb(Cm, Y, To)
    Cg = F(Cm)
    hq.lib().tile_dlarfb(n, n, P(Cg), n, P(Y), P(To), n, None)
    assert rel(Cg, Co) < TOL
    # dssrfb with a TS reflector block
    R1 = F(np.triu(rng.uniform(-1, 1, (n, n))))
    _, V2, TAU2, _ = orc.tile_dtsqrt(R1, F(rng.uniform(-1, 1, (n, n))), n)
    T2 = orc.tile_dlarft(V2, TAU2, 1)
    B1, B2 = F(rng.uniform(-1, 1, (n, n))), F(rng.uniform(-1, 1, (n, n)))
    o1, o2 = orc.tile_dssrfb(B1, B2, V2, T2)
    g1, g2 = F(B1), F(B2)
    hq.lib().tile_dssrfb(n, n, P(g1), P(g2), n, P(V2), P(T2), n, None)
    assert rel(g1, o1) < TOL and rel(g2, o2) < TOL


@pytest.mark.parametrize("n", [128, 256])
def test_tile_dlarfb_zero_reflector_column(hq, orc, n):
    """tile_dgeqrt leaves Y's last column zero; tile_dlarfb then ignores the
    matching row and column of T whatever they hold (here a random lower T),
    on the matrix-core path too."""
    rng = np.random.default_rng(5 * n)
    _, _, Y, _ = orc.tile_dgeqrt(F(rng.uniform(-1, 1, (n, n))), n)
    assert np.all(Y[:, -1] == 0.0)
    T = F(np.tril(rng.uniform(-1, 1, (n, n))))
    Cm = F(rng.uniform(-1, 1, (n, n)))
    Co = orc.tile_dlarfb(Cm, Y, T)
    Cg = F(Cm)
    print(f"tile_dlarfb zero column n={n}", flush=True)
    hq.lib().tile_dlarfb(n, n, P(Cg), n, P(Y), P(T), n, None)
    assert rel(Cg, Co) < TOL


@pytest.mark.parametrize("row,col", [(24, 16), (40, 32), (96, 64)])
def test_tile_dssrfb_row_above_col(hq, orc, row, col):
    """tile_dssrfb with more rows than reflectors (tile_kernel.c:121-136):
    A1's rows past col come back zero, as in the reference."""
    rng = np.random.default_rng(row + col)
    Y = F(rng.uniform(-1, 1, (row, col)))
    T = F(np.tril(rng.uniform(-1, 1, (col, col))))
    B1, B2 = F(rng.uniform(-1, 1, (row, col))), F(rng.uniform(-1, 1, (row, col)))
    o1, o2 = orc.tile_dssrfb(B1, B2, Y, T)
    g1, g2 = F(B1), F(B2)
    print(f"tile_dssrfb row={row} col={col}", flush=True)
    hq.lib().tile_dssrfb(row, col, P(g1), P(g2), row, P(Y), P(T), col, None)
    assert rel(g1, o1) < TOL and rel(g2, o2) < TOL
    assert np.all(g1[col:] == 0.0)


def test_make_i_and_dgemm(hq, orc):
    MT = NT = 3
    nb = 64
    n = MT * nb
    st = hq.mat_init(n, n, nb)
    L = hq.lib()
    Id = np.full(n * n, 7.0)
    L.Hybrid_make_I(C.byref(st), P(Id), 0, 1)
    assert np.array_equal(orc.tile_to_cm(Id, MT, NT, nb), np.eye(n))
    A = orc.fill_uniform_cm(n, n, 1)
    B = orc.fill_uniform_cm(n, n, 2)
    At, Bt = orc.cm_to_tile(A, nb), orc.cm_to_tile(B, nb)
    Ct = np.zeros(n * n)
    L.Hybrid_dgemm(C.byref(st), P(At), P(Bt), P(Ct), 0, 1)
    assert hq.last_error()[0] == 0
    assert np.abs(orc.tile_to_cm(Ct, MT, NT, nb) - A.T @ B).max() < 1e-12
