"""Pins the oracle's tile-layout drivers: Hybrid_tile_QR (flat-tree tile QR,
Hybrid_tile_QR.c:307-1461), Hybrid_make_Q (Hybrid_make_Q.c:56-237),
Hybrid_dorgqr (Hybrid_dorgqr.c:22-192), Hybrid_dormqr (Hybrid_dormqr.c:20-171),
the TileSch-style dynamic schedule and the input generators."""
import ctypes

import numpy as np
import pytest


def r_from_tiles(A, MT, NT, nb, orc):
    return np.triu(orc.tile_to_cm(A, MT, NT, nb))[: NT * nb]


def sign_normalise(R):
    s = np.sign(np.diag(R))
    s[s == 0] = 1
    return R * s[:, None]


@pytest.mark.parametrize("MT,NT,nb,ib", [(4, 4, 8, 4), (3, 3, 16, 5), (5, 2, 8, 3), (6, 6, 4, 4)])
def test_tile_qr_r_matches_lapack_up_to_row_signs(orc, MT, NT, nb, ib):
    A = orc.fill_uniform_cm(MT * nb, NT * nb, 11)
    At = orc.cm_to_tile(A, nb)
    RY, TAU = orc.tile_qr(At, MT, NT, nb, ib)
    R = r_from_tiles(RY, MT, NT, nb, orc)
    R_np = np.linalg.qr(A, mode="r")
    assert np.abs(sign_normalise(R) - sign_normalise(R_np)).max() / np.abs(R_np).max() < 1e-13


@pytest.mark.parametrize("MT,NT,nb,ib", [(4, 4, 8, 4), (3, 3, 16, 8)])
def test_make_q_and_dorgqr_give_qt(orc, MT, NT, nb, ib):
    n = NT * nb
    A = orc.fill_uniform_cm(MT * nb, n, 12)
    At = orc.cm_to_tile(A, nb)
    RY, TAU = orc.tile_qr(At, MT, NT, nb, ib)
    R = r_from_tiles(RY, MT, NT, nb, orc)
    Qt_tiles, RY0 = orc.make_Q(RY, TAU, MT, NT, nb)
    Qt = orc.tile_to_cm(Qt_tiles, MT, NT, nb)     # the array holds Q^T
    Q = Qt.T
    assert np.linalg.norm(A - Q @ R) / np.linalg.norm(A) < 1e-14
    assert np.linalg.norm(np.eye(n) - Q.T @ Q) < 1e-13
    # make_Q leaves R only (reflectors zeroed)
    assert np.abs(orc.tile_to_cm(RY0, MT, NT, nb) - R).max() == 0.0
    # dorgqr (homegrown kernels, RY untouched) produces the same Q^T
    Qt2 = orc.tile_to_cm(orc.dorgqr(RY, TAU, MT, NT, nb), MT, NT, nb)
    assert np.abs(Qt2 - Qt).max() < 1e-13
    # dormqr: C <- Q^T C
    Cm = orc.fill_uniform_cm(MT * nb, n, 13)
    out = orc.tile_to_cm(orc.dormqr(RY, TAU, orc.cm_to_tile(Cm, nb), MT, NT, nb), MT, NT, nb)
    assert np.abs(out - Qt @ Cm).max() < 1e-13


def test_tall_tile_qr_terminates_and_factors(orc):
    MT, NT, nb, ib = 6, 2, 8, 4
    A = orc.fill_uniform_cm(MT * nb, NT * nb, 14)
    At = orc.cm_to_tile(A, nb)
    RY, TAU = orc.tile_qr(At, MT, NT, nb, ib, nthreads=3)
    R = r_from_tiles(RY, MT, NT, nb, orc)
    R_np = np.linalg.qr(A, mode="r")
    assert np.abs(sign_normalise(R) - sign_normalise(R_np)).max() < 1e-13


@pytest.mark.parametrize("threads", [2, 4, 8])
def test_dynamic_schedule_is_bitwise_identical(orc, threads):
    MT, NT, nb, ib = 6, 6, 8, 4
    At = orc.cm_to_tile(orc.fill_uniform_cm(MT * nb, NT * nb, 15), nb)
    a_seq, t_seq, T_seq = orc.tile_qr(At, MT, NT, nb, ib, nthreads=1, want_T=True)
    a_dyn, t_dyn, T_dyn = orc.tile_qr(At, MT, NT, nb, ib, nthreads=threads, want_T=True)
    assert np.array_equal(a_seq, a_dyn) and np.array_equal(t_seq, t_dyn)
    assert np.array_equal(T_seq, T_dyn)


def test_splitmix_known_answer(orc):
    # SplitMix64 (Steele, Lea, Flood 2014), state 0: first output 0xE220A8397B1DCDAF
    z = 0xE220A8397B1DCDAF
    assert orc.uniform_pm1(0, 0) == 2.0 * ((z >> 11) * 2.0 ** -53) - 1.0
    z2 = 0x6E789E6AA1B965F4
    assert orc.uniform_pm1(0, 1) == 2.0 * ((z2 >> 11) * 2.0 ** -53) - 1.0
    A = orc.fill_uniform_cm(7, 5, 99)
    assert A.min() >= -1.0 and A.max() < 1.0 and A[3, 2] == orc.uniform_pm1(99, 2 * 7 + 3)


def test_glibc_generator_layout(orc):
    """mat_prod.c:157-182 walks tile columns, columns, tile rows, rows: the
    global column-major order of the rand() stream."""
    MT, NT, nb = 3, 2, 4
    At = orc.mat_make_glibc(MT, NT, nb, 20151130)
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(20151130)
    M, N = MT * nb, NT * nb
    seq = np.array([libc.rand() / 2147483647.0 for _ in range(M * N)])
    A = seq.reshape(N, M).T
    assert np.array_equal(orc.tile_to_cm(At, MT, NT, nb), A)
