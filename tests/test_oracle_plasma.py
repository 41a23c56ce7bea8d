"""Pins the PLASMA-convention oracle kernels (GEQRT/TSQRT/LARFB/SSRFB as
called at Hybrid_tile_QR.c:985,1135,1276,1356) against LAPACK golden vectors
(tests/golden/lapack_tile_kernels.npz) and against mathematical identities."""
import os

import numpy as np
import pytest

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "lapack_tile_kernels.npz"))
TOL = 1e-13


@pytest.mark.parametrize("tag", ["sq", "tall", "wide", "ib1", "ibn"])
def test_geqrt_matches_lapack(orc, tag):
    A = G[f"geqrt_{tag}_A"]
    ib = int(G[f"geqrt_{tag}_ib"][0])
    a, T, tau = orc.dgeqrt(A, ib)
    assert np.abs(a - G[f"geqrt_{tag}_out"]).max() < TOL
    k = min(A.shape)
    # LAPACK stores T as ib x k; compare the blocks that exist
    assert np.abs(T[:, :k] - G[f"geqrt_{tag}_T"][:, :k]).max() < TOL


@pytest.mark.parametrize("tag", ["sq", "rect", "ib1"])
def test_tsqrt_matches_lapack(orc, tag):
    ib = int(G[f"tpqrt_{tag}_ib"][0])
    a1, a2, T, tau = orc.dtsqrt(G[f"tpqrt_{tag}_A1"], G[f"tpqrt_{tag}_B"], ib)
    assert np.abs(a1 - G[f"tpqrt_{tag}_A1out"]).max() < TOL
    assert np.abs(a2 - G[f"tpqrt_{tag}_Bout"]).max() < TOL
    assert np.abs(T - G[f"tpqrt_{tag}_T"]).max() < TOL


@pytest.mark.parametrize("tag", ["sq", "rect", "ib1"])
def test_tsmqr_matches_lapack(orc, tag):
    ib = int(G[f"tpqrt_{tag}_ib"][0])
    c1, c2 = orc.dtsmqr_lt(G[f"tpmqrt_{tag}_C1"], G[f"tpmqrt_{tag}_C2"],
                           G[f"tpqrt_{tag}_Bout"], G[f"tpqrt_{tag}_T"], ib)
    assert np.abs(c1 - G[f"tpmqrt_{tag}_C1out"]).max() < TOL
    assert np.abs(c2 - G[f"tpmqrt_{tag}_C2out"]).max() < TOL


@pytest.mark.parametrize("tag", ["sq", "ib3"])
def test_ormqr_matches_lapack(orc, tag):
    ib = int(G[f"gemqrt_{tag}_ib"][0])
    c = orc.dormqr_lt(G[f"gemqrt_{tag}_V"], G[f"gemqrt_{tag}_T"], G[f"gemqrt_{tag}_C"], ib)
    assert np.abs(c - G[f"gemqrt_{tag}_Cout"]).max() < TOL


def _explicit_q(V, tau):
    """Q = H(0) H(1) ... from explicit reflector matrices (V unit lower)."""
    m, k = V.shape
    Q = np.eye(m)
    for i in range(k):
        v = np.zeros(m)
        v[i] = 1.0
        v[i + 1:] = V[i + 1:, i]
        Q = Q @ (np.eye(m) - tau[i] * np.outer(v, v))
    return Q


def test_geqrt_sign_convention_and_reconstruction(orc):
    rng = np.random.default_rng(7)
    A = rng.uniform(-1, 1, (24, 24))
    a, T, tau = orc.dgeqrt(A, 6)
    R = np.triu(a)
    Q = _explicit_q(a, tau)
    assert np.linalg.norm(Q @ R - A) / np.linalg.norm(A) < 1e-14
    assert np.linalg.norm(Q.T @ Q - np.eye(24)) < 1e-13
    # LAPACK dlarfg: beta = -sign(alpha)*norm -> |R| equals numpy's |R|
    R_np = np.linalg.qr(A, mode="r")
    assert np.abs(np.abs(R) - np.abs(R_np)).max() < 1e-13
    # last reflector of a square tile is the identity (n=1 dlarfg)
    assert tau[-1] == 0.0


def test_dlarfg_closed_form(orc):
    import ctypes as C
    for alpha0, x0 in [(3.0, [4.0]), (-3.0, [4.0]), (0.5, [0.0, 0.0]), (0.0, [2.0, 0.0])]:
        x = np.array(x0, dtype=np.float64)
        alpha = C.c_double(alpha0)
        tau = C.c_double(0)
        orc.lib().orc_dlarfg(len(x) + 1, C.byref(alpha), x.ctypes.data_as(C.POINTER(C.c_double)),
                             1, C.byref(tau))
        nrm = np.hypot(alpha0, np.linalg.norm(x0))
        if np.linalg.norm(x0) == 0:
            assert tau.value == 0.0 and alpha.value == alpha0
            continue
        beta = -np.copysign(nrm, alpha0)
        assert abs(alpha.value - beta) < 1e-15
        assert abs(tau.value - (beta - alpha0) / beta) < 1e-15
        v = np.concatenate([[1.0], x])
        H = np.eye(len(v)) - tau.value * np.outer(v, v)
        y = H @ np.concatenate([[alpha0], x0])
        assert abs(y[0] - beta) < 1e-14 and np.abs(y[1:]).max() < 1e-14


def test_planted_mistake_in_oracle_is_caught():
    """A wrong sign in T (a plausible slip) must break the LAPACK comparison."""
    A = G["gemqrt_sq_V"]
    T = G["gemqrt_sq_T"].copy()
    T[0, 1] = -T[0, 1]
    import oracle
    c = oracle.dormqr_lt(A, T, G["gemqrt_sq_C"], 8)
    assert np.abs(c - G["gemqrt_sq_Cout"]).max() > 1e-6
