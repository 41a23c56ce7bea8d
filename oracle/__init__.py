"""CPU oracle for the tiled Householder QR path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only as the checker (or the timed CPU
baseline).  The shipped library never calls it.

The C restatement lives in ``hq_oracle.c`` (see its header for what each
routine restates and how it is pinned); this module wraps it with ctypes and
numpy.  Arrays are column-major float64 (``np.asfortranarray``); tile-layout
arrays are flat float64 vectors laid out as in mat_prod.h:57-67.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(_SO):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        i, d, u64 = C.c_int, C.c_double, C.c_uint64
        sig = {
            "orc_dlarfg": (None, [i, _dp, _dp, i, _dp]),
            "orc_dgeqrt": (None, [i, i, i, _dp, i, _dp, i, _dp]),
            "orc_dtsqrt": (None, [i, i, i, _dp, i, _dp, i, _dp, i, _dp]),
            "orc_dormqr_lt": (None, [i, i, i, i, _dp, i, _dp, i, _dp, i]),
            "orc_dtsmqr_lt": (None, [i, i, i, i, i, i, _dp, i, _dp, i, _dp, i, _dp, i]),
            "orc_dlarft_fc": (None, [i, i, _dp, i, _dp, _dp, i]),
            "orc_dtsrft": (None, [i, i, i, _dp, i, _dp, _dp, i]),
            "orc_tile_dgeqrt": (None, [i, i, i, _dp, i, _dp, _dp, _dp, i, _dp]),
            "orc_tile_dtsqrt": (None, [i, i, i, _dp, i, _dp, i, _dp, _dp, i, _dp]),
            "orc_tile_dlarfb": (None, [i, i, _dp, i, _dp, _dp, i, _dp]),
            "orc_tile_dssrfb": (None, [i, i, _dp, _dp, i, _dp, _dp, i, _dp]),
            "orc_tile_dlarft": (None, [i, i, i, _dp, i, _dp, _dp, i, _dp]),
            "orc_tile_qr": (i, [i, i, i, i, _dp, _dp, _dp, i]),
            "orc_make_Q": (None, [i, i, i, _dp, _dp, _dp]),
            "orc_dorgqr": (None, [i, i, i, _dp, _dp, _dp]),
            "orc_dormqr": (None, [i, i, i, _dp, _dp, _dp]),
            "orc_uniform_pm1": (d, [u64, u64]),
            "orc_fill_uniform_cm": (None, [i, i, u64, _dp]),
            "orc_mat_make_glibc": (None, [i, i, i, C.c_uint, _dp]),
            "orc_cm_to_tile": (None, [i, i, i, _dp, _dp]),
            "orc_tile_to_cm": (None, [i, i, i, _dp, _dp]),
            "orc_time_tile_qr": (d, [i, i, i, i, u64, i, i]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(_dp)


def _f(a) -> np.ndarray:
    return np.array(a, dtype=np.float64, order="F", copy=True)


# ---------------------------------------------------------------- PLASMA style
def dgeqrt(A, ib):
    """CORE_dgeqrt: returns (A_out with R/V, T (ib x n), tau)."""
    A = _f(A)
    m, n = A.shape
    k = min(m, n)
    T = np.zeros((ib, n), order="F")
    tau = np.zeros(max(k, 1))
    lib().orc_dgeqrt(m, n, ib, _p(A), m, _p(T), ib, _p(tau))
    return A, T, tau[:k]


def dtsqrt(A1, A2, ib):
    """CORE_dtsqrt: QR of [triu(A1); A2] -> (A1_out, A2_out=V2, T, tau)."""
    A1, A2 = _f(A1), _f(A2)
    n = A1.shape[1]
    m = A2.shape[0]
    T = np.zeros((ib, n), order="F")
    tau = np.zeros(n)
    lib().orc_dtsqrt(m, n, ib, _p(A1), A1.shape[0], _p(A2), m, _p(T), ib, _p(tau))
    return A1, A2, T, tau


def dormqr_lt(V, T, Cm, ib, k=None):
    """CORE_dormqr(Left, Trans): C <- Q^T C."""
    V, T, Cm = _f(V), _f(T), _f(Cm)
    m, n = Cm.shape
    k = V.shape[1] if k is None else k
    lib().orc_dormqr_lt(m, n, k, ib, _p(V), V.shape[0], _p(T), T.shape[0], _p(Cm), m)
    return Cm


def dtsmqr_lt(A1, A2, V, T, ib, k=None):
    """CORE_dtsmqr(Left, Trans): [A1; A2] <- Q^T [A1; A2], Q = [I; V] block."""
    A1, A2, V, T = _f(A1), _f(A2), _f(V), _f(T)
    m1, n1 = A1.shape
    m2, n2 = A2.shape
    k = V.shape[1] if k is None else k
    lib().orc_dtsmqr_lt(m1, n1, m2, n2, k, ib, _p(A1), m1, _p(A2), m2, _p(V), V.shape[0],
                        _p(T), T.shape[0])
    return A1, A2


def dlarft_fc(V, tau):
    V = _f(V)
    n, k = V.shape
    T = np.zeros((k, k), order="F")
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    lib().orc_dlarft_fc(n, k, _p(V), n, _p(tau), _p(T), k)
    return T


def dtsrft(V, tau, ib):
    V = _f(V)
    m, n = V.shape
    T = np.zeros((ib, n), order="F")
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    lib().orc_dtsrft(m, n, ib, _p(V), m, _p(tau), _p(T), ib)
    return T


# ------------------------------------------------------------- homegrown API
def tile_dgeqrt(A, block, Y=None, T=None, TAU=None):
    A = _f(A)
    row, col = A.shape
    Y = np.zeros((row, col), order="F") if Y is None else _f(Y)
    T = np.zeros((col, col), order="F") if T is None else _f(T)
    TAU = np.zeros(col) if TAU is None else np.array(TAU, dtype=np.float64)
    work = np.zeros(3 * max(row, col) ** 2 + 16)
    lib().orc_tile_dgeqrt(row, col, block, _p(A), row, _p(TAU), _p(Y), _p(T), T.shape[0],
                          _p(work))
    return A, TAU, Y, T


def tile_dtsqrt(A1, A2, block, T=None, TAU=None):
    A1, A2 = _f(A1), _f(A2)
    row, col = A2.shape
    T = np.zeros((col, col), order="F") if T is None else _f(T)
    TAU = np.zeros(col) if TAU is None else np.array(TAU, dtype=np.float64)
    work = np.zeros(3 * max(row, col) ** 2 + 16)
    lib().orc_tile_dtsqrt(row, col, block, _p(A1), A1.shape[0], _p(A2), row, _p(TAU),
                          _p(T), T.shape[0], _p(work))
    return A1, A2, TAU, T


def tile_dlarfb(A, Y, T):
    A, Y, T = _f(A), _f(Y), _f(T)
    row, col = A.shape
    work = np.zeros(3 * max(row, col) ** 2 + 16)
    lib().orc_tile_dlarfb(row, col, _p(A), row, _p(Y), _p(T), T.shape[0], _p(work))
    return A


def tile_dssrfb(A1, A2, Y, T):
    A1, A2, Y, T = _f(A1), _f(A2), _f(Y), _f(T)
    row, col = A1.shape
    work = np.zeros(3 * max(row, col) ** 2 + 16)
    lib().orc_tile_dssrfb(row, col, _p(A1), _p(A2), row, _p(Y), _p(T), T.shape[0], _p(work))
    return A1, A2


def tile_dlarft(Y, TAU, block, T=None):
    Y = _f(Y)
    row, col = Y.shape
    T = np.zeros((col, col), order="F") if T is None else _f(T)
    TAU = np.ascontiguousarray(TAU, dtype=np.float64)
    work = np.zeros(3 * max(row, col) ** 2 + 16)
    lib().orc_tile_dlarft(row, col, block, _p(Y), row, _p(TAU), _p(T), T.shape[0], _p(work))
    return T


# ------------------------------------------------------------------- drivers
def cm_to_tile(A, nb):
    A = _f(A)
    M, N = A.shape
    out = np.zeros(M * N)
    lib().orc_cm_to_tile(M // nb, N // nb, nb, _p(A), _p(out))
    return out


def tile_to_cm(t, MT, NT, nb):
    out = np.zeros((MT * nb, NT * nb), order="F")
    t = np.ascontiguousarray(t, dtype=np.float64)
    lib().orc_tile_to_cm(MT, NT, nb, _p(t), _p(out))
    return out


def tile_qr(At, MT, NT, nb, ib, nthreads=1, want_T=False):
    """Hybrid_tile_QR semantics on a tile-layout vector.  Returns (A, TAU[, T])."""
    A = np.array(At, dtype=np.float64, copy=True)
    TAU = np.zeros(MT * NT * nb)
    T = np.zeros(MT * NT * ib * nb) if want_T else None
    rc = lib().orc_tile_qr(MT, NT, nb, ib, _p(A), _p(TAU), _p(T), nthreads)
    assert rc == 0
    return (A, TAU, T) if want_T else (A, TAU)


def make_Q(RY, TAU, MT, NT, nb):
    RY = np.array(RY, dtype=np.float64, copy=True)
    Q = np.zeros(MT * NT * nb * nb)
    lib().orc_make_Q(MT, NT, nb, _p(Q), _p(RY), _p(np.ascontiguousarray(TAU)))
    return Q, RY


def dorgqr(RY, TAU, MT, NT, nb):
    Q = np.zeros(MT * NT * nb * nb)
    lib().orc_dorgqr(MT, NT, nb, _p(Q), _p(np.ascontiguousarray(RY)),
                     _p(np.ascontiguousarray(TAU)))
    return Q


def dormqr(RY, TAU, Ct, MT, NT, nb):
    Cc = np.array(Ct, dtype=np.float64, copy=True)
    lib().orc_dormqr(MT, NT, nb, _p(np.ascontiguousarray(RY)), _p(np.ascontiguousarray(TAU)),
                     _p(Cc))
    return Cc


def uniform_pm1(seed, index):
    return lib().orc_uniform_pm1(seed, index)


def fill_uniform_cm(m, n, seed):
    A = np.zeros((m, n), order="F")
    lib().orc_fill_uniform_cm(m, n, seed, _p(A))
    return A


def mat_make_glibc(MT, NT, nb, seed=20151130):
    A = np.zeros(MT * NT * nb * nb)
    lib().orc_mat_make_glibc(MT, NT, nb, seed, _p(A))
    return A


def time_tile_qr(MT, NT, nb, ib, seed, nthreads, reps=1):
    return lib().orc_time_tile_qr(MT, NT, nb, ib, seed, nthreads, reps)
