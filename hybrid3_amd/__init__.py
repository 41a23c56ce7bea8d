"""hybrid3_amd -- tiled Householder QR engine for one MI355X (gfx950).

The product is the C-ABI shared library ``libhybrid3.so`` built from
``csrc/`` (HIP kernels + host runtime); its entry points are declared in
``include/*.h``.  This module is the Python-side binding used by the tests
and ``bench.py``: it loads the library with ctypes and mirrors the
reference's interface names (``mat_init``, ``Hybrid_tile_QR`` ...).

There is no fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhybrid3.so")

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class mat_stat_t(C.Structure):
    """Same fields as the reference's mat_stat_t (mat_prod.h:12-25)."""
    _fields_ = [
        ("mat_M", C.c_ulong), ("mat_N", C.c_ulong),
        ("my_mat_M", C.c_ulong), ("my_mat_N", C.c_ulong),
        ("tile_M", C.c_uint), ("tile_N", C.c_uint),
        ("glob_tile_M", C.c_uint), ("glob_tile_N", C.c_uint),
        ("my_tile_M", C.c_uint), ("my_tile_N", C.c_uint),
    ]


# name -> (restype, argtypes); every symbol declared in include/*.h
SIGNATURES = {
    # mat_prod.h
    "mat_init": (None, [C.POINTER(mat_stat_t), C.c_long, C.c_long, C.c_int, C.c_int, C.c_int, C.c_int]),
    # Hybrid_tile_QR.h
    "Hybrid_tile_QR": (None, [C.POINTER(mat_stat_t), _dp, _dp, C.c_int, C.c_int, C.c_int]),
    "Hybrid_make_Q": (None, [C.POINTER(mat_stat_t), _dp, _dp, _dp, C.c_int, C.c_int, C.c_int]),
    "Hybrid_dorgqr": (None, [C.POINTER(mat_stat_t), _dp, _dp, _dp, C.c_int, C.c_int]),
    "Hybrid_dormqr": (None, [C.POINTER(mat_stat_t), _dp, _dp, _dp, C.c_int, C.c_int]),
    "Hybrid_make_I": (None, [C.POINTER(mat_stat_t), _dp, C.c_int, C.c_int]),
    "Hybrid_dgemm": (None, [C.POINTER(mat_stat_t), _dp, _dp, _dp, C.c_int, C.c_int]),
    # tile_kernel.h
    "tile_init": (_vp, [C.c_char, C.c_int, C.c_int]),
    "tile_dgeqrt": (None, [C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, _dp, C.c_int, _dp]),
    "tile_dtsqrt": (None, [C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int, _dp]),
    "tile_dlarfb": (None, [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int, _dp]),
    "tile_dssrfb": (None, [C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_int, _dp]),
    "tile_dlarft": (None, [C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int, _dp]),
    # hybrid3_dev.h
    "hq_last_error": (C.c_int, []),
    "hq_last_error_message": (C.c_char_p, []),
    "hq_version": (C.c_char_p, []),
    "hq_tile_qr_dev": (C.c_int, [C.c_int] * 4 + [_vp, _vp, _vp, _vp]),
    "hq_tile_qr_prepare_dev": (C.c_int, [C.c_int] * 5),
    "hq_tile_qr_dev_serial": (C.c_int, [C.c_int] * 4 + [_vp, _vp, _vp, _vp]),
    "hq_schedule_info": (C.c_int, [C.c_int] * 4 + [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "hq_schedule_info_prec": (C.c_int, [C.c_int] * 5 + [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "hq_dgeqrt_dev": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "hq_dtsqrt_dev": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp]),
    "hq_dormqr_dev": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "hq_dtsmqr_dev": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp]),
    "hq_fill_uniform_dev": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint64, _vp, _vp]),
    "hq_tile_qr_dev_traced": (C.c_int, [C.c_int] * 4 + [_vp, _vp, _vp, _vp, _vp]),
    "hq_tile_qr_traced_prec_dev": (C.c_int, [C.c_int] * 5 + [_vp, _vp, _vp, _vp, _vp]),
    "hq_schedule_tasks": (C.c_int, [C.c_int] * 4 + [C.POINTER(C.c_int), C.c_int]),
    "hq_schedule_tasks_prec": (C.c_int, [C.c_int] * 5 + [C.POINTER(C.c_int), C.c_int]),
    "hq_wall_clock_khz": (C.c_int, [C.POINTER(C.c_int)]),
    "hq_tile_build_t_dev": (C.c_int, [C.c_int] * 4 + [_vp, _vp, _vp, _vp]),
    "hq_tile_ormqr_dev": (C.c_int, [C.c_int] * 4 + [_vp, _vp, _vp, _vp, C.c_int, _vp]),
    "hq_tile_ormqr_cols_dev": (C.c_int, [C.c_int] * 4 + [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp]),
    "hq_tile_ormqr_cols_traced_dev": (C.c_int, [C.c_int] * 4 + [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int,
                                                                _vp, _vp]),
    "hq_tile_orgqr_dev": (C.c_int, [C.c_int] * 4 + [_vp, _vp, _vp, _vp, _vp]),
    "hq_tile_make_q_dev": (C.c_int, [C.c_int] * 4 + [_vp, _vp, _vp, _vp]),
    "hq_dag_host": (C.c_int, [C.c_int] * 6 + [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                                             C.POINTER(C.c_int)]),
    "hq_dag_host_apply": (C.c_int, [C.c_int] * 8 + [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                                                 C.POINTER(C.c_int)]),
    "hq_dag_host_prec": (C.c_int, [C.c_int] * 7 + [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                                                  C.POINTER(C.c_int)]),
    "hq_batched_tile_qr_dev": (C.c_int, [C.c_int] * 3 + [C.POINTER(C.c_int)] * 2 + [C.POINTER(_vp)] * 3 + [_vp]),
    "hq_batched_plan_dev": (C.c_int, [C.c_int] * 3 + [C.POINTER(C.c_int)] * 2 + [C.POINTER(_vp)] * 3
                            + [C.POINTER(C.c_int)]),
    "hq_batched_run_dev": (C.c_int, [C.c_int, _vp]),
    "hq_batched_plan_check": (C.c_int, [C.c_int, _vp]),
    "hq_batched_plan_free": (C.c_int, [C.c_int]),
    "hq_batched_fill_uniform_dev": (C.c_int, [C.c_int] * 2 + [C.POINTER(C.c_int)] * 2
                                    + [C.POINTER(C.c_uint64), C.POINTER(_vp), _vp]),
    "hq_cm_to_tile_dev": (C.c_int, [C.c_int] * 2 + [C.POINTER(C.c_int)] * 2
                          + [C.POINTER(_vp), C.POINTER(C.c_int), C.POINTER(_vp), _vp]),
    "hq_tile_to_cm_dev": (C.c_int, [C.c_int] * 2 + [C.POINTER(C.c_int)] * 2
                          + [C.POINTER(_vp), C.POINTER(_vp), C.POINTER(C.c_int), _vp]),
    "hq_tile_qr_mixed_dev": (C.c_int, [C.c_int] * 4 + [_vp, _vp, _vp, _vp]),
    "hq_batched_tile_qr_prec_dev": (C.c_int, [C.c_int] * 4 + [C.POINTER(C.c_int)] * 2 + [C.POINTER(_vp)] * 3
                                    + [_vp]),
    "hq_batched_prepare_dev": (C.c_int, [C.c_int] * 4 + [C.POINTER(C.c_int)] * 2),
    "hq_batched_tile_qr_ex_dev": (C.c_int, [C.c_int] * 5 + [C.POINTER(C.c_int)] * 3 + [C.POINTER(_vp)] * 3
                                  + [_vp]),
    "hq_batched_prepare_ex_dev": (C.c_int, [C.c_int] * 5 + [C.POINTER(C.c_int)] * 2),
    "hq_batched_plan_ex_dev": (C.c_int, [C.c_int] * 4 + [C.POINTER(C.c_int)] * 3 + [C.POINTER(_vp)] * 3
                               + [C.POINTER(C.c_int)]),
    "hq_cm32_to_tile_dev": (C.c_int, [C.c_int] * 2 + [C.POINTER(C.c_int)] * 2
                            + [C.POINTER(_vp), C.POINTER(C.c_int), C.POINTER(_vp), _vp]),
    "hq_tile_to_cm32_dev": (C.c_int, [C.c_int] * 2 + [C.POINTER(C.c_int)] * 2
                            + [C.POINTER(_vp), C.POINTER(_vp), C.POINTER(C.c_int), _vp]),
    "hq_tile_qr_split_dev": (C.c_int, [C.c_int] * 4 + [_vp, _vp, _vp, _vp]),
    "hq_tile_qr_mn_dev": (C.c_int, [C.c_int] * 4 + [_vp, _vp, _vp, _vp]),
    # mat_prod.h harness helpers
    "mat_make": (None, [C.POINTER(mat_stat_t), _dp, C.c_uint, C.c_int]),
    "mat_remv": (None, [C.POINTER(mat_stat_t), _dp, C.c_uint, C.c_int]),
    "mat_comp3": (C.c_double, [C.POINTER(mat_stat_t), _dp, C.c_int, C.c_int]),
    "mat_dprint": (None, [C.POINTER(mat_stat_t), _dp, _dp]),
    "mat_print": (None, [C.POINTER(mat_stat_t), _vp, _dp, C.c_char_p]),
    # core_blas.h (PLASMA core_blas shim)
    "CORE_dgeqrt": (C.c_int, [C.c_int] * 3 + [_dp, C.c_int, _dp, C.c_int, _dp, _dp]),
    "CORE_dtsqrt": (C.c_int, [C.c_int] * 3 + [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp]),
    "CORE_dormqr": (C.c_int, [C.c_int] * 6 + [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int]),
    "CORE_dtsmqr": (C.c_int, [C.c_int] * 8 + [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp,
                                             C.c_int]),
    "hq_set_wait_limit": (C.c_int, [C.c_uint]),
    "hq_tile_qr_noalloc_dev": (C.c_int, [C.c_int] * 4 + [_vp, _vp, _vp, _vp]),
    "hq_tile_qr_noalloc_check": (C.c_int, [_vp]),
    "hq_sync": (C.c_int, [_vp]),
    "hq_finalize": (None, []),
}

_lib = None


def lib():
    """Load libhybrid3.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                " (there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if not hasattr(L, name):
                continue
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def build(verbose: bool = False) -> str:
    from . import _build
    return _build.build(verbose=verbose)


class HQError(RuntimeError):
    pass


def check(rc: int):
    if rc != 0:
        msg = lib().hq_last_error_message()
        raise HQError(f"hybrid3 error {rc}: {msg.decode() if msg else ''}")


def last_error():
    L = lib()
    return L.hq_last_error(), (L.hq_last_error_message() or b"").decode()


def mat_init(M, N, nb, my_rank=0, proc_num=1) -> mat_stat_t:
    st = mat_stat_t()
    lib().mat_init(C.byref(st), M, N, nb, nb, my_rank, proc_num)
    return st


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def Hybrid_tile_QR(st: mat_stat_t, A: np.ndarray, TAU: np.ndarray, my_rank=0, proc_num=1, ib=32):
    """Host-pointer drop-in (Hybrid_tile_QR.h:26-30) on flat tile-layout arrays."""
    lib().Hybrid_tile_QR(C.byref(st), _p(A), _p(TAU), my_rank, proc_num, ib)
    return last_error()


# ---------------------------------------------------------------- device API
def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream(stream=None):
    if stream is not None:
        return C.c_void_p(stream)
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def tile_qr_dev(MT, NT, nb, ib, A, TAU, T=None, stream=None, serial=False):
    """Factor the device tile-layout tensor A in place (torch float64 tensors)."""
    fn = lib().hq_tile_qr_dev_serial if serial else lib().hq_tile_qr_dev
    check(fn(MT, NT, nb, ib, _ptr(A), _ptr(TAU), _ptr(T), _stream(stream)))


def tile_qr_mixed_dev(MT, NT, nb, ib, A, TAU, T=None, stream=None):
    """FP64 panels, FP32 LARFB / SSRFB updates (mixed precision)."""
    check(lib().hq_tile_qr_mixed_dev(MT, NT, nb, ib, _ptr(A), _ptr(TAU), _ptr(T), _stream(stream)))


def tile_qr_mn_dev(m, n, nb, ib, A, TAU, T=None, stream=None):
    """Tile QR of an m x n matrix with edge tiles (padding zeroed by the call)."""
    check(lib().hq_tile_qr_mn_dev(m, n, nb, ib, _ptr(A), _ptr(TAU), _ptr(T), _stream(stream)))


def tile_qr_split_dev(MT, NT, nb, ib, A, TAU, T=None, stream=None):
    """FP64 panels, LARFB / SSRFB on FP16 hi+lo splits with FP32 accumulation."""
    check(lib().hq_tile_qr_split_dev(MT, NT, nb, ib, _ptr(A), _ptr(TAU), _ptr(T), _stream(stream)))


def fill_uniform_dev(MT, NT, nb, seed, A, stream=None):
    check(lib().hq_fill_uniform_dev(MT, NT, nb, seed, _ptr(A), _stream(stream)))


def _iarr(xs):
    return (C.c_int * len(xs))(*[int(x) for x in xs])


def _parr(ptrs):
    return (_vp * len(ptrs))(*[p if isinstance(p, int) else _ptr(p) for p in ptrs])


def batched_tile_qr_dev(nb, ib, MT, NT, A, TAU, T=None, stream=None):
    """Factor a batch in one launch.  MT, NT: per-matrix tile grids; A, TAU
    (and T): per-matrix device tensors or raw device addresses (int)."""
    n = len(MT)
    Tp = None if T is None else _parr(T)
    check(lib().hq_batched_tile_qr_dev(n, nb, ib, _iarr(MT), _iarr(NT), _parr(A), _parr(TAU), Tp,
                                       _stream(stream)))


def batched_tile_qr_prec_dev(prec, nb, ib, MT, NT, A, TAU, T=None, stream=None):
    """batched_tile_qr_dev with updates in FP64 (0), FP32 (1) or FP16 hi+lo splits (2)."""
    n = len(MT)
    Tp = None if T is None else _parr(T)
    check(lib().hq_batched_tile_qr_prec_dev(prec, n, nb, ib, _iarr(MT), _iarr(NT), _parr(A), _parr(TAU), Tp,
                                            _stream(stream)))


def batched_tile_qr_ex_dev(prec, split_panels, nb, ib, MT, NT, A, TAU, T=None, stream=None, ncol=None):
    """batched_tile_qr_prec_dev with the panel mapping (0 one task per tile
    panel QR, 1 one task per ib block plus per-slab trailing tasks) and,
    optionally, each matrix's column count (the rest of its last tile column
    is zero padding, whose tasks are skipped)."""
    n = len(MT)
    Tp = None if T is None else _parr(T)
    nc = None if ncol is None else _iarr(ncol)
    check(lib().hq_batched_tile_qr_ex_dev(prec, split_panels, n, nb, ib, _iarr(MT), _iarr(NT), nc, _parr(A),
                                          _parr(TAU), Tp, _stream(stream)))


def batched_prepare(prec, nb, ib, MT, NT, split_panels=0):
    """Build the schedule of a batch ahead of its first batched factorisation call."""
    check(lib().hq_batched_prepare_ex_dev(prec, split_panels, len(MT), nb, ib, _iarr(MT), _iarr(NT)))


def cm32_to_tile_dev(nb, m, n, Cs, ldc, A, stream=None):
    check(lib().hq_cm32_to_tile_dev(len(m), nb, _iarr(m), _iarr(n), _parr(Cs), _iarr(ldc), _parr(A),
                                    _stream(stream)))


def tile_to_cm32_dev(nb, m, n, A, Cs, ldc, stream=None):
    check(lib().hq_tile_to_cm32_dev(len(m), nb, _iarr(m), _iarr(n), _parr(A), _parr(Cs), _iarr(ldc),
                                    _stream(stream)))


def batched_plan_dev(nb, ib, MT, NT, A, TAU, T=None, split_panels=0, ncol=None):
    """Prepare a batch for repeated (graph-capturable) runs; returns the plan id."""
    n = len(MT)
    Tp = None if T is None else _parr(T)
    nc = None if ncol is None else _iarr(ncol)
    plan = C.c_int(0)
    check(lib().hq_batched_plan_ex_dev(split_panels, n, nb, ib, _iarr(MT), _iarr(NT), nc, _parr(A), _parr(TAU), Tp,
                                       C.byref(plan)))
    return plan.value


def batched_run_dev(plan, stream=None):
    """Enqueue one run of a prepared batch (kernel launches only)."""
    check(lib().hq_batched_run_dev(plan, _stream(stream)))


def batched_plan_check(plan, stream=None):
    check(lib().hq_batched_plan_check(plan, _stream(stream)))


def batched_plan_free(plan):
    check(lib().hq_batched_plan_free(plan))


def batched_fill_uniform_dev(nb, m, n, seeds, A, stream=None):
    cnt = len(m)
    sd = (C.c_uint64 * cnt)(*[int(x) for x in seeds])
    check(lib().hq_batched_fill_uniform_dev(cnt, nb, _iarr(m), _iarr(n), sd, _parr(A), _stream(stream)))


def cm_to_tile_dev(nb, m, n, Cs, ldc, A, stream=None):
    check(lib().hq_cm_to_tile_dev(len(m), nb, _iarr(m), _iarr(n), _parr(Cs), _iarr(ldc), _parr(A),
                                  _stream(stream)))


def tile_to_cm_dev(nb, m, n, A, Cs, ldc, stream=None):
    check(lib().hq_tile_to_cm_dev(len(m), nb, _iarr(m), _iarr(n), _parr(A), _parr(Cs), _iarr(ldc),
                                  _stream(stream)))


def tile_qr_prepare(MT, NT, nb, ib, prec=0):
    """Build the schedule and reserve the scratch of a factorisation ahead of its first call."""
    check(lib().hq_tile_qr_prepare_dev(MT, NT, nb, ib, prec))


def schedule_info(MT, NT, nb, ib, prec=0):
    """(tasks, workgroups) of the factorisation DAG with updates in precision prec."""
    n, g = C.c_int(0), C.c_int(0)
    check(lib().hq_schedule_info_prec(MT, NT, nb, ib, prec, C.byref(n), C.byref(g)))
    return n.value, g.value


def dgeqrt_dev(nb, ib, A, T, TAU, stream=None):
    check(lib().hq_dgeqrt_dev(nb, ib, _ptr(A), _ptr(T), _ptr(TAU), _stream(stream)))


def dtsqrt_dev(nb, ib, A1, A2, T, TAU, stream=None):
    check(lib().hq_dtsqrt_dev(nb, ib, _ptr(A1), _ptr(A2), _ptr(T), _ptr(TAU), _stream(stream)))


def dormqr_dev(nb, ib, V, T, Cm, stream=None):
    check(lib().hq_dormqr_dev(nb, ib, _ptr(V), _ptr(T), _ptr(Cm), _stream(stream)))


def dtsmqr_dev(nb, ib, A1, A2, V, T, stream=None):
    check(lib().hq_dtsmqr_dev(nb, ib, _ptr(A1), _ptr(A2), _ptr(V), _ptr(T), _stream(stream)))


def sync(stream=None):
    """Wait for the stream and raise if the device scheduler reported a timeout."""
    check(lib().hq_sync(_stream(stream)))


def tile_ormqr_dev(MT, NT, nb, ib, RY, TAU, X, XNT, T=None, stream=None):
    """X <- Q^T X on the device (X has MT x XNT tiles)."""
    check(lib().hq_tile_ormqr_dev(MT, NT, nb, ib, _ptr(RY), _ptr(TAU), _ptr(T), _ptr(X), XNT,
                                  _stream(stream)))


def tile_ormqr_cols_dev(MT, NT, nb, ib, RY, TAU, X, XNT, ncol, T=None, width=0, stream=None):
    """X <- Q^T X on the device for an X whose first `ncol` columns hold
    entries and whose other columns hold zeros; width 16 / 64 / 0 (pick)."""
    check(lib().hq_tile_ormqr_cols_dev(MT, NT, nb, ib, _ptr(RY), _ptr(TAU), _ptr(T), _ptr(X), XNT, ncol, width,
                                       _stream(stream)))


def tile_orgqr_dev(MT, NT, nb, ib, RY, TAU, Q, T=None, stream=None):
    """Q <- Q^T (the reference's dorgqr output array) on the device."""
    check(lib().hq_tile_orgqr_dev(MT, NT, nb, ib, _ptr(RY), _ptr(TAU), _ptr(T), _ptr(Q),
                                  _stream(stream)))


def tile_make_q_dev(MT, NT, nb, ib, RY, TAU, Q, stream=None):
    check(lib().hq_tile_make_q_dev(MT, NT, nb, ib, _ptr(RY), _ptr(TAU), _ptr(Q), _stream(stream)))


def tile_build_t_dev(MT, NT, nb, ib, RY, TAU, T, stream=None):
    check(lib().hq_tile_build_t_dev(MT, NT, nb, ib, _ptr(RY), _ptr(TAU), _ptr(T), _stream(stream)))


def dag_host(MT, NT, nb, ib, XNT=0, P=256, prec=0):
    """Task DAG in claim order (no GPU): (tasks[n,5], deps[n,6] with -1 padding);
    prec: precision of the factorisation's updates (their slab width)."""
    L = lib()
    n = C.c_int(0)
    check(L.hq_dag_host_prec(MT, NT, nb, ib, prec, XNT, P, None, None, 0, C.byref(n)))
    tasks = np.zeros(5 * n.value, dtype=np.int32)
    deps = np.zeros(6 * n.value, dtype=np.int32)
    ip = C.POINTER(C.c_int)
    check(L.hq_dag_host_prec(MT, NT, nb, ib, prec, XNT, P, tasks.ctypes.data_as(ip), deps.ctypes.data_as(ip),
                             n.value, C.byref(n)))
    return tasks.reshape(-1, 5), deps.reshape(-1, 6)


def dag_host_apply(MT, NT, nb, ib, XNT=1, ncol=0, width=16, P=256):
    """Task list of a Q^T application in claim order as the device receives it
    (no GPU): (tasks[n,5], deps[n,6]); a pipelined predecessor is -2 - index,
    -1 pads."""
    L = lib()
    n = C.c_int(0)
    check(L.hq_dag_host_apply(MT, NT, nb, ib, XNT, ncol, width, P, None, None, 0, C.byref(n)))
    tasks = np.zeros(5 * n.value, dtype=np.int32)
    deps = np.zeros(6 * n.value, dtype=np.int32)
    ip = C.POINTER(C.c_int)
    check(L.hq_dag_host_apply(MT, NT, nb, ib, XNT, ncol, width, P, tasks.ctypes.data_as(ip),
                              deps.ctypes.data_as(ip), n.value, C.byref(n)))
    return tasks.reshape(-1, 5), deps.reshape(-1, 6)
