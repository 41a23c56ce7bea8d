"""C-ABI contract tests that need no GPU: the library loads, exports every
function include/*.h declares, validates arguments before touching the
device, and mat_init follows the reference's geometry (mat_prod.c:15-37)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def hq():
    import hybrid3_amd
    hybrid3_amd.lib()
    return hybrid3_amd


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", src)
        for m in re.finditer(r"\b(?:void|int|double|char)\s*\*?\s*(\w+)\s*\(", src):
            names.append(m.group(1))
    return sorted(set(names))


def test_every_declared_symbol_is_exported(hq):
    names = declared_functions()
    assert len(names) >= 30
    lib = C.CDLL(hq.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_reference_entry_points_present(hq):
    lib = C.CDLL(hq.LIB_PATH)
    for n in ["Hybrid_tile_QR", "Hybrid_make_Q", "Hybrid_dorgqr", "Hybrid_dormqr", "Hybrid_make_I",
              "Hybrid_dgemm", "mat_init", "tile_init", "tile_dgeqrt", "tile_dtsqrt", "tile_dlarfb",
              "tile_dssrfb", "tile_dlarft"]:
        assert hasattr(lib, n), n


def test_mat_init_matches_reference_formula(hq):
    for M, N, nb, rank, P in [(2048, 2048, 128, 0, 1), (1000, 700, 128, 0, 1),
                              (1024, 1024, 128, 1, 3), (512, 256, 64, 0, 1)]:
        st = hq.mat_init(M, N, nb, rank, P)
        mt, nt = M // nb, N // nb
        mine = nt // P + (1 if nt % P > rank else 0)
        assert (st.mat_M, st.mat_N, st.my_mat_M) == (M, N, M)
        assert (st.glob_tile_M, st.glob_tile_N, st.my_tile_M) == (mt, nt, mt)
        assert (st.my_tile_N, st.my_mat_N, st.tile_M, st.tile_N) == (mine, nb * mine, nb, nb)


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_rank_and_geometry_errors_leave_arrays_untouched(hq):
    L = hq.lib()
    st = hq.mat_init(256, 256, 64)
    A = np.random.default_rng(0).uniform(-1, 1, 256 * 256)
    TAU = np.zeros(128 * 128)
    A0 = A.copy()
    L.Hybrid_tile_QR(C.byref(st), _p(A), _p(TAU), 0, 2, 16)   # Proc_num != 1
    assert hq.last_error()[0] == 2 and np.array_equal(A, A0)
    L.Hybrid_tile_QR(C.byref(st), _p(A), _p(TAU), 1, 1, 16)   # my_rank != 0
    assert hq.last_error()[0] == 2
    L.Hybrid_tile_QR(C.byref(st), _p(A), _p(TAU), 0, 1, 0)    # ib <= 0 (main.c:75-84)
    assert hq.last_error()[0] == 1 and np.array_equal(A, A0)
    L.Hybrid_tile_QR(C.byref(st), _p(A), _p(TAU), 0, 1, 33)   # ib above the supported 32
    assert hq.last_error()[0] == 1
    st2 = hq.mat_init(300, 300, 100)                           # nb = 100 unsupported
    L.Hybrid_tile_QR(C.byref(st2), _p(A), _p(TAU), 0, 1, 16)
    code, msg = hq.last_error()
    assert code == 1 and "nb=100" in msg
    L.Hybrid_dorgqr(C.byref(st), _p(A), _p(A0), _p(TAU), 0, 4)
    assert hq.last_error()[0] == 2
    L.tile_dgeqrt(0, 8, 8, _p(A), 8, _p(TAU), _p(A), _p(A), 8, None)
    assert hq.last_error()[0] == 1
    assert np.array_equal(A, A0)


def test_device_entry_points_validate_before_launch(hq):
    L = hq.lib()
    assert L.hq_tile_qr_dev(4, 4, 96, 16, None, None, None, None) == 1
    assert L.hq_tile_qr_dev(0, 4, 64, 16, None, None, None, None) == 1
    assert L.hq_dgeqrt_dev(64, 0, None, None, None, None) == 1
    assert L.hq_tile_ormqr_dev(4, 4, 64, 16, None, None, None, None, 0, None) == 1


def test_tile_init_allocates_three_tiles_of_zeros(hq):
    L = hq.lib()
    p = L.tile_init(b"d", 16, 8)
    arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(3 * 16 * 8,))
    assert np.all(arr == 0.0)
    C.CDLL("libc.so.6").free(C.c_void_p(p))
