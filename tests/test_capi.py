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
    L.Hybrid_tile_QR(C.byref(st), _p(A), _p(TAU), 0, 1, 65)   # ib above nb = 64
    assert hq.last_error()[0] == 1
    st2 = hq.mat_init(1200, 1200, 600)                         # nb = 600 unsupported (above 512)
    L.Hybrid_tile_QR(C.byref(st2), _p(A), _p(TAU), 0, 1, 16)
    code, msg = hq.last_error()
    assert code == 1 and "nb=600" in msg
    L.Hybrid_dorgqr(C.byref(st), _p(A), _p(A0), _p(TAU), 0, 4)
    assert hq.last_error()[0] == 2
    L.tile_dgeqrt(0, 8, 8, _p(A), 8, _p(TAU), _p(A), _p(A), 8, None)
    assert hq.last_error()[0] == 1
    assert np.array_equal(A, A0)


def test_device_entry_points_validate_before_launch(hq):
    L = hq.lib()
    assert L.hq_tile_qr_dev(4, 4, 600, 16, None, None, None, None) == 1
    assert L.hq_tile_qr_dev(0, 4, 64, 16, None, None, None, None) == 1
    assert L.hq_dgeqrt_dev(64, 0, None, None, None, None) == 1
    assert L.hq_tile_ormqr_dev(4, 4, 64, 16, None, None, None, None, 0, None) == 1
    # ncol outside X's last tile column, an update width other than 0 / 16 /
    # 32 / 64, 32-column updates above nb = 256
    assert L.hq_tile_ormqr_cols_dev(4, 4, 64, 16, None, None, None, None, 1, 65, 0, None) == 1
    assert L.hq_tile_ormqr_cols_dev(4, 4, 64, 16, None, None, None, None, 2, 64, 0, None) == 1
    assert L.hq_tile_ormqr_cols_dev(4, 4, 64, 16, None, None, None, None, 1, 64, 48, None) == 1
    code, msg = hq.last_error()
    assert code == 1 and "width 48" in msg
    assert L.hq_tile_ormqr_cols_dev(2, 2, 384, 16, None, None, None, None, 1, 64, 32, None) == 1
    assert "nb <= 256" in hq.last_error()[1]


def test_tile_init_allocates_three_tiles_of_zeros(hq):
    L = hq.lib()
    p = L.tile_init(b"d", 16, 8)
    arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(3 * 16 * 8,))
    assert np.all(arr == 0.0)
    C.CDLL("libc.so.6").free(C.c_void_p(p))


# ---------------------------------------------------------------- mat_prod helpers
def glibc_rand(seed, count):
    """glibc random_r TYPE_3 (the generator behind rand(), published algorithm):
    r[0] = seed, r[i] = 16807 r[i-1] mod (2^31 - 1) for i < 31, r[31..33] =
    r[0..2], r[i] = r[i-3] + r[i-31] mod 2^32; output k is r[344 + k] >> 1."""
    r = [seed & 0xFFFFFFFF]
    for i in range(1, 31):
        x = (16807 * r[i - 1]) % 2147483647
        r.append(x)
    for i in range(31, 34):
        r.append(r[i - 31])
    out = []
    i = 34
    while len(out) < count:
        r.append((r[i - 3] + r[i - 31]) & 0xFFFFFFFF)
        if i >= 344:
            out.append(r[i] >> 1)
        i += 1
    return out


def test_glibc_model_matches_known_rand_values():
    # srand(1): glibc's first rand() values
    assert glibc_rand(1, 3) == [1804289383, 846930886, 1681692777]


def test_mat_make_is_the_reference_generator(hq, orc):
    L = hq.lib()
    nb, MT = 4, 3
    st = hq.mat_init(MT * nb, MT * nb, nb)
    A = np.zeros(MT * MT * nb * nb)
    L.mat_make(C.byref(st), _p(A), 20151130, 0)
    vals = np.array(glibc_rand(20151130, A.size), dtype=np.float64) / 2147483647.0
    # loop order of mat_prod.c:165-180: tile column, column, tile row, row
    order = [((i + j * MT) * nb * nb + ii + jj * nb) for j in range(MT) for jj in range(nb)
             for i in range(MT) for ii in range(nb)]
    expect = np.zeros_like(A)
    expect[order] = vals
    assert np.array_equal(A, expect)
    assert np.array_equal(A, orc.mat_make_glibc(MT, MT, nb, 20151130))
    L.mat_make(C.byref(st), _p(A), 20151129, 1)            # seed + my_rank
    assert np.array_equal(A, expect)
    L.mat_remv(C.byref(st), _p(A), 20151130, 0)
    assert np.all(A == 0.0)


def test_mat_comp3_is_the_one_norm(hq):
    L = hq.lib()
    nb = 8
    st = hq.mat_init(24, 24, nb)
    E = np.random.default_rng(5).uniform(-1, 1, (24, 24))
    flat = np.ascontiguousarray(E.T).ravel()            # column-major storage
    assert abs(L.mat_comp3(C.byref(st), _p(flat), 0, 1) - np.linalg.norm(E, 1)) < 1e-14
    flat[100] = np.nan
    assert np.isnan(L.mat_comp3(C.byref(st), _p(flat), 0, 1))
    assert np.isnan(L.mat_comp3(C.byref(st), _p(flat), 0, 2))


def test_mat_dprint_is_tile_to_column_major(hq, orc):
    L = hq.lib()
    MT, nb = 3, 8
    st = hq.mat_init(MT * nb, MT * nb, nb)
    t = np.random.default_rng(6).uniform(-1, 1, MT * MT * nb * nb)
    out = np.zeros_like(t)
    L.mat_dprint(C.byref(st), _p(out), _p(t))
    cm = orc.tile_to_cm(t, MT, MT, nb)
    assert np.array_equal(out, cm.ravel(order="F"))


def test_mat_print_format(hq, tmp_path):
    L = hq.lib()
    libc = C.CDLL("libc.so.6")
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    st = hq.mat_init(4, 4, 2)
    t = np.arange(16, dtype=np.float64)
    path = str(tmp_path / "m.txt")
    f = libc.fopen(path.encode(), b"w")
    L.mat_print(C.byref(st), f, _p(t), b"A")
    libc.fclose(f)
    lines = open(path).read().split("\n")
    assert lines[0] == "[    0] A"
    # row 0: tiles (0,0) and (0,1), columns 0..1 each: entries 0, 2 | 8, 10
    assert lines[1].split("\t")[:4] == ["0.000000", "2.000000", "8.000000", "10.000000"]


def test_core_shim_validates_without_device(hq):
    L = hq.lib()
    A = np.zeros(64 * 64)
    assert L.CORE_dgeqrt(-1, 64, 32, _p(A), 64, _p(A), 64, _p(A), None) == -1
    assert L.CORE_dgeqrt(64, 64, 32, _p(A), 10, _p(A), 64, _p(A), None) == -5
    assert L.CORE_dgeqrt(96, 80, 32, _p(A), 96, _p(A), 96, _p(A), None) == -1   # not square
    assert L.CORE_dgeqrt(600, 600, 32, _p(A), 600, _p(A), 600, _p(A), None) == -1   # above 512
    assert L.CORE_dgeqrt(0, 64, 32, _p(A), 1, _p(A), 32, _p(A), None) == 0      # quick return
    assert L.CORE_dtsqrt(64, 64, 32, _p(A), 64, _p(A), 64, _p(A), 8, _p(A), None) == -9
    assert L.CORE_dormqr(142, 112, 64, 64, 64, 32, _p(A), 64, _p(A), 64, _p(A), 64, None, 64) == -1
    assert L.CORE_dormqr(141, 111, 64, 64, 64, 32, _p(A), 64, _p(A), 64, _p(A), 64, None, 64) == -2
    assert L.CORE_dtsmqr(141, 112, 64, 64, 64, 32, 64, 32, _p(A), 64, _p(A), 64, _p(A), 64, _p(A), 64,
                         None, 32) == -6
    assert L.CORE_dtsmqr(141, 112, 64, 64, 64, 64, 64, 48, _p(A), 64, _p(A), 64, _p(A), 64, _p(A), 64,
                         None, 32) == -8    # nb = 64 takes ib <= 32 in the updates
    code, msg = hq.last_error()
    assert code == 1 and "argument 8" in msg


def test_driver_program_is_built(hq):
    import subprocess
    exe = os.path.join(ROOT, "hybrid3_amd", "hybrid3")
    assert os.access(exe, os.X_OK)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)   # no arguments: no GPU use
    assert r.returncode == 0 and "This program need two arguments." in r.stderr
