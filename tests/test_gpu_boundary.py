"""GPU tests of the drop-in boundary beyond the tile kernels themselves:

* the PLASMA core_blas shim (include/core_blas.h) with the reference's own
  argument pattern (LDT = nb for an ib x nb T, Hybrid_tile_QR.c:985-990,
  1135-1141,1276-1283,1356-1363) against the oracle's PLASMA restatement;
* any 1 <= ib <= nb (main.c:85-88 defaults ib to nb) in the tile QR and the
  single-tile kernels, T included, against the oracle at the same ib;
* edge tiles (m, n not multiples of nb) against the oracle on the
  zero-padded matrix and LAPACK's R of the unpadded one;
* tile_kernel.h calls from several host threads at once (the reference calls
  them inside `omp for`, Hybrid_dorgqr.c:92-108);
* the scheduler's bounded dependency wait: a short bound set from the host
  makes a run give up, every workgroup exits, the error is reported, and the
  next run is correct;
* the reference's driver program (main.c) end to end on the library.
FP64; 1e-12 relative to the largest entry."""
import ctypes as C
import os
import re
import subprocess
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-12
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def hq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import hybrid3_amd
    hybrid3_amd.lib()
    return hybrid3_amd


def P(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def F(a):
    return np.array(a, dtype=np.float64, order="F", copy=True)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def padded(a, ld):
    """a (m x n) stored column-major with leading dimension ld > m (junk in the gap)."""
    m, n = a.shape
    buf = np.full((ld, n), 12345.0, order="F")
    buf[:m] = a
    return buf


@pytest.mark.parametrize("nb,ib", [(128, 32), (128, 48), (128, 128), (256, 32), (256, 64), (64, 16)])
def test_core_shim_matches_plasma_restatement(hq, orc, nb, ib):
    L = hq.lib()
    rng = np.random.default_rng(nb * 7 + ib)
    ld = nb + 5
    print(f"CORE_* nb={nb} ib={ib}", flush=True)
    # CORE_dgeqrt with T of ib x nb at LDT = nb (as the reference passes it)
    A = rng.uniform(-1, 1, (nb, nb))
    Ao, To, tauo = orc.dgeqrt(A, ib)
    Ah, Th, tauh = padded(A, ld), np.zeros((nb, nb), order="F"), np.zeros(nb)
    assert L.CORE_dgeqrt(nb, nb, ib, P(Ah), ld, P(Th), nb, P(tauh), None) == 0
    assert rel(Ah[:nb], Ao) < TOL and rel(Th[:ib], To) < TOL and np.abs(tauh - tauo).max() < TOL
    assert np.all(Ah[nb:] == 12345.0) and np.all(Th[ib:] == 0.0)
    # CORE_dtsqrt on [R; B]
    R = np.triu(Ao)
    B = rng.uniform(-1, 1, (nb, nb))
    r1o, vo, t2o, tau2o = orc.dtsqrt(R, B, ib)
    R1h, Bh, T2h, tau2h = padded(R, ld), padded(B, ld), np.zeros((nb, nb), order="F"), np.zeros(nb)
    assert L.CORE_dtsqrt(nb, nb, ib, P(R1h), ld, P(Bh), ld, P(T2h), nb, P(tau2h), None) == 0
    assert rel(R1h[:nb], r1o) < TOL and rel(Bh[:nb], vo) < TOL and rel(T2h[:ib], t2o) < TOL
    # CORE_dormqr(Left, Trans) with the GEQRT reflectors
    Cm = rng.uniform(-1, 1, (nb, nb))
    co = orc.dormqr_lt(Ao, To, Cm, ib)
    Ch = padded(Cm, ld)
    Tin = np.zeros((nb, nb), order="F")
    Tin[:ib] = To
    if nb < 128 and ib > 32:
        return
    assert L.CORE_dormqr(141, 112, nb, nb, nb, ib, P(padded(Ao, ld)), ld, P(Tin), nb, P(Ch), ld, None, nb) == 0
    assert rel(Ch[:nb], co) < TOL
    # CORE_dtsmqr(Left, Trans) with the TSQRT reflectors
    X1, X2 = rng.uniform(-1, 1, (nb, nb)), rng.uniform(-1, 1, (nb, nb))
    x1o, x2o = orc.dtsmqr_lt(X1, X2, vo, t2o, ib)
    X1h, X2h = padded(X1, ld), padded(X2, ld)
    T2in = np.zeros((nb, nb), order="F")
    T2in[:ib] = t2o
    assert L.CORE_dtsmqr(141, 112, nb, nb, nb, nb, nb, ib, P(X1h), ld, P(X2h), ld, P(padded(vo, ld)), ld,
                         P(T2in), nb, None, ib) == 0
    assert rel(X1h[:nb], x1o) < TOL and rel(X2h[:nb], x2o) < TOL


@pytest.mark.parametrize("MT,NT,nb,ib", [(3, 3, 128, 128), (4, 2, 128, 64), (2, 3, 256, 48), (3, 3, 64, 64)])
def test_tile_qr_any_ib(hq, orc, MT, NT, nb, ib):
    import torch
    A = orc.fill_uniform_cm(MT * nb, NT * nb, 900 + ib)
    At = orc.cm_to_tile(A, nb)
    RYo, TAUo, To = orc.tile_qr(At, MT, NT, nb, ib, want_T=True)
    dA = torch.from_numpy(At.copy()).cuda()
    dTAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    dT = torch.zeros(MT * NT * ib * nb, dtype=torch.float64, device="cuda")
    print(f"tile QR {MT}x{NT} nb={nb} ib={ib}", flush=True)
    hq.tile_qr_dev(MT, NT, nb, ib, dA, dTAU, dT)
    hq.sync()
    assert rel(dA.cpu().numpy(), RYo) < TOL
    assert np.abs(dTAU.cpu().numpy() - TAUo).max() < TOL
    # T of the reflector tiles (r >= k) at the caller's ib
    T = dT.cpu().numpy().reshape(NT, MT, nb, ib)
    To = To.reshape(NT, MT, nb, ib)
    for k in range(min(MT, NT)):
        for r in range(k, MT):
            assert rel(T[k, r], To[k, r]) < TOL, (r, k)


def test_host_tile_qr_with_default_ib(hq, orc):
    """main.c's default: ib = nb (Hybrid3 N NB without IB)."""
    L = hq.lib()
    nb, MT = 128, 3
    st = hq.mat_init(MT * nb, MT * nb, nb)
    A = orc.mat_make_glibc(MT, MT, nb, 20151130)
    RYo, TAUo = orc.tile_qr(A, MT, MT, nb, nb)
    TAU = np.zeros((MT * nb // 2) ** 2)
    L.Hybrid_tile_QR(C.byref(st), P(A), P(TAU), 0, 1, nb)
    assert hq.last_error()[0] == 0
    assert rel(A, RYo) < TOL and np.abs(TAU[:MT * MT * nb] - TAUo).max() < TOL


@pytest.mark.parametrize("m,n,nb,ib", [(300, 200, 128, 32), (200, 450, 128, 32), (700, 613, 256, 32),
                                       (129, 129, 64, 16)])
def test_edge_tiles(hq, orc, m, n, nb, ib):
    import torch
    MT, NT = -(-m // nb), -(-n // nb)
    A = orc.fill_uniform_cm(m, n, m + n)
    Ap = np.zeros((MT * nb, NT * nb), order="F")
    Ap[:m, :n] = A
    At = orc.cm_to_tile(Ap, nb)
    RYo, TAUo = orc.tile_qr(At, MT, NT, nb, ib)
    junk = orc.cm_to_tile(np.where(np.arange(MT * nb)[:, None] < m, 0.0, 7.0) +
                          np.where(np.arange(NT * nb)[None, :] < n, 0.0, 5.0), nb)
    dA = torch.from_numpy(At + junk).cuda()      # junk in the padding: the call zeroes it
    dTAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    print(f"edge tiles {m}x{n} nb={nb}", flush=True)
    hq.tile_qr_mn_dev(m, n, nb, ib, dA, dTAU)
    hq.sync()
    RY = dA.cpu().numpy()
    assert rel(RY, RYo) < TOL
    assert np.abs(dTAU.cpu().numpy() - TAUo).max() < TOL
    R = np.triu(orc.tile_to_cm(RY, MT, NT, nb))[:min(m, n), :n]
    Rl = np.linalg.qr(A, mode="r")
    s = np.sign(np.diag(R)) * np.sign(np.diag(Rl))
    assert rel(R * s[:, None], Rl) < 1e-12


def test_tile_kernels_from_several_threads(hq, orc):
    """tile_dssrfb on disjoint tiles from 4 host threads at once."""
    L = hq.lib()
    n = 128
    rng = np.random.default_rng(77)
    R1 = F(np.triu(rng.uniform(-1, 1, (n, n))))
    _, V2, TAU2, _ = orc.tile_dtsqrt(R1, F(rng.uniform(-1, 1, (n, n))), n)
    T2 = orc.tile_dlarft(V2, TAU2, 1)
    jobs = []
    for t in range(4):
        B1, B2 = F(rng.uniform(-1, 1, (n, n))), F(rng.uniform(-1, 1, (n, n)))
        jobs.append((F(B1), F(B2), orc.tile_dssrfb(B1, B2, V2, T2)))
    errs = []

    def run(j):
        g1, g2, _ = jobs[j]
        for _ in range(5):
            a1, a2 = F(g1), F(g2)
            L.tile_dssrfb(n, n, P(a1), P(a2), n, P(V2), P(T2), n, None)
            if L.hq_last_error() != 0:
                errs.append(L.hq_last_error())
        jobs[j] = (g1, g2, jobs[j][2], a1, a2)

    th = [threading.Thread(target=run, args=(j,)) for j in range(4)]
    print("tile_dssrfb from 4 threads", flush=True)
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errs
    for g1, g2, (o1, o2), a1, a2 in jobs:
        assert rel(a1, o1) < TOL and rel(a2, o2) < TOL


def test_scheduler_wait_bound_reports_and_recovers(hq, orc):
    import torch
    MT = NT = 4
    nb, ib = 128, 32
    A = orc.fill_uniform_cm(MT * nb, NT * nb, 31)
    At = orc.cm_to_tile(A, nb)
    RYo, TAUo = orc.tile_qr(At, MT, NT, nb, ib)
    dTAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    L = hq.lib()
    print("scheduler wait bound: 1 poll", flush=True)
    assert L.hq_set_wait_limit(1) == 0
    dA = torch.from_numpy(At.copy()).cuda()
    hq.tile_qr_dev(MT, NT, nb, ib, dA, dTAU)
    with pytest.raises(hq.HQError, match="timed out"):
        hq.sync()
    assert L.hq_set_wait_limit(0) == 0
    dA = torch.from_numpy(At.copy()).cuda()
    hq.tile_qr_dev(MT, NT, nb, ib, dA, dTAU)
    hq.sync()
    assert rel(dA.cpu().numpy(), RYo) < TOL


@pytest.mark.parametrize("args", [["512", "128", "32"], ["1024", "256"]])
def test_reference_driver_end_to_end(hq, args):
    exe = os.path.join(ROOT, "hybrid3_amd", "hybrid3")
    print("hybrid3", *args, flush=True)
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    n, nb = int(args[0]), int(args[1])
    ib = int(args[2]) if len(args) > 2 else nb
    m = re.search(r"node= 1\tN= (\d+)\tNB= (\d+)\tIB= (\d+)\ttime= ([0-9.]+) s\tflops= ([0-9.]+) Gflops", out)
    assert m and (int(m.group(1)), int(m.group(2)), int(m.group(3))) == (n, nb, ib), out
    assert "CRAYJ: Elapsed time for Hybrid_make_Q() ==" in out
    norms = {k: float(v) for k, v in re.findall(r"\|\|(Q|I-Q'Q|A-QR)\|\|oo=\t(\S+)", out)}
    # ||Q||oo is the largest column sum of |Q^T| in tile layout: between 1 and sqrt(n)
    assert 1.0 <= norms["Q"] <= np.sqrt(n) + 1e-9
    assert norms["I-Q'Q"] < 1e-13 * n
    assert norms["A-QR"] < 1e-13 * n * n
    assert "end ||A-QR||oo" in r.stderr
