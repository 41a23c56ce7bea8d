"""Tile sizes without a specialised kernel set (nb = 384 and 512, the larger
tiles the reference's command line accepts, and sizes that are not multiples
of 16), run by the generic kernels: the four tile kernels, the scheduled and
one-task-per-launch tile QR, T for a wider ib, Q generation, edge tiles and a
batch, each against the CPU oracle at 1e-12 relative to the largest entry
(FP64), and a whole-chip run through the column-norm invariant."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12
KERNEL_CASES = [(384, 32), (512, 32), (512, 7), (100, 12), (200, 20)]


def dev(a):
    import torch
    return torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).to("cuda")


def host(t, shape):
    return t.cpu().numpy().reshape(shape, order="F")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def hq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import hybrid3_amd
    hybrid3_amd.lib()
    return hybrid3_amd


@pytest.mark.parametrize("nb,ib", KERNEL_CASES)
def test_tile_kernels(hq, orc, nb, ib):
    import torch
    print(f"generic tile kernels nb={nb} ib={ib}", flush=True)
    rng = np.random.default_rng(nb + ib)
    # GEQRT
    A = rng.uniform(-1, 1, (nb, nb))
    Ao, To, tauo = orc.dgeqrt(A, ib)
    dA, dT = dev(A), torch.zeros(ib * nb, dtype=torch.float64, device="cuda")
    dtau = torch.zeros(nb, dtype=torch.float64, device="cuda")
    hq.dgeqrt_dev(nb, ib, dA, dT, dtau)
    torch.cuda.synchronize()
    assert rel(host(dA, (nb, nb)), Ao) < TOL
    assert rel(host(dT, (ib, nb)), To) < TOL
    assert np.abs(dtau.cpu().numpy() - tauo).max() < TOL
    # LARFB with those reflectors
    Cm = rng.uniform(-1, 1, (nb, nb))
    Co = orc.dormqr_lt(Ao, To, Cm, ib, k=nb)
    dC = dev(Cm)
    hq.dormqr_dev(nb, ib, dev(Ao), dev(To), dC)
    torch.cuda.synchronize()
    assert rel(host(dC, (nb, nb)), Co) < TOL
    # TSQRT on [R; A2]
    A1 = np.triu(rng.uniform(-1, 1, (nb, nb))) + np.tril(rng.uniform(-1, 1, (nb, nb)), -1)
    A2 = rng.uniform(-1, 1, (nb, nb))
    a1o, a2o, T2o, tau2o = orc.dtsqrt(A1, A2, ib)
    dA1, dA2 = dev(A1), dev(A2)
    dT2 = torch.zeros(ib * nb, dtype=torch.float64, device="cuda")
    dtau2 = torch.zeros(nb, dtype=torch.float64, device="cuda")
    hq.dtsqrt_dev(nb, ib, dA1, dA2, dT2, dtau2)
    torch.cuda.synchronize()
    g1 = host(dA1, (nb, nb))
    assert rel(g1, a1o) < TOL
    assert np.array_equal(np.tril(g1, -1), np.tril(A1, -1))   # GEQRT's reflectors untouched
    assert rel(host(dA2, (nb, nb)), a2o) < TOL
    assert rel(host(dT2, (ib, nb)), T2o) < TOL
    assert np.abs(dtau2.cpu().numpy() - tau2o).max() < TOL
    # SSRFB with the TSQRT reflectors
    B1, B2 = rng.uniform(-1, 1, (nb, nb)), rng.uniform(-1, 1, (nb, nb))
    o1, o2 = orc.dtsmqr_lt(B1, B2, a2o, T2o, ib, k=nb)
    d1, d2 = dev(B1), dev(B2)
    hq.dtsmqr_dev(nb, ib, d1, d2, dev(a2o), dev(T2o))
    torch.cuda.synchronize()
    assert rel(host(d1, (nb, nb)), o1) < TOL
    assert rel(host(d2, (nb, nb)), o2) < TOL


def run(hq, At, MT, NT, nb, ib, serial=False):
    import torch
    dA = torch.from_numpy(At.copy()).to("cuda")
    dTAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    dT = torch.zeros(MT * NT * ib * nb, dtype=torch.float64, device="cuda")
    hq.tile_qr_dev(MT, NT, nb, ib, dA, dTAU, dT, serial=serial)
    hq.sync()
    return dA.cpu().numpy(), dTAU.cpu().numpy(), dT.cpu().numpy()


GEOMS = [(2, 2, 384, 32), (3, 2, 512, 32), (2, 3, 512, 16), (3, 3, 100, 12), (4, 3, 200, 24),
         (2, 2, 512, 64), (2, 2, 384, 48)]


@pytest.mark.parametrize("serial", [False, True])
@pytest.mark.parametrize("MT,NT,nb,ib", GEOMS)
def test_tile_qr_matches_oracle(hq, orc, MT, NT, nb, ib, serial):
    """Scheduled and one-task-per-launch; ib > 32 runs with 32-wide blocks and
    returns T for the caller's ib (rebuilt from V and TAU)."""
    print(f"generic tile_qr MT={MT} NT={NT} nb={nb} ib={ib} serial={serial}", flush=True)
    At = orc.cm_to_tile(orc.fill_uniform_cm(MT * nb, NT * nb, 31 + MT + NT + nb), nb)
    RY, TAU, T = run(hq, At, MT, NT, nb, ib, serial)
    RYo, TAUo, To = orc.tile_qr(At, MT, NT, nb, min(ib, 32), want_T=True)
    assert rel(RY, RYo) < TOL
    assert np.abs(TAU - TAUo).max() < TOL
    if ib <= 32:
        assert np.abs(T - To).max() / max(np.abs(To).max(), 1) < TOL
    else:
        _, _, Tw = orc.tile_qr(At, MT, NT, nb, ib, want_T=True)
        assert np.abs(T - Tw).max() / max(np.abs(Tw).max(), 1) < 1e-11


@pytest.mark.parametrize("MT,NT,nb,ib", [(2, 2, 384, 32), (3, 2, 200, 24)])
def test_orgqr_ormqr_make_q(hq, orc, MT, NT, nb, ib):
    import torch
    print(f"generic orgqr MT={MT} NT={NT} nb={nb} ib={ib}", flush=True)
    At = orc.cm_to_tile(orc.fill_uniform_cm(MT * nb, NT * nb, 5 + nb), nb)
    dA = torch.from_numpy(At.copy()).to("cuda")
    dTAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    hq.tile_qr_dev(MT, NT, nb, ib, dA, dTAU, None)
    hq.sync()
    RY, TAU = dA.cpu().numpy(), dTAU.cpu().numpy()
    dQ = torch.empty_like(dA)
    hq.tile_orgqr_dev(MT, NT, nb, ib, dA, dTAU, dQ)
    hq.sync()
    assert np.abs(dQ.cpu().numpy() - orc.dorgqr(RY, TAU, MT, NT, nb)).max() < TOL
    Ct = orc.cm_to_tile(orc.fill_uniform_cm(MT * nb, NT * nb, 99), nb)
    dC = torch.from_numpy(Ct.copy()).to("cuda")
    hq.tile_ormqr_dev(MT, NT, nb, ib, dA, dTAU, dC, NT)
    hq.sync()
    Co = orc.dormqr(RY, TAU, Ct, MT, NT, nb)
    assert rel(dC.cpu().numpy(), Co) < TOL
    dQ3 = torch.zeros_like(dA)
    hq.tile_make_q_dev(MT, NT, nb, ib, dA.clone(), dTAU, dQ3)
    hq.sync()
    Qm, _ = orc.make_Q(RY, TAU, MT, NT, nb)
    assert np.abs(dQ3.cpu().numpy() - Qm).max() < TOL


def test_edge_tiles_and_lapack(hq, orc):
    """1000 x 700 with nb = 384 (edge tiles): the oracle's factor of the
    zero-padded matrix, and LAPACK's R up to row signs."""
    import torch
    m, n, nb, ib = 1000, 700, 384, 32
    MT, NT = -(-m // nb), -(-n // nb)
    A = orc.fill_uniform_cm(m, n, 1700)
    Ap = np.zeros((MT * nb, NT * nb), order="F")
    Ap[:m, :n] = A
    At = orc.cm_to_tile(Ap, nb)
    RYo, TAUo = orc.tile_qr(At, MT, NT, nb, ib)
    dA = torch.from_numpy(At.copy()).cuda()
    dTAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    print("generic edge tiles 1000x700 nb=384", flush=True)
    hq.tile_qr_mn_dev(m, n, nb, ib, dA, dTAU)
    hq.sync()
    RY = dA.cpu().numpy()
    assert rel(RY, RYo) < TOL
    assert np.abs(dTAU.cpu().numpy() - TAUo).max() < TOL
    R = np.triu(orc.tile_to_cm(RY, MT, NT, nb))[:n, :n]
    Rl = np.linalg.qr(A, mode="r")
    s = np.sign(np.diag(R)) * np.sign(np.diag(Rl))
    assert rel(R * s[:, None], Rl) < TOL


def test_batch_with_generic_tiles(hq, orc):
    import torch
    nb, ib = 96, 16
    sizes = [(100, 70), (300, 150), (96, 96), (250, 249)]
    print("generic batch nb=96", flush=True)
    tiles = [((m + nb - 1) // nb, (n + nb - 1) // nb) for m, n in sizes]
    A = [torch.empty(MT * NT * nb * nb, dtype=torch.float64, device="cuda") for MT, NT in tiles]
    hq.batched_fill_uniform_dev(nb, [m for m, _ in sizes], [n for _, n in sizes], [60 + i for i in range(4)], A)
    A0 = [a.cpu().numpy() for a in A]
    TAU = [torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda") for MT, NT in tiles]
    for mapping in (0, 1):
        for a, a0 in zip(A, A0):
            a.copy_(torch.from_numpy(a0))
        hq.batched_tile_qr_ex_dev(0, mapping, nb, ib, [t[0] for t in tiles], [t[1] for t in tiles], A, TAU)
        hq.sync()
        for i, (MT, NT) in enumerate(tiles):
            RYo, TAUo = orc.tile_qr(A0[i], MT, NT, nb, ib)
            assert rel(A[i].cpu().numpy(), RYo) < TOL, (i, mapping)
            assert np.abs(TAU[i].cpu().numpy() - TAUo).max() < TOL


@pytest.mark.parametrize("n,nb,ib", [(8192, 512, 32), (6144, 384, 32)])
def test_large_colnorm_invariant(hq, n, nb, ib):
    """Whole chip: every column of R keeps the norm of the same column of A."""
    import torch
    MT = NT = n // nb
    print(f"generic colnorm n={n} nb={nb} ib={ib}", flush=True)
    A = torch.empty(MT * NT * nb * nb, dtype=torch.float64, device="cuda")
    hq.fill_uniform_dev(MT, NT, nb, 4343 + n, A)
    A0 = A.clone()
    TAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    hq.tile_qr_dev(MT, NT, nb, ib, A, TAU, None)
    hq.sync()
    a = A0.view(NT, MT, nb, nb)
    f = A.view(NT, MT, nb, nb)
    tj = torch.arange(NT, device="cuda").view(NT, 1, 1, 1)
    ti = torch.arange(MT, device="cuda").view(1, MT, 1, 1)
    col = torch.arange(nb, device="cuda").view(1, 1, nb, 1)
    row = torch.arange(nb, device="cuda").view(1, 1, 1, nb)
    upper = (ti * nb + row) <= (tj * nb + col)
    an = (a * a).sum(dim=(1, 3)).sqrt()
    rn = (f * f * upper).sum(dim=(1, 3)).sqrt()
    assert bool(torch.isfinite(A).all())
    assert ((an - rn).abs() / an).max().item() < 1e-12
