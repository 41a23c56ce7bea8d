"""GPU parity of the four PLASMA-convention tile kernels (GEQRT, TSQRT, LARFB,
SSRFB as called at Hybrid_tile_QR.c:985,1135,1276,1356) against the CPU
oracle, through the C-ABI (hybrid3_dev.h).  FP64 throughout; tolerance
1e-12 relative to the largest entry (same algorithm, different summation
order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12
CASES = [(64, 16), (64, 12), (128, 32), (128, 8), (128, 5), (256, 32), (256, 16), (256, 24), (256, 20)]


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


def factors(orc, nb, ib, seed):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (nb, nb))
    Ao, T, tau = orc.dgeqrt(A, ib)
    return A, Ao, T, tau


@pytest.mark.parametrize("nb,ib", CASES)
def test_geqrt(hq, orc, nb, ib):
    import torch
    A, Ao, To, tauo = factors(orc, nb, ib, nb + ib)
    dA, dT = dev(A), torch.zeros(ib * nb, dtype=torch.float64, device="cuda")
    dtau = torch.zeros(nb, dtype=torch.float64, device="cuda")
    print(f"geqrt nb={nb} ib={ib}", flush=True)
    hq.dgeqrt_dev(nb, ib, dA, dT, dtau)
    torch.cuda.synchronize()
    assert rel(host(dA, (nb, nb)), Ao) < TOL
    assert rel(host(dT, (ib, nb)), To) < TOL
    assert np.abs(dtau.cpu().numpy() - tauo).max() < TOL


@pytest.mark.parametrize("nb,ib", CASES)
def test_tsqrt(hq, orc, nb, ib):
    import torch
    rng = np.random.default_rng(2 * nb + ib)
    A1 = np.triu(rng.uniform(-1, 1, (nb, nb))) + np.tril(rng.uniform(-1, 1, (nb, nb)), -1)
    A2 = rng.uniform(-1, 1, (nb, nb))
    a1o, a2o, To, tauo = orc.dtsqrt(A1, A2, ib)
    dA1, dA2 = dev(A1), dev(A2)
    dT = torch.zeros(ib * nb, dtype=torch.float64, device="cuda")
    dtau = torch.zeros(nb, dtype=torch.float64, device="cuda")
    print(f"tsqrt nb={nb} ib={ib}", flush=True)
    hq.dtsqrt_dev(nb, ib, dA1, dA2, dT, dtau)
    torch.cuda.synchronize()
    g1 = host(dA1, (nb, nb))
    assert rel(g1, a1o) < TOL
    # the strictly lower part of A1 (GEQRT's reflectors) is not touched
    assert np.array_equal(np.tril(g1, -1), np.tril(A1, -1))
    assert rel(host(dA2, (nb, nb)), a2o) < TOL
    assert rel(host(dT, (ib, nb)), To) < TOL
    assert np.abs(dtau.cpu().numpy() - tauo).max() < TOL


@pytest.mark.parametrize("nb,ib", CASES)
def test_larfb(hq, orc, nb, ib):
    import torch
    A, V, T, tau = factors(orc, nb, ib, 3 * nb + ib)
    Cm = np.random.default_rng(5).uniform(-1, 1, (nb, nb))
    Co = orc.dormqr_lt(V, T, Cm, ib, k=nb)
    dC = dev(Cm)
    print(f"larfb nb={nb} ib={ib}", flush=True)
    hq.dormqr_dev(nb, ib, dev(V), dev(T), dC)
    torch.cuda.synchronize()
    assert rel(host(dC, (nb, nb)), Co) < TOL


@pytest.mark.parametrize("nb,ib", CASES)
def test_ssrfb(hq, orc, nb, ib):
    import torch
    rng = np.random.default_rng(4 * nb + ib)
    R1 = np.triu(rng.uniform(-1, 1, (nb, nb)))
    _, V2, T, tau = orc.dtsqrt(R1, rng.uniform(-1, 1, (nb, nb)), ib)
    B1 = rng.uniform(-1, 1, (nb, nb))
    B2 = rng.uniform(-1, 1, (nb, nb))
    o1, o2 = orc.dtsmqr_lt(B1, B2, V2, T, ib, k=nb)
    d1, d2 = dev(B1), dev(B2)
    print(f"ssrfb nb={nb} ib={ib}", flush=True)
    hq.dtsmqr_dev(nb, ib, d1, d2, dev(V2), dev(T))
    torch.cuda.synchronize()
    assert rel(host(d1, (nb, nb)), o1) < TOL
    assert rel(host(d2, (nb, nb)), o2) < TOL


def test_fill_uniform_bit_exact(hq, orc):
    import torch
    MT, NT, nb = 3, 2, 64
    d = torch.empty(MT * NT * nb * nb, dtype=torch.float64, device="cuda")
    hq.fill_uniform_dev(MT, NT, nb, 1234, d)
    torch.cuda.synchronize()
    ref = orc.cm_to_tile(orc.fill_uniform_cm(MT * nb, NT * nb, 1234), nb)
    assert np.array_equal(d.cpu().numpy(), ref)
