"""Mixed-precision tile QR (no reference counterpart): FP64 panels (GEQRT /
TSQRT), trailing updates (LARFB / SSRFB) either on FP32 MFMA with FP32
accumulation ("fp32") or on FP16 MFMA over FP16 hi + lo splits with FP32
accumulation ("split"), through the C-ABI.

Pinned by the FP64 oracle: R / V / TAU agree to FP32 level (a few 1e-6 of the
largest entry) and differ from the FP64 result by more than FP64 rounding,
which shows a reduced-precision path ran; the split result also differs from
the FP32 one, which shows it is the split path.  A matrix whose columns span
10^-30 .. 10^30 checks the per-column power-of-two scaling of the split path
(without it FP16 would flush the small columns to zero and overflow the large
ones).  At whole-chip sizes, Q's orthogonality fixes the column norms of R to
FP32 level."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import hybrid3_amd
    hybrid3_amd.lib()
    return hybrid3_amd


def factor(hq, prec, MT, NT, nb, ib, At):
    import torch
    dA = torch.from_numpy(At.copy()).cuda()
    dTAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    fn = hq.tile_qr_split_dev if prec == "split" else hq.tile_qr_mixed_dev
    fn(MT, NT, nb, ib, dA, dTAU, None)
    hq.sync()
    return dA.cpu().numpy(), dTAU.cpu().numpy()


@pytest.mark.parametrize("prec", ["fp32", "split"])
@pytest.mark.parametrize("MT,NT,nb,ib", [(3, 3, 128, 32), (4, 4, 256, 32), (3, 2, 256, 20),
                                         (4, 3, 128, 24), (2, 3, 128, 16)])
def test_mixed_matches_fp64_oracle_to_fp32_level(hq, orc, MT, NT, nb, ib, prec):
    print(f"mixed {prec} MT={MT} NT={NT} nb={nb} ib={ib}", flush=True)
    A = orc.fill_uniform_cm(MT * nb, NT * nb, 31 + MT + NT)
    At = orc.cm_to_tile(A, nb)
    RY, TAU = factor(hq, prec, MT, NT, nb, ib, At)
    RYo, TAUo = orc.tile_qr(At, MT, NT, nb, ib)
    err = np.abs(RY - RYo).max() / np.abs(RYo).max()
    assert err < 2e-5, err
    assert err > 1e-12, "no reduced-precision rounding visible: the update path did not run"
    if prec == "split":
        RY32, _ = factor(hq, "fp32", MT, NT, nb, ib, At)
        assert np.abs(RY - RY32).max() / np.abs(RYo).max() > 1e-12, "split result equals the FP32 path's"
    assert np.abs(TAU - TAUo).max() < 2e-4
    # the first panel (tile column 0) is computed before any update: FP64 exact
    R0 = orc.tile_to_cm(RY, MT, NT, nb)[:, :nb]
    R0o = orc.tile_to_cm(RYo, MT, NT, nb)[:, :nb]
    assert np.abs(R0 - R0o).max() / np.abs(R0o).max() < 1e-12


@pytest.mark.parametrize("nb", [128, 256])
def test_split_graded_columns(hq, orc, nb):
    """Columns scaled by 10^u, u uniform in [-30, 30]: R = R(A0) diag(10^u) exactly,
    so each column of R must match the oracle to FP32 level relative to its norm."""
    MT, NT, ib = 3, 3, 32
    print(f"split graded columns nb={nb}", flush=True)
    A = orc.fill_uniform_cm(MT * nb, NT * nb, 77)
    d = 10.0 ** np.random.default_rng(78).uniform(-30, 30, NT * nb)
    A = np.asfortranarray(A * d[None, :])
    At = orc.cm_to_tile(A, nb)
    RY, _ = factor(hq, "split", MT, NT, nb, ib, At)
    RYo, _ = orc.tile_qr(At, MT, NT, nb, ib)
    R = np.triu(orc.tile_to_cm(RY, MT, NT, nb))
    Ro = np.triu(orc.tile_to_cm(RYo, MT, NT, nb))
    assert np.all(np.isfinite(R))
    colerr = np.abs(R - Ro).max(axis=0) / np.linalg.norm(A, axis=0)
    assert colerr.max() < 2e-5, colerr.max()


@pytest.mark.parametrize("prec,n", [("fp32", 4096), ("split", 4096), ("fp32", 16384), ("split", 16384)])
def test_mixed_large_colnorm(hq, prec, n):
    """Whole-chip runs (16384: 85k update tasks in flight on every CU, the
    pipelined split update's copies two passes deep) keep R's column norms
    equal to A's to FP32 level."""
    import torch
    nb, ib = 256, 32
    MT = NT = n // nb
    print(f"mixed {prec} colnorm {n}", flush=True)
    A = torch.empty(MT * NT * nb * nb, dtype=torch.float64, device="cuda")
    hq.fill_uniform_dev(MT, NT, nb, 99, A)
    A0 = A.clone()
    TAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    (hq.tile_qr_split_dev if prec == "split" else hq.tile_qr_mixed_dev)(MT, NT, nb, ib, A, TAU, None)
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
    rel = ((an - rn).abs() / an).max().item()
    assert rel < 1e-4, rel
