"""GPU parity of Q^T applied to a narrow block of right-hand sides
(hq_tile_ormqr_cols_dev: X's first ncol columns hold entries, the rest
zeros), the C argument of Hybrid_dormqr (Hybrid_dormqr.c:20-171), against the
CPU oracle's dormqr, through the C-ABI, for both update widths (16-column
tasks on the generic kernels, the factorisation's slabs)."""
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


def factor(hq, orc, MT, NT, nb, ib, seed):
    import torch
    A = orc.fill_uniform_cm(MT * nb, NT * nb, seed)
    dA = torch.from_numpy(orc.cm_to_tile(A, nb).copy()).to("cuda")
    dTAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    dT = torch.zeros(MT * NT * ib * nb, dtype=torch.float64, device="cuda")
    hq.tile_qr_dev(MT, NT, nb, ib, dA, dTAU, dT)
    hq.sync()
    return dA, dTAU, dT


def narrow_block(orc, MT, NT, nb, XNT, ncol, seed):
    """X (MT x XNT tiles, tile layout) with entries in its first ncol
    columns and zeros after them, and the same X as the first XNT tile
    columns of an MT x NT tile array (the oracle's dormqr takes RY's shape)."""
    Xc = np.zeros((MT * nb, XNT * nb), order="F")
    Xc[:, :ncol] = orc.fill_uniform_cm(MT * nb, ncol, seed)
    Xt = orc.cm_to_tile(Xc, nb)
    full = np.zeros(MT * NT * nb * nb)
    full[: Xt.size] = Xt
    return Xt, full


# (MT, NT, nb, ib): specialised tile sizes, generic ones, ib above the update
# kernels' block width and ib not a multiple of 16 (16-column tasks rebuild
# T with ib = 32)
GEOMS = [(4, 3, 64, 16), (4, 4, 128, 32), (5, 3, 256, 32), (3, 2, 96, 32), (2, 2, 512, 32), (3, 2, 128, 64),
         (3, 2, 128, 24)]


@pytest.mark.parametrize("width", [16, 32, 64])
@pytest.mark.parametrize("MT,NT,nb,ib", GEOMS)
def test_ormqr_cols_matches_oracle(hq, orc, MT, NT, nb, ib, width):
    import torch
    if width == 32 and nb > 256:
        pytest.skip("32-column updates are for nb <= 256")
    dA, dTAU, dT = factor(hq, orc, MT, NT, nb, ib, 11)
    RY, TAU = dA.cpu().numpy(), dTAU.cpu().numpy()
    cases = [(1, 1), (1, 17), (1, min(64, nb)), (1, nb), (2, nb + 5)]
    for XNT, ncol in cases:
        print(f"ormqr_cols {MT}x{NT} nb={nb} ib={ib} width={width} XNT={XNT} ncol={ncol}", flush=True)
        Xt, full = narrow_block(orc, MT, NT, nb, XNT, ncol, 100 + ncol)
        expect = orc.dormqr(RY, TAU, full, MT, NT, nb)[: Xt.size]
        for T in (None, dT):
            dX = torch.from_numpy(Xt.copy()).to("cuda")
            hq.tile_ormqr_cols_dev(MT, NT, nb, ib, dA, dTAU, dX, XNT, ncol, T=T, width=width)
            hq.sync()
            got = orc.tile_to_cm(dX.cpu().numpy(), MT, XNT, nb)
            ref = orc.tile_to_cm(expect, MT, XNT, nb)
            assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12, (XNT, ncol, T is None)
            assert np.all(got[:, ncol:] == 0.0), (XNT, ncol)


@pytest.mark.parametrize("MT,NT,nb,ib", [(6, 2, 256, 32), (4, 4, 128, 32)])
def test_ormqr_cols_widths_agree_and_auto_is_narrow(hq, orc, MT, NT, nb, ib):
    """16-column and slab-wide updates agree to rounding; width 0 up to 256
    columns is the 16-column run, bit for bit, and so is the whole-X entry on
    a one-tile-column X (its zero columns stay zero)."""
    import torch
    dA, dTAU, dT = factor(hq, orc, MT, NT, nb, ib, 12)
    Xt, _ = narrow_block(orc, MT, NT, nb, 1, 64, 13)
    out = {}
    for width in (16, 32, 64, 0):
        dX = torch.from_numpy(Xt.copy()).to("cuda")
        hq.tile_ormqr_cols_dev(MT, NT, nb, ib, dA, dTAU, dX, 1, 64, T=dT, width=width)
        hq.sync()
        out[width] = dX.cpu().numpy()
    dX = torch.from_numpy(Xt.copy()).to("cuda")
    hq.tile_ormqr_dev(MT, NT, nb, ib, dA, dTAU, dX, 1, T=dT)
    hq.sync()
    whole = dX.cpu().numpy()
    scale = np.abs(whole).max()
    assert np.array_equal(out[0], out[16])
    assert np.abs(out[16] - out[64]).max() / scale < 1e-13
    assert np.array_equal(out[16], out[32])   # a strip's arithmetic does not depend on its neighbour
    assert np.array_equal(out[16], whole)


def test_ormqr_cols_repeat_calls_reuse_schedules(hq, orc):
    """Alternating narrow widths and column counts on one context: each call
    gets its own DAG (cached by width and columns) and the same results."""
    import torch
    MT, NT, nb, ib = 5, 2, 128, 32
    dA, dTAU, dT = factor(hq, orc, MT, NT, nb, ib, 14)
    RY, TAU = dA.cpu().numpy(), dTAU.cpu().numpy()
    for ncol, width in [(40, 16), (128, 64), (40, 64), (40, 16), (3, 16), (40, 32), (3, 32)]:
        Xt, full = narrow_block(orc, MT, NT, nb, 1, ncol, 15)
        expect = orc.tile_to_cm(orc.dormqr(RY, TAU, full, MT, NT, nb)[: Xt.size], MT, 1, nb)
        dX = torch.from_numpy(Xt.copy()).to("cuda")
        hq.tile_ormqr_cols_dev(MT, NT, nb, ib, dA, dTAU, dX, 1, ncol, T=dT, width=width)
        hq.sync()
        got = orc.tile_to_cm(dX.cpu().numpy(), MT, 1, nb)
        assert np.abs(got - expect).max() / np.abs(expect).max() < 1e-12, (ncol, width)


def test_ormqr_cols_wait_bound_reports_and_recovers(hq, orc):
    """With the wait bound at one poll, the chain's waits (the scheduler's and
    the sub-block waits inside a pipelined SSRFB) time out: the launch drains,
    hq_sync reports it, and the next call with the bound restored is right."""
    import torch
    MT, NT, nb, ib = 12, 2, 128, 32
    dA, dTAU, dT = factor(hq, orc, MT, NT, nb, ib, 16)
    RY, TAU = dA.cpu().numpy(), dTAU.cpu().numpy()
    Xt, full = narrow_block(orc, MT, NT, nb, 1, 16, 17)
    expect = orc.tile_to_cm(orc.dormqr(RY, TAU, full, MT, NT, nb)[: Xt.size], MT, 1, nb)
    L = hq.lib()
    print("narrow apply wait bound: 1 poll", flush=True)
    assert L.hq_set_wait_limit(1) == 0
    try:
        dX = torch.from_numpy(Xt.copy()).to("cuda")
        hq.tile_ormqr_cols_dev(MT, NT, nb, ib, dA, dTAU, dX, 1, 16, T=dT, width=16)
        with pytest.raises(hq.HQError, match="timed out"):
            hq.sync()
    finally:
        assert L.hq_set_wait_limit(0) == 0
    dX = torch.from_numpy(Xt.copy()).to("cuda")
    hq.tile_ormqr_cols_dev(MT, NT, nb, ib, dA, dTAU, dX, 1, 16, T=dT, width=16)
    hq.sync()
    got = orc.tile_to_cm(dX.cpu().numpy(), MT, 1, nb)
    assert np.abs(got - expect).max() / np.abs(expect).max() < 1e-12
