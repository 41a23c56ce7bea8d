"""The allocation-free tile QR (hq_tile_qr_noalloc_dev): its scheduler state
and task list live in the library's device array, not in hipMalloc'ed
memory.  Checked through the C-ABI against the allocating entry point
hq_tile_qr_dev on two calls in a row, the second with a smaller matrix, and
against the oracle (FP64, 1e-12 relative: the two entry points apply LARFB
with different but equivalent products, so they agree to rounding).  Also:
a repeated call on the resident geometry is bitwise reproducible, a call
captured in a hipGraph replays correctly, and a geometry whose task list
does not fit is refused before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def hq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import hybrid3_amd
    hybrid3_amd.lib()
    return hybrid3_amd


def rel(a, b):
    return (a - b).abs().max().item() / max(b.abs().max().item(), 1e-300)


def run(hq, fn, MT, NT, nb, ib, pristine):
    import torch
    A = pristine.clone()
    TAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    T = torch.zeros(MT * NT * ib * nb, dtype=torch.float64, device="cuda")
    hq.check(fn(MT, NT, nb, ib, hq._ptr(A), hq._ptr(TAU), hq._ptr(T), hq._stream()))
    return A, TAU, T


def test_noalloc_matches_allocating_entry_point(hq, orc):
    import torch
    L = hq.lib()
    # the second call smaller; the third larger than the second again (its flag
    # range then covers the second geometry's task list)
    for MT, NT, nb, ib, seed in [(12, 10, 128, 32, 5), (4, 3, 128, 16, 6), (7, 6, 128, 32, 7)]:
        print(f"noalloc {MT}x{NT} nb={nb} ib={ib}", flush=True)
        pristine = torch.empty(MT * NT * nb * nb, dtype=torch.float64, device="cuda")
        hq.fill_uniform_dev(MT, NT, nb, seed, pristine)
        A1, TAU1, T1 = run(hq, L.hq_tile_qr_dev, MT, NT, nb, ib, pristine)
        hq.sync()
        A2, TAU2, T2 = run(hq, L.hq_tile_qr_noalloc_dev, MT, NT, nb, ib, pristine)
        hq.check(L.hq_tile_qr_noalloc_check(hq._stream()))
        assert rel(A2, A1) < TOL and rel(TAU2, TAU1) < TOL and rel(T2, T1) < TOL
        if MT * NT <= 12:
            RYo, TAUo = orc.tile_qr(pristine.cpu().numpy(), MT, NT, nb, ib)
            assert rel(A2.cpu(), torch.from_numpy(RYo)) < TOL
        # the resident geometry again: no copy, a new epoch, the same bits
        A3, TAU3, T3 = run(hq, L.hq_tile_qr_noalloc_dev, MT, NT, nb, ib, pristine)
        hq.check(L.hq_tile_qr_noalloc_check(hq._stream()))
        assert torch.equal(A3, A2) and torch.equal(TAU3, TAU2) and torch.equal(T3, T2)


def test_noalloc_graph_replays(hq):
    import torch
    L = hq.lib()
    MT, NT, nb, ib = 6, 6, 128, 32
    pristine = torch.empty(MT * NT * nb * nb, dtype=torch.float64, device="cuda")
    hq.fill_uniform_dev(MT, NT, nb, 11, pristine)
    ref, _, _ = run(hq, L.hq_tile_qr_noalloc_dev, MT, NT, nb, ib, pristine)   # makes the geometry resident
    hq.check(L.hq_tile_qr_noalloc_check(hq._stream()))
    ref_alloc, _, _ = run(hq, L.hq_tile_qr_dev, MT, NT, nb, ib, pristine)
    hq.sync()
    assert rel(ref, ref_alloc) < TOL
    A = pristine.clone()
    TAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    T = torch.zeros(MT * NT * ib * nb, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    print("noalloc graph capture", flush=True)
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            hq.check(L.hq_tile_qr_noalloc_dev(MT, NT, nb, ib, hq._ptr(A), hq._ptr(TAU), hq._ptr(T),
                                              hq._stream()))
    for rep in range(3):
        A.copy_(pristine)
        junk = torch.randn(1000, 1000, dtype=torch.float64, device="cuda") @ torch.randn(
            1000, 1000, dtype=torch.float64, device="cuda")   # other eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        hq.check(L.hq_tile_qr_noalloc_check(hq._stream()))
        assert torch.equal(A, ref), rep
        del junk


def test_noalloc_refuses_a_schedule_beyond_its_array(hq):
    L = hq.lib()
    MT = NT = 140
    nb, ib = 64, 16
    n, _ = hq.schedule_info(MT, NT, nb, ib)
    assert n * 52 > 32 << 20
    rc = L.hq_tile_qr_noalloc_dev(MT, NT, nb, ib, 1, 1, 1, None)
    code, msg = hq.last_error()
    assert rc == 1 and code == 1 and "at most 32 MiB" in msg
    assert L.hq_tile_qr_noalloc_dev(4, 4, 128, 48, 1, 1, 1, None) == 1   # ib > 32
    assert L.hq_tile_qr_noalloc_dev(4, 4, 128, 32, 1, 1, None, None) == 1  # T required
