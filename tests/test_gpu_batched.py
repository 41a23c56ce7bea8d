"""Batched uneven QR (no reference counterpart): many independent matrices
factored by ONE persistent launch of the DAG scheduler, through the C-ABI.

Pinned by:
* the oracle's flat-tree tile QR of each zero-padded matrix (1e-12 of the
  largest entry, same as the single-matrix parity tests);
* bitwise equality with one hq_tile_qr_dev launch per matrix (same kernels,
  same per-tile update order);
* LAPACK's QR of the UNPADDED m x n matrix (numpy.linalg.qr): R unique up to
  the signs of its rows, so the padding is exact;
* Q orthogonal: every column of R keeps the norm of the same column of A
  (whole-chip batch, sizes the oracle does not finish quickly)."""
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


def make_batch(hq, sizes, nb, seed0):
    import torch
    tiles = [((m + nb - 1) // nb, (n + nb - 1) // nb) for m, n in sizes]
    A = [torch.empty(MT * NT * nb * nb, dtype=torch.float64, device="cuda") for MT, NT in tiles]
    hq.batched_fill_uniform_dev(nb, [m for m, _ in sizes], [n for _, n in sizes],
                                [seed0 + i for i in range(len(sizes))], A)
    return tiles, A


SIZES = [(100, 70), (64, 64), (200, 33), (130, 129), (64, 32), (300, 150), (257, 200), (65, 65)]


@pytest.mark.parametrize("nb,ib", [(64, 16), (64, 12), (128, 32)])
def test_batched_matches_oracle_and_single(hq, orc, nb, ib):
    import torch
    print(f"batched nb={nb} ib={ib}", flush=True)
    tiles, A = make_batch(hq, SIZES, nb, 900)
    A0 = [a.clone() for a in A]
    TAU = [torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda") for MT, NT in tiles]
    hq.batched_tile_qr_dev(nb, ib, [t[0] for t in tiles], [t[1] for t in tiles], A, TAU)
    hq.sync()
    for i, ((m, n), (MT, NT)) in enumerate(zip(SIZES, tiles)):
        # the oracle on the zero-padded matrix built by the same generator
        cm = np.zeros((MT * nb, NT * nb), order="F")
        cm[:m, :n] = orc.fill_uniform_cm(m, n, 900 + i)
        At = orc.cm_to_tile(cm, nb)
        assert np.array_equal(A0[i].cpu().numpy(), At), "device fill differs from the column-major fill"
        RYo, TAUo = orc.tile_qr(At, MT, NT, nb, ib)
        RY = A[i].cpu().numpy()
        assert np.abs(RY - RYo).max() / np.abs(RYo).max() < 1e-12, (m, n)
        assert np.abs(TAU[i].cpu().numpy() - TAUo).max() < 1e-12, (m, n)
        # one launch per matrix gives the same bits
        a1 = A0[i].clone()
        t1 = torch.zeros_like(TAU[i])
        hq.tile_qr_dev(MT, NT, nb, ib, a1, t1, None)
        hq.sync()
        assert torch.equal(a1, A[i]) and torch.equal(t1, TAU[i]), (m, n)


@pytest.mark.parametrize("nb,ib", [(64, 16), (128, 32), (256, 32)])
def test_split_panel_mapping_matches_oracle(hq, orc, nb, ib):
    """Panel mapping 1 (one task per ib block plus per-slab trailing tasks)
    gives the oracle's factor as mapping 0 does."""
    import torch
    print(f"batched split panels nb={nb} ib={ib}", flush=True)
    sizes = SIZES + [(600, 300), (513, 513)]
    tiles, A = make_batch(hq, sizes, nb, 930)
    A0 = [a.clone() for a in A]
    TAU = [torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda") for MT, NT in tiles]
    hq.batched_tile_qr_ex_dev(0, 1, nb, ib, [t[0] for t in tiles], [t[1] for t in tiles], A, TAU)
    hq.sync()
    A1 = [a.clone() for a in A0]
    TAU1 = [torch.zeros_like(t) for t in TAU]
    hq.batched_tile_qr_ex_dev(0, 0, nb, ib, [t[0] for t in tiles], [t[1] for t in tiles], A1, TAU1)
    hq.sync()
    same = all(torch.equal(a, b) for a, b in zip(A, A1))
    print(f"  bitwise equal to whole-tile panels: {same}", flush=True)
    for i, ((m, n), (MT, NT)) in enumerate(zip(sizes, tiles)):
        if m * n > 300 * 300:
            continue   # the oracle's time; the column-norm test below covers large ones
        RYo, TAUo = orc.tile_qr(A0[i].cpu().numpy(), MT, NT, nb, ib)
        assert np.abs(A[i].cpu().numpy() - RYo).max() / np.abs(RYo).max() < 1e-12, (m, n)
        assert np.abs(TAU[i].cpu().numpy() - TAUo).max() < 1e-12, (m, n)
    for a, b in zip(A, A1):
        assert (a - b).abs().max().item() <= 1e-12 * b.abs().max().item()


@pytest.mark.parametrize("nb,ib,prec", [(64, 16, 0), (128, 32, 0), (256, 32, 0), (96, 16, 0), (256, 32, 1),
                                         (128, 32, 2)])
def test_padding_skip_is_bitwise(hq, nb, ib, prec):
    """With each matrix's column count given, the tasks wholly inside the
    zero padding of its last tile column only write zeros: A, TAU and T are
    bitwise those of the full run, for both panel mappings."""
    import torch
    sizes = SIZES + [(600, 300), (513, 40), (300, 257)]
    print(f"batched padding skip nb={nb} ib={ib} prec={prec}", flush=True)
    tiles, A0 = make_batch(hq, sizes, nb, 970)
    MTs, NTs = [t[0] for t in tiles], [t[1] for t in tiles]
    for mapping in (0, 1):
        out = []
        for ncol in (None, [n for _, n in sizes]):
            A = [a.clone() for a in A0]
            TAU = [torch.full((MT * NT * nb,), 5.0, dtype=torch.float64, device="cuda") for MT, NT in tiles]
            T = [torch.full((MT * NT * ib * nb,), 7.0, dtype=torch.float64, device="cuda") for MT, NT in tiles]
            hq.batched_tile_qr_ex_dev(prec, mapping, nb, ib, MTs, NTs, A, TAU, T, ncol=ncol)
            hq.sync()
            out.append((A, TAU, T))
        for name, x, y in zip(("A", "TAU", "T"), out[0], out[1]):
            assert all(torch.equal(a, b) for a, b in zip(x, y)), (name, mapping)
    with pytest.raises(hq.HQError):   # a column count that does not end in the last tile column
        hq.batched_tile_qr_ex_dev(0, 0, nb, ib, MTs[:1], NTs[:1], A0[:1], TAU[:1], ncol=[NTs[0] * nb + 1])


def test_fp32_batch_is_the_rounded_fp64_factor(hq, orc):
    """FP32 matrices: widened to the FP64 tile layout, factored, rounded back.
    The stored FP32 factor is the oracle's FP64 factor of the FP32 inputs
    rounded to FP32 (to one unit in the last place, the FP64 results agree
    to 1e-12), and the FP32 factor is a QR of the input to FP32 accuracy."""
    import torch
    from hybrid3_amd import batched as B
    sizes = SIZES + [(64, 33), (191, 64)]
    plan = [(64, 16, i % 2) if i % 3 else (128, 32, i % 2) for i in range(len(sizes))]
    print("batched FP32 matrices", flush=True)
    bt = B.Batch(sizes, plan, seed0=960, fp32=True)
    inputs = bt.pristine.cpu().numpy()
    bt.factor()
    sync_out = bt.C32.cpu().numpy()
    for i, ((m, n), (MT, NT), (nb, ib, _)) in enumerate(zip(bt.sizes, bt.tiles, bt.plan)):
        lo = int(bt.c_off[i])
        a32 = inputs[lo:lo + m * n].reshape(n, m).T
        # the FP32 inputs are the generator's values rounded
        assert np.array_equal(a32, orc.fill_uniform_cm(m, n, 960 + i).astype(np.float32)), (m, n)
        cm = np.zeros((MT * nb, NT * nb), order="F")
        cm[:m, :n] = a32.astype(np.float64)
        RYo, _ = orc.tile_qr(orc.cm_to_tile(cm, nb), MT, NT, nb, ib)
        want = orc.tile_to_cm(RYo, MT, NT, nb)[:m, :n]
        got = sync_out[lo:lo + m * n].reshape(n, m).T.astype(np.float64)
        ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(got - want) <= ulp + 1e-12 * np.abs(want).max()), (m, n)
    res = bt.residuals(range(len(sizes)))
    assert max(r for r, _ in res) < 1e-5 and max(o for _, o in res) < 1e-5, res
    assert bt.colnorm_error() < 1e-6


def test_batch_residual_and_orthogonality(hq):
    """Per-matrix ||A - QR|| / ||A|| and ||I - Q^T Q|| of an FP64 batch
    (dormqr / dorgqr of the engine on the batch's reflectors)."""
    from hybrid3_amd import batched as B
    sizes = B.uneven_sizes(40, seed=12, m_range=(64, 900))
    plan = [(64, 16, 0) if m < 256 else (128, 32, 1) for m, _ in sizes]
    print("batched residuals", flush=True)
    bt = B.Batch(sizes, plan, seed0=40)
    bt.factor()
    res = bt.residuals(range(len(sizes)))
    assert max(r for r, _ in res) < 1e-13 and max(o for _, o in res) < 1e-12, res


def test_padding_is_exact_vs_lapack(hq, orc):
    import torch
    nb, ib = 64, 16
    sizes = [(150, 90), (97, 97), (300, 41)]
    print("batched padding vs numpy.linalg.qr", flush=True)
    tiles, A = make_batch(hq, sizes, nb, 4100)
    TAU = [torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda") for MT, NT in tiles]
    hq.batched_tile_qr_dev(nb, ib, [t[0] for t in tiles], [t[1] for t in tiles], A, TAU)
    hq.sync()
    for i, ((m, n), (MT, NT)) in enumerate(zip(sizes, tiles)):
        a = orc.fill_uniform_cm(m, n, 4100 + i)
        R_lapack = np.linalg.qr(a, mode="r")[:n, :n]
        R = np.triu(orc.tile_to_cm(A[i].cpu().numpy(), MT, NT, nb))[:n, :n]
        d = np.sign(np.diag(R)) * np.sign(np.diag(R_lapack))
        assert np.abs(R - d[:, None] * R_lapack).max() / np.abs(R_lapack).max() < 1e-12, (m, n)


def test_layout_round_trip(hq):
    import torch
    nb = 64
    sizes = [(70, 33), (128, 64), (1, 1), (200, 130)]
    print("batched layout round trip", flush=True)
    rng = np.random.default_rng(5)
    cms = [torch.from_numpy(np.asfortranarray(rng.uniform(-1, 1, (m + 3, n))).ravel(order="F").copy()).cuda()
           for m, n in sizes]
    ld = [m + 3 for m, _ in sizes]
    tiles = [((m + nb - 1) // nb, (n + nb - 1) // nb) for m, n in sizes]
    A = [torch.full((MT * NT * nb * nb,), 7.0, dtype=torch.float64, device="cuda") for MT, NT in tiles]
    hq.cm_to_tile_dev(nb, [m for m, _ in sizes], [n for _, n in sizes], cms, ld, A)
    back = [torch.full_like(c, -3.0) for c in cms]
    hq.tile_to_cm_dev(nb, [m for m, _ in sizes], [n for _, n in sizes], A, back, ld)
    hq.sync()
    for (m, n), (MT, NT), c, b, a in zip(sizes, tiles, cms, back, A):
        cm = c.cpu().numpy().reshape(n, m + 3).T
        bm = b.cpu().numpy().reshape(n, m + 3).T
        assert np.array_equal(bm[:m, :], cm[:m, :])      # entries copied back exactly
        assert np.all(bm[m:, :] == -3.0)                  # rows past m untouched
        full = a.cpu().numpy().reshape(NT, MT, nb, nb)    # [tile col][tile row][col][row]
        dense = full.transpose(1, 3, 0, 2).reshape(MT * nb, NT * nb)
        assert np.array_equal(dense[:m, :n], cm[:m, :])
        assert np.all(dense[m:, :] == 0.0) and np.all(dense[:, n:] == 0.0)


def test_large_batch_colnorm_invariant(hq):
    import torch
    nb, ib = 64, 16
    rng = np.random.default_rng(77)
    count = 600
    ms = rng.integers(64, 513, count)
    ns = np.array([rng.integers(32, m + 1) for m in ms])
    sizes = list(zip(ms.tolist(), ns.tolist()))
    print(f"batched colnorm count={count}", flush=True)
    tiles, A = make_batch(hq, sizes, nb, 7000)
    A0 = [a.clone() for a in A]
    TAU = [torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda") for MT, NT in tiles]
    hq.batched_tile_qr_dev(nb, ib, [t[0] for t in tiles], [t[1] for t in tiles], A, TAU)
    hq.sync()
    worst = 0.0
    for (MT, NT), a, f in zip(tiles, A0, A):
        a = a.view(NT, MT, nb, nb)
        f = f.view(NT, MT, nb, nb)
        tj = torch.arange(NT, device="cuda").view(NT, 1, 1, 1)
        ti = torch.arange(MT, device="cuda").view(1, MT, 1, 1)
        col = torch.arange(nb, device="cuda").view(1, 1, nb, 1)
        row = torch.arange(nb, device="cuda").view(1, 1, 1, nb)
        upper = (ti * nb + row) <= (tj * nb + col)
        an = (a * a).sum(dim=(1, 3)).sqrt()
        rn = (f * f * upper).sum(dim=(1, 3)).sqrt()
        live = an > 0          # zero padding columns have zero norm on both sides
        assert torch.all(rn[~live] == 0)
        worst = max(worst, ((an[live] - rn[live]).abs() / an[live]).max().item())
    assert worst < 1e-12


def test_autotuner_plan_and_batch(hq):
    """The tuner returns one measured (nb, ib) per matrix from the candidate
    set, and the batch it plans factors correctly (launch per group)."""
    from hybrid3_amd import batched as B
    sizes = B.uneven_sizes(90, seed=11, m_range=(64, 700))
    print("autotune 90 matrices", flush=True)
    plan, report = B.autotune(sizes, sample=24, reps=1)
    assert len(plan) == len(sizes)
    assert all(tuple(p) in B.CANDIDATES for p in plan)
    assert report["chosen"] in report["plans"]
    bt = B.Batch(sizes, plan, seed0=3)
    B.time_factor(bt, reps=1)
    assert bt.colnorm_error() < 1e-12
    # the JSON table of the report gives back the same plan
    table = B.table_from_report(report)
    assert B.plan_from_table(sizes, table) == [B._key(p) for p in plan]


def test_plan_runs_and_graph_replays(hq):
    """A batch prepared once (hq_batched_plan_dev) is run eagerly, then from a
    captured graph replayed three times with other eager batched work and
    synchronisations between the replays: every run gives the bits of the
    one-launch hq_batched_tile_qr_dev call (same kernels, same per-tile update
    order)."""
    import torch
    nb, ib = 128, 32
    print("batched plan + graph replays", flush=True)
    tiles, A = make_batch(hq, SIZES, nb, 950)
    MTs, NTs = [t[0] for t in tiles], [t[1] for t in tiles]
    A0 = [a.clone() for a in A]
    TAU = [torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda") for MT, NT in tiles]
    Aref = [a.clone() for a in A0]
    TAUref = [torch.zeros_like(t) for t in TAU]
    hq.batched_tile_qr_dev(nb, ib, MTs, NTs, Aref, TAUref)
    hq.sync()
    # other matrices factored eagerly between the replays
    otiles, other = make_batch(hq, SIZES[:4], nb, 960)
    other0 = [a.clone() for a in other]
    otau = [torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda") for MT, NT in otiles]
    plan = hq.batched_plan_dev(nb, ib, MTs, NTs, A, TAU)
    try:
        hq.batched_run_dev(plan)
        hq.batched_plan_check(plan)
        assert all(torch.equal(a, r) for a, r in zip(A, Aref))
        assert all(torch.equal(t, r) for t, r in zip(TAU, TAUref))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            hq.batched_run_dev(plan)
        for rep in range(3):
            for a, a0 in zip(A, A0):
                a.copy_(a0)
            for t in TAU:
                t.zero_()
            for a, a0 in zip(other, other0):
                a.copy_(a0)
            hq.batched_tile_qr_dev(nb, ib, [t[0] for t in otiles], [t[1] for t in otiles], other, otau)
            hq.sync()
            g.replay()
            hq.batched_plan_check(plan)
            torch.cuda.synchronize()
            assert all(torch.equal(a, r) for a, r in zip(A, Aref)), rep
            assert all(torch.equal(t, r) for t, r in zip(TAU, TAUref)), rep
    finally:
        hq.batched_plan_free(plan)


def test_plan_errors(hq):
    """Unknown or released plans are rejected with an argument error; a plan
    rejects invalid geometry at creation (nothing allocated)."""
    import torch
    nb, ib = 64, 16
    tiles, A = make_batch(hq, [(64, 64)], nb, 5)
    TAU = [torch.zeros(nb, dtype=torch.float64, device="cuda")]
    with pytest.raises(hq.HQError):
        hq.batched_run_dev(987654)
    with pytest.raises(hq.HQError):
        hq.batched_plan_dev(nb, 33, [1], [1], A, TAU)   # ib > 32
    plan = hq.batched_plan_dev(nb, ib, [1], [1], A, TAU)
    hq.batched_run_dev(plan)
    hq.batched_plan_check(plan)
    hq.batched_plan_free(plan)
    for fn in (lambda: hq.batched_run_dev(plan), lambda: hq.batched_plan_check(plan),
               lambda: hq.batched_plan_free(plan)):
        with pytest.raises(hq.HQError):
            fn()
    assert hq.lib().hq_last_error() == 1   # HQ_ERR_ARG
