"""CPU audit of the DAG the device scheduler executes (the reference's
MEM / LIN / YTM dependency bits, Hybrid_tile_QR.c:36-40,496-601,1008-1440):
claim order is topological, every pair of tasks whose footprints conflict is
ordered, and executing the DAG in random topological orders with the
oracle's tile kernels reproduces the sequential factorisation bit for bit.
GEQRT / TSQRT appear as one panel task per ib block and one trailing task per
(block, 64-column slab); the oracle runs them as the same LAPACK-style steps
on the block's sub-matrices."""
import numpy as np
import pytest

GEQRT, TSQRT, LARFB, SSRFB, GEQRT_U, TSQRT_U = 0, 1, 2, 3, 4, 5


@pytest.fixture(scope="module")
def hq():
    import hybrid3_amd
    hybrid3_amd.lib()
    return hybrid3_amd


@pytest.mark.parametrize("MT,NT,nb,XNT", [(4, 4, 128, 0), (6, 3, 256, 0), (3, 5, 64, 0),
                                          (4, 4, 128, 4), (5, 2, 256, 3)])
def test_claim_order_is_topological_and_complete(hq, MT, NT, nb, XNT):
    tasks, deps = hq.dag_host(MT, NT, nb, 16, XNT, 64)
    n = len(tasks)
    KT, S = min(MT, NT), max(1, nb // 64)
    W = XNT if XNT > 0 else NT
    if XNT > 0:
        expect = sum(W * S + (MT - k - 1) * W * S for k in range(KT))
    else:
        ib = 16
        nblk = (nb + ib - 1) // ib
        # panel tasks of one tile: a panel per block + a trailing task per
        # slab right of the block
        per_tile = sum(1 + (S - (b * ib + ib) // 64 if b * ib + ib < nb else 0)
                       for b in range(nblk))
        expect = sum((MT - k) * per_tile + (NT - k - 1) * S + (NT - k - 1) * (MT - k - 1) * S
                     for k in range(KT))
    assert n == expect
    if XNT == 0:
        assert hq.schedule_info(MT, NT, nb, 16)[0] == n
    for i in range(n):
        for d in deps[i]:
            assert d < i
    seen = {tuple(t) for t in tasks}
    assert len(seen) == n


def footprint(t, nb, ib, apply_mode, uw=64):
    """(array, ti, tj, part, rows, cols, write) items: rows / cols are
    half-open index ranges inside the tile; part U / L / A = the entries with
    row <= col / row > col / all of that rectangle.  uw: column width of a
    LARFB / SSRFB slab (the panels' trailing tasks use 64)."""
    kind, k, r, j, s = t
    full = (0, nb)
    sw = uw if kind in (LARFB, SSRFB) else 64
    slab = (sw * s, min(nb, sw * s + sw))
    X = "X" if apply_mode else "A"
    if kind in (GEQRT, TSQRT, GEQRT_U, TSQRT_U):
        ii, j0 = j * ib, min(nb, j * ib + ib)
        blk = (ii, j0)
        right = (max(slab[0], j0), slab[1])
    if kind == GEQRT:
        return [("A", k, k, "A", (ii, nb), blk, True), ("T", k, k, "A", full, blk, True)]
    if kind == GEQRT_U:
        return [("A", k, k, "L", (ii, nb), blk, False), ("T", k, k, "A", full, blk, False),
                ("A", k, k, "A", (ii, nb), right, True)]
    if kind == TSQRT:
        return [("A", k, k, "U", blk, blk, True), ("A", r, k, "A", full, blk, True),
                ("T", r, k, "A", full, blk, True)]
    if kind == TSQRT_U:
        return [("A", r, k, "A", full, blk, False), ("T", r, k, "A", full, blk, False),
                ("A", k, k, "U", blk, right, True), ("A", r, k, "A", full, right, True)]
    if kind == LARFB:
        return [("A", k, k, "L", full, full, False), ("T", k, k, "A", full, full, False),
                (X, k, j, "A", full, slab, True)]
    return [("A", r, k, "A", full, full, False), ("T", r, k, "A", full, full, False),
            (X, k, j, "A", full, slab, True), (X, r, j, "A", full, slab, True)]


def conflicts(fa, fb):
    def meet(x, y):
        return x[0] < y[1] and y[0] < x[1]
    for a in fa:
        for b in fb:
            if a[:3] != b[:3] or not (a[6] or b[6]):
                continue
            if a[3] != "A" and b[3] != "A" and a[3] != b[3]:
                continue
            if meet(a[4], b[4]) and meet(a[5], b[5]):
                return True
    return False


@pytest.mark.parametrize("MT,NT,nb,XNT", [(3, 3, 128, 0), (4, 2, 128, 0), (2, 4, 64, 0),
                                          (3, 3, 128, 3)])
def test_every_conflicting_pair_is_ordered(hq, MT, NT, nb, XNT):
    ib = 16 if nb != 128 else 24
    tasks, deps = hq.dag_host(MT, NT, nb, ib, XNT, 8)
    n = len(tasks)
    reach = [set() for _ in range(n)]   # ancestors
    for i in range(n):
        for d in deps[i]:
            if d >= 0:
                reach[i] |= reach[d] | {d}
    fps = [footprint(t, nb, ib, XNT > 0) for t in tasks]
    for i in range(n):
        for j in range(i):
            if conflicts(fps[i], fps[j]):
                assert j in reach[i], (tasks[j], tasks[i])


@pytest.mark.parametrize("MT,NT,nb,ib,prec", [(3, 3, 128, 24, 1), (4, 3, 256, 32, 2), (2, 4, 256, 16, 1),
                                               (5, 2, 128, 32, 2)])
def test_wide_update_slabs_are_ordered(hq, MT, NT, nb, ib, prec):
    """FP32-update factorisations run LARFB / SSRFB on 128-column slabs
    beside the panels' 64-column trailing tasks: the task count matches
    the schedule, the order is topological and every conflicting pair of
    tasks is ordered."""
    tasks, deps = hq.dag_host(MT, NT, nb, ib, 0, 8, prec)
    n = len(tasks)
    assert hq.schedule_info(MT, NT, nb, ib, prec)[0] == n
    uw = min(128 if prec == 1 else 64, nb)
    upd = (tasks[:, 0] == LARFB) | (tasks[:, 0] == SSRFB)
    assert tasks[upd, 4].max() == nb // uw - 1
    KT, SU = min(MT, NT), nb // uw
    assert upd.sum() == sum((NT - k - 1) * SU * (MT - k) for k in range(KT))
    reach = [set() for _ in range(n)]
    for i in range(n):
        for d in deps[i]:
            assert d < i
            if d >= 0:
                reach[i] |= reach[d] | {d}
    fps = [footprint(t, nb, ib, False, uw) for t in tasks]
    for i in range(n):
        for j in range(i):
            if conflicts(fps[i], fps[j]):
                assert j in reach[i], (tasks[j], tasks[i])


def run_task(orc, t, A, TAU, T, MT, nb, ib):
    kind, k, r, j, s = t
    ts = nb * nb

    def tile(i, jj):
        o = (i + jj * MT) * ts
        return A[o:o + ts].reshape(nb, nb, order="F")

    def tf(i, jj):
        o = (i + jj * MT) * ib * nb
        return T[o:o + ib * nb].reshape(ib, nb, order="F")

    def tau(i, jj):
        o = (i + jj * MT) * nb
        return TAU[o:o + nb]

    cs = slice(64 * s, min(nb, 64 * s + 64))
    if kind in (GEQRT, TSQRT, GEQRT_U, TSQRT_U):
        ii, j0 = j * ib, min(nb, j * ib + ib)
        sb = j0 - ii
        rc = slice(max(64 * s, j0), min(nb, 64 * s + 64))
    if kind == GEQRT:      # dgeqr2 + dlarft of the block (rows ii..)
        a, tt, tv = orc.dgeqrt(tile(k, k)[ii:, ii:j0], sb)
        tile(k, k)[ii:, ii:j0] = a
        tf(k, k)[:sb, ii:j0] = tt
        tau(k, k)[ii:j0] = tv
    elif kind == GEQRT_U:  # dlarfb of the block on the slab's columns right of it
        tile(k, k)[ii:, rc] = orc.dormqr_lt(tile(k, k)[ii:, ii:j0], tf(k, k)[:sb, ii:j0],
                                            tile(k, k)[ii:, rc], sb, k=sb)
    elif kind == TSQRT:    # tsqr2 + T of the block of [R(k,k); A(r,k)]
        a1, a2, tt, tv = orc.dtsqrt(tile(k, k)[ii:j0, ii:j0], tile(r, k)[:, ii:j0], sb)
        tile(k, k)[ii:j0, ii:j0] = np.triu(a1) + np.tril(tile(k, k)[ii:j0, ii:j0], -1)
        tile(r, k)[:, ii:j0] = a2
        tf(r, k)[:sb, ii:j0] = tt
        tau(r, k)[ii:j0] = tv
    elif kind == TSQRT_U:  # parfb of the block on [R rows ii..j0; A(r,k)], slab columns
        a1, a2 = orc.dtsmqr_lt(tile(k, k)[ii:j0, rc], tile(r, k)[:, rc], tile(r, k)[:, ii:j0],
                               tf(r, k)[:sb, ii:j0], sb, k=sb)
        tile(k, k)[ii:j0, rc] = a1
        tile(r, k)[:, rc] = a2
    elif kind == LARFB:
        tile(k, j)[:, cs] = orc.dormqr_lt(tile(k, k), tf(k, k), tile(k, j)[:, cs], ib, k=nb)
    else:
        a1, a2 = orc.dtsmqr_lt(tile(k, j)[:, cs], tile(r, j)[:, cs], tile(r, k), tf(r, k), ib, k=nb)
        tile(k, j)[:, cs] = a1
        tile(r, j)[:, cs] = a2


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_interleavings_reproduce_sequential_result(hq, orc, seed):
    MT, NT, nb, ib = 4, 3, 128, 16
    tasks, deps = hq.dag_host(MT, NT, nb, ib, 0, 16)
    At = orc.cm_to_tile(orc.fill_uniform_cm(MT * nb, NT * nb, 40 + seed), nb)
    ref, tref, Tref = orc.tile_qr(At, MT, NT, nb, ib, want_T=True)
    A, TAU, T = At.copy(), np.zeros(MT * NT * nb), np.zeros(MT * NT * ib * nb)
    rng = np.random.default_rng(seed)
    n = len(tasks)
    done = np.zeros(n, bool)
    ready = [i for i in range(n) if (deps[i] < 0).all()]
    while ready:
        i = ready.pop(rng.integers(len(ready)))
        run_task(orc, tasks[i], A, TAU, T, MT, nb, ib)
        done[i] = True
        for q in range(n):
            if not done[q] and q not in ready and all(d < 0 or done[d] for d in deps[q]):
                ready.append(q)
    assert done.all()
    assert np.array_equal(A, ref) and np.array_equal(TAU, tref) and np.array_equal(T, Tref)


def test_trace_critical_path_matches_brute_force(hq):
    """hybrid3_amd.trace.critical_path on synthetic stamps: the longest
    dependency chain equals an independent longest-path search (memoised
    recursion over predecessors), and a task started late is reported."""
    import sys
    from hybrid3_amd import trace
    MT, NT, nb, ib, P = 3, 3, 64, 16, 4
    tasks, deps = hq.dag_host(MT, NT, nb, ib, 0, P)
    n = len(tasks)
    rng = np.random.default_rng(7)
    dur = rng.uniform(1.0, 50.0, n)
    # a valid schedule: each task starts 2 us after its last dependency ends
    us = np.zeros((n, 4))
    for i in range(n):
        d = deps[i][deps[i] >= 0]
        st = (us[d, 2].max() if len(d) else 0.0) + 2.0
        us[i] = (st - 1.0, st, st + dur[i], i % P)
    late = n - 1
    us[late, 1:3] += 40.0                       # the last task is held back
    sys.setrecursionlimit(10000)
    memo = {}

    def longest(i):
        if i not in memo:
            d = deps[i][deps[i] >= 0]
            memo[i] = (us[i, 2] - us[i, 1]) + max((longest(int(j)) for j in d), default=0.0)
        return memo[i]
    want = max(longest(i) for i in range(n))
    got = trace.critical_path(tasks, us, MT, NT, nb, ib, P)
    assert abs(got["critical_path_us"] - want) < 1e-9
    assert abs(got["work_per_workgroup_us"] - (us[:, 2] - us[:, 1]).sum() / P) < 1e-9
    d = deps[late][deps[late] >= 0]
    assert got["last_held_ready_task_us"] == us[d, 2].max()
