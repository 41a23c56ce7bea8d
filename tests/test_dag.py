"""CPU audit of the DAG the device scheduler executes (the reference's
MEM / LIN / YTM dependency bits, Hybrid_tile_QR.c:36-40,496-601,1008-1440):
claim order is topological, every pair of tasks whose footprints conflict is
ordered, and executing the DAG in random topological orders with the
oracle's tile kernels reproduces the sequential factorisation bit for bit."""
import numpy as np
import pytest

GEQRT, TSQRT, LARFB, SSRFB = 0, 1, 2, 3


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
        expect = sum(1 + (MT - k - 1) + (NT - k - 1) * S + (NT - k - 1) * (MT - k - 1) * S
                     for k in range(KT))
    assert n == expect
    for i in range(n):
        for d in deps[i]:
            assert d < i
    seen = {tuple(t) for t in tasks}
    assert len(seen) == n


def footprint(t, nb, apply_mode):
    """(array, ti, tj, part, cols, write) items; part U/L/A = upper incl.
    diagonal / strictly lower / whole tile; cols = column slab range."""
    kind, k, r, j, s = t
    full = (0, nb)
    slab = (64 * s, min(nb, 64 * s + 64))
    X = "X" if apply_mode else "A"
    if kind == GEQRT:
        return [("A", k, k, "A", full, True), ("T", k, k, "A", full, True)]
    if kind == TSQRT:
        return [("A", k, k, "U", full, True), ("A", r, k, "A", full, True), ("T", r, k, "A", full, True)]
    if kind == LARFB:
        return [("A", k, k, "L", full, False), ("T", k, k, "A", full, False),
                (X, k, j, "A", slab, True)]
    return [("A", r, k, "A", full, False), ("T", r, k, "A", full, False),
            (X, k, j, "A", slab, True), (X, r, j, "A", slab, True)]


def conflicts(fa, fb):
    for a in fa:
        for b in fb:
            if a[:3] != b[:3] or not (a[4] or b[4]):
                continue
            if a[3] != "A" and b[3] != "A" and a[3] != b[3]:
                continue
            if a[4][0] < b[4][1] and b[4][0] < a[4][1] and (a[5] or b[5]):
                return True
    return False


@pytest.mark.parametrize("MT,NT,nb,XNT", [(3, 3, 128, 0), (4, 2, 128, 0), (2, 4, 64, 0),
                                          (3, 3, 128, 3)])
def test_every_conflicting_pair_is_ordered(hq, MT, NT, nb, XNT):
    tasks, deps = hq.dag_host(MT, NT, nb, 16, XNT, 8)
    n = len(tasks)
    reach = [set() for _ in range(n)]   # ancestors
    for i in range(n):
        for d in deps[i]:
            if d >= 0:
                reach[i] |= reach[d] | {d}
    fps = [footprint(t, nb, XNT > 0) for t in tasks]
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
    if kind == GEQRT:
        a, tt, tv = orc.dgeqrt(tile(k, k), ib)
        tile(k, k)[:] = a
        tf(k, k)[:] = tt
        tau(k, k)[:] = tv
    elif kind == TSQRT:
        a1, a2, tt, tv = orc.dtsqrt(tile(k, k), tile(r, k), ib)
        tile(k, k)[:] = a1
        tile(r, k)[:] = a2
        tf(r, k)[:] = tt
        tau(r, k)[:] = tv
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
