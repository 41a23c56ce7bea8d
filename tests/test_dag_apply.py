"""Host-side properties of the task lists a Q^T application sends to the
device scheduler (hq_dag_host_apply; no GPU): the sweep of Hybrid_dormqr.c:
64-167 as LARFB / SSRFB chains, one per update slab that holds columns, and,
on the 16- / 32-column tasks, each SSRFB below the first of its chain marked
to follow the previous one sub-block by sub-block."""
import numpy as np
import pytest

LARFB, SSRFB = 2, 3


@pytest.fixture(scope="module")
def hq():
    import hybrid3_amd
    hybrid3_amd.lib()
    return hybrid3_amd


def decode(d):
    """(claim index, pipelined?) of a dependency entry; None for padding."""
    if d == -1:
        return None
    return (-2 - d, True) if d <= -2 else (d, False)


@pytest.mark.parametrize("MT,NT,nb,ib,XNT,ncol,width", [
    (12, 3, 128, 32, 1, 64, 32), (12, 3, 128, 32, 1, 40, 16), (9, 9, 64, 16, 2, 100, 16),
    (20, 4, 256, 32, 1, 256, 32), (7, 2, 96, 24, 1, 0, 16), (12, 3, 128, 32, 1, 0, 0)])
def test_apply_task_list(hq, MT, NT, nb, ib, XNT, ncol, width):
    tasks, deps = hq.dag_host_apply(MT, NT, nb, ib, XNT=XNT, ncol=ncol, width=width, P=256)
    KT = min(MT, NT)
    uw = width if width else min(64, nb)
    cols = ncol if ncol else XNT * nb
    # one chain per update slab holding columns: LARFB(k) and SSRFB(k, r > k)
    slabs = [(j, s) for j in range(XNT) for s in range(-(-nb // uw)) if j * nb + s * uw < cols]
    assert len(tasks) == len(slabs) * sum(MT - k for k in range(KT))
    index = {}
    for i, (kind, k, r, j, s) in enumerate(tasks):
        assert kind in (LARFB, SSRFB) and (j, s) in slabs
        index[(kind, k, r, j, s)] = i
    piped = 0
    for i, (kind, k, r, j, s) in enumerate(tasks):
        got = [decode(d) for d in deps[i] if d != -1]
        # every dependency, pipelined or not, is claimed earlier: the claim
        # order is topological, so no workgroup waits on an unclaimed task
        assert all(e[0] < i for e in got)
        want = set()
        if kind == SSRFB:
            want.add(index[(LARFB, k, k, j, s)] if r == k + 1 else index[(SSRFB, k, r - 1, j, s)])
        if k > 0:
            want.add(index[(SSRFB, k - 1, k if kind == LARFB else r, j, s)])
        assert {e[0] for e in got} == want
        marks = [e for e in got if e[1]]
        if width and kind == SSRFB and r > k + 1:
            # the previous SSRFB of the chain, as dep[0]
            assert len(marks) == 1 and marks[0][0] == index[(SSRFB, k, r - 1, j, s)]
            assert decode(deps[i][0]) == marks[0]
            piped += 1
        else:
            assert not marks
    assert piped == (len(slabs) * sum(MT - k - 2 for k in range(KT) if MT - k - 2 > 0) if width else 0)


def test_apply_task_list_rejects_bad_arguments(hq):
    L = hq.lib()
    assert L.hq_dag_host_apply(4, 2, 64, 16, 1, 65, 16, 256, None, None, 0, None) == 1
    assert L.hq_dag_host_apply(4, 2, 64, 16, 1, 64, 64, 256, None, None, 0, None) == 1
    assert L.hq_dag_host_apply(4, 2, 64, 16, 0, 0, 16, 256, None, None, 0, None) == 1
