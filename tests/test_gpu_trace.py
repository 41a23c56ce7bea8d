"""The device scheduler's own timeline (the counterpart of the reference's
CRAYJ_TIMING / CRAYJ_TIMELINE, Hybrid_tile_QR.c:53-74) as a check of its
dependency handling: a traced factorisation gives bitwise the results of an
untraced one, runs the host DAG's tasks in its claim order, and starts no
task before every task it depends on (the reference's MEM / LIN / YTM bits,
Hybrid_tile_QR.c:496-601) has ended."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# A task's end is stamped just after its flag is published, so a successor
# that sees the flag at once may stamp its start slightly earlier; a real
# ordering error shows up as most of a task's duration (20-80 us here).
SKEW_US = 2.0


@pytest.fixture(scope="module")
def hq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import hybrid3_amd
    hybrid3_amd.lib()
    return hybrid3_amd


@pytest.mark.parametrize("MT,NT,nb,ib", [(6, 5, 128, 32), (5, 5, 256, 32), (8, 3, 64, 16)])
def test_traced_run_follows_the_dag(hq, MT, NT, nb, ib):
    import torch

    from hybrid3_amd import trace
    print(f"traced tile QR MT={MT} NT={NT} nb={nb} ib={ib}", flush=True)
    A = torch.empty(MT * NT * nb * nb, dtype=torch.float64, device="cuda")
    TAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    hq.fill_uniform_dev(MT, NT, nb, 11, A)
    A0 = A.clone()
    hq.tile_qr_dev(MT, NT, nb, ib, A, TAU)
    hq.sync()
    ref_A, ref_TAU = A.cpu(), TAU.cpu()
    A.copy_(A0)
    TAU.zero_()
    tasks, us = trace.run_traced(MT, NT, nb, ib, A, TAU)
    assert torch.equal(A.cpu(), ref_A) and torch.equal(TAU.cpu(), ref_TAU)

    ntasks, grid = hq.schedule_info(MT, NT, nb, ib)
    t2, deps = hq.dag_host(MT, NT, nb, ib, 0, grid)
    assert len(tasks) == ntasks and (t2 == tasks).all()
    assert (us[:, 1] >= us[:, 0]).all() and (us[:, 2] > us[:, 1]).all()
    assert ((us[:, 3] >= 0) & (us[:, 3] < grid)).all()
    worst = np.inf
    for i, d in enumerate(deps):
        d = d[d >= 0]
        if len(d):
            worst = min(worst, float((us[i, 1] - us[d, 2]).min()))
    assert worst > -SKEW_US, f"a task started {-worst:.2f} us before a dependency ended"
