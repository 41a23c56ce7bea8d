"""Per-task timeline of the on-device DAG scheduler.

Counterpart of the reference's CRAYJ_TIMING (per-kernel elapsed time and call
counts, Hybrid_tile_QR.c:53-63,1465-1588) and CRAYJ_TIMELINE (per-thread event
files, Hybrid_tile_QR.c:64-74): one factorisation is run with device
wall-clock stamps (claim / start / end per task, and the workgroup that ran
it), then summarised per task kind and optionally written as a Chrome trace.

    python -m hybrid3_amd.trace --n 16384 --nb 256 --ib 32 [--chrome out.json] [--critical-path]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json

import numpy as np

KINDS = ["GEQRT", "TSQRT", "LARFB", "SSRFB", "GEQRT_U", "TSQRT_U"]


def run_traced(MT, NT, nb, ib, A, TAU, T=None, stream=None, prec=0):
    """Factor A (device tensor) with timestamps; returns (tasks[n,5], trace[n,4] in us).
    prec: trailing updates in FP64 (0), FP32 (1) or FP16 hi+lo splits (2)."""
    import torch

    import hybrid3_amd as hq
    L = hq.lib()
    ntasks, _ = hq.schedule_info(MT, NT, nb, ib, prec)
    tasks = np.zeros(5 * ntasks, dtype=np.int32)
    hq.check(L.hq_schedule_tasks_prec(MT, NT, nb, ib, prec, tasks.ctypes.data_as(C.POINTER(C.c_int)),
                                      5 * ntasks))
    tr = torch.zeros(4 * ntasks, dtype=torch.int64, device=A.device)
    hq.check(L.hq_tile_qr_traced_prec_dev(MT, NT, nb, ib, prec, hq._ptr(A), hq._ptr(TAU), hq._ptr(T),
                                          hq._ptr(tr), hq._stream(stream)))
    hq.sync(stream)
    khz = C.c_int(0)
    hq.check(L.hq_wall_clock_khz(C.byref(khz)))
    raw = tr.cpu().numpy().reshape(ntasks, 4).astype(np.float64)
    t0 = raw[:, 0].min()
    us = raw.copy()
    us[:, :3] = (raw[:, :3] - t0) * 1e3 / khz.value
    return tasks.reshape(ntasks, 5), us


def summarize(tasks, us, grid):
    make = us[:, 2].max()
    dur = us[:, 2] - us[:, 1]
    wait = us[:, 1] - us[:, 0]
    out = {"makespan_us": float(make), "tasks": int(len(tasks)), "workgroups": int(grid),
           "busy_fraction": float(dur.sum() / (grid * make)),
           "wait_fraction": float(wait.sum() / (grid * make)), "kinds": {}}
    for k, name in enumerate(KINDS):
        sel = tasks[:, 0] == k
        if not sel.any():
            continue
        d = dur[sel]
        out["kinds"][name] = {"count": int(sel.sum()), "mean_us": float(d.mean()),
                              "median_us": float(np.median(d)), "max_us": float(d.max()),
                              "total_us": float(d.sum()),
                              "mean_wait_us": float(wait[sel].mean())}
    # panel chain: time from first to last panel task of each step
    pan = tasks[:, 0] <= 1
    steps = []
    for k in np.unique(tasks[pan, 1]):
        s = pan & (tasks[:, 1] == k)
        steps.append((int(k), float(us[s, 1].min()), float(us[s, 2].max())))
    out["panel_steps"] = steps[:4] + steps[-2:]
    return out


def critical_path(tasks, us, MT, NT, nb, ib, grid, prec=0):
    """Longest dependency chain (us) with the measured task durations, and the
    last time a task became ready and then waited more than 5 us to start
    (claim and acquire take 1-3 us; after this time the run is limited by
    the DAG, not by the claim order).  None if the host DAG's order
    differs from the traced one."""
    import hybrid3_amd as hq
    t2, deps = hq.dag_host(MT, NT, nb, ib, 0, grid, prec)
    if t2.shape != tasks.shape or not (t2 == tasks).all():
        return None
    dur = us[:, 2] - us[:, 1]
    fin = np.zeros(len(tasks))
    ready = np.zeros(len(tasks))
    for i, d in enumerate(deps):
        d = d[d >= 0]
        if len(d):
            fin[i] = fin[d].max()
            ready[i] = us[d, 2].max()
        fin[i] += dur[i]
    late = us[:, 1] - ready > 5.0          # ready, but held behind other claims
    tail = float(ready[late].max()) if late.any() else 0.0
    return {"critical_path_us": float(fin.max()),
            "work_per_workgroup_us": float(dur.sum() / grid),
            "last_held_ready_task_us": tail}


def chrome(tasks, us, path):
    ev = []
    for (kind, k, r, j, s), (c, b, e, wg) in zip(tasks, us):
        ev.append({"name": f"{KINDS[kind]}({k},{r},{j},{s})", "ph": "X", "ts": float(b),
                   "dur": float(e - b), "pid": 0, "tid": int(wg), "cat": KINDS[kind]})
    with open(path, "w") as f:
        json.dump({"traceEvents": ev}, f)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--m", type=int, default=0)
    ap.add_argument("--nb", type=int, default=256)
    ap.add_argument("--ib", type=int, default=32)
    ap.add_argument("--seed", type=int, default=20151130)
    ap.add_argument("--chrome", default="")
    ap.add_argument("--critical-path", action="store_true",
                    help="also report the DAG's critical path with the measured durations")
    ap.add_argument("--raw", default="", help="write the task list and stamps (us) to this .npz")
    ap.add_argument("--prec", type=int, default=0,
                    help="trailing updates in FP64 (0), FP32 (1) or FP16 hi+lo splits (2)")
    args = ap.parse_args()
    import torch

    import hybrid3_amd as hq
    m = args.m or args.n
    MT, NT = m // args.nb, args.n // args.nb
    A = torch.empty(MT * NT * args.nb * args.nb, dtype=torch.float64, device="cuda")
    TAU = torch.zeros(MT * NT * args.nb, dtype=torch.float64, device="cuda")
    hq.fill_uniform_dev(MT, NT, args.nb, args.seed, A)
    run_traced(MT, NT, args.nb, args.ib, A, TAU, prec=args.prec)   # warm-up (schedule build, code load)
    hq.fill_uniform_dev(MT, NT, args.nb, args.seed, A)
    tasks, us = run_traced(MT, NT, args.nb, args.ib, A, TAU, prec=args.prec)
    _, grid = hq.schedule_info(MT, NT, args.nb, args.ib, args.prec)
    s = summarize(tasks, us, grid)
    flops = 4.0 / 3.0 * args.n ** 3 if m == args.n else 2.0 * m * args.n ** 2 - 2.0 * args.n ** 3 / 3
    s["tflops_traced"] = flops / (s["makespan_us"] * 1e-6) / 1e12
    if args.critical_path:
        s["critical_path"] = critical_path(tasks, us, MT, NT, args.nb, args.ib, grid, args.prec)
    print(json.dumps(s, indent=1))
    if args.chrome:
        chrome(tasks, us, args.chrome)
    if args.raw:
        np.savez_compressed(args.raw, tasks=tasks, us=us)


if __name__ == "__main__":
    main()
