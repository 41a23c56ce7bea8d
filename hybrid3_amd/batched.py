"""Batched uneven QR on one GPU: batch storage, the autotuner, and the
measured launch paths.

A batch is a list of (m, n) sizes.  A plan gives each matrix a tile size nb
and inner block ib.  Matrices with the same (nb, ib) are factored by ONE
persistent launch of the DAG scheduler (``hq_batched_tile_qr_dev``); a plan
with several (nb, ib) groups issues one launch per group on the same
stream.

The autotuner chooses the plan by measurement:
  1. per size bucket (by m), time every candidate (nb, ib) on a sample of
     the bucket's matrices, factored as one batch;
  2. time the whole batch under the per-bucket plan and under every
     uniform plan, and keep the fastest.
Choices are measured on the device that runs them; nothing is predicted.
"""
from __future__ import annotations

import time

import numpy as np

from . import (batched_fill_uniform_dev, batched_plan_dev, batched_plan_free, batched_run_dev,
               batched_tile_qr_dev, sync, tile_qr_dev)

CANDIDATES = ((64, 16), (64, 32), (128, 32))
BUCKETS = ((1, 256), (256, 512), (512, 1 << 30))  # half-open ranges of m


def uneven_sizes(count=4000, seed=4000, m_range=(64, 1024), n_min=32):
    """count sizes with m uniform in m_range and n uniform in [n_min, m]."""
    rng = np.random.default_rng(seed)
    ms = rng.integers(m_range[0], m_range[1] + 1, count)
    ns = [int(rng.integers(n_min, m + 1)) for m in ms]
    return [(int(m), n) for m, n in zip(ms, ns)]


def flops(sizes):
    """Algorithmic Householder QR work, 2mn^2 - 2n^3/3 per m x n matrix."""
    return float(sum(2.0 * m * n * n - 2.0 * n ** 3 / 3.0 for m, n in sizes))


def bucket_of(m):
    for b, (lo, hi) in enumerate(BUCKETS):
        if lo <= m < hi:
            return b
    return len(BUCKETS) - 1


class Batch:
    """Device storage of a batch in tile layout (zero padding) for a plan.

    Matrix i holds the m x n values of a column-major uniform[-1,1) fill
    from seed0 + i (the same values for every plan)."""

    def __init__(self, sizes, plan, seed0=1):
        import torch
        self.sizes = list(sizes)
        self.plan = list(plan)
        self.tiles = [((m + nb - 1) // nb, (n + nb - 1) // nb) for (m, n), (nb, _) in zip(self.sizes, self.plan)]
        elems = [MT * NT * nb * nb for (MT, NT), (nb, _) in zip(self.tiles, self.plan)]
        taus = [MT * NT * nb for (MT, NT), (nb, _) in zip(self.tiles, self.plan)]
        self.a_off = np.concatenate([[0], np.cumsum(elems)[:-1]]).astype(np.int64)
        self.t_off = np.concatenate([[0], np.cumsum(taus)[:-1]]).astype(np.int64)
        self.elems, self.taus = elems, taus
        self.A = torch.empty(int(sum(elems)), dtype=torch.float64, device="cuda")
        self.TAU = torch.zeros(int(sum(taus)), dtype=torch.float64, device="cuda")
        base_a, base_t = self.A.data_ptr(), self.TAU.data_ptr()
        self.a_ptr = [base_a + 8 * int(o) for o in self.a_off]
        self.t_ptr = [base_t + 8 * int(o) for o in self.t_off]
        self.groups = {}
        for i, p in enumerate(self.plan):
            self.groups.setdefault(tuple(p), []).append(i)
        for (nb, _), idx in self.groups.items():
            batched_fill_uniform_dev(nb, [self.sizes[i][0] for i in idx], [self.sizes[i][1] for i in idx],
                                     [seed0 + i for i in idx], [self.a_ptr[i] for i in idx])
        self.pristine = self.A.clone()

    def restore(self):
        self.A.copy_(self.pristine)

    def factor(self):
        """One persistent launch per (nb, ib) group."""
        for (nb, ib), idx in self.groups.items():
            batched_tile_qr_dev(nb, ib, [self.tiles[i][0] for i in idx], [self.tiles[i][1] for i in idx],
                                [self.a_ptr[i] for i in idx], [self.t_ptr[i] for i in idx])

    def capture(self):
        """A hipGraph of one run per (nb, ib) group of prepared plans
        (hq_batched_plan_dev): replaying it factors the batch in place."""
        import torch
        self.release()
        self.plans = []
        for (nb, ib), idx in self.groups.items():
            self.plans.append(batched_plan_dev(nb, ib, [self.tiles[i][0] for i in idx],
                                               [self.tiles[i][1] for i in idx],
                                               [self.a_ptr[i] for i in idx], [self.t_ptr[i] for i in idx]))
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            for p in self.plans:
                batched_run_dev(p)
        return self.graph

    def release(self):
        for p in getattr(self, "plans", []):
            batched_plan_free(p)
        self.plans = []
        self.graph = None

    def factor_per_matrix(self, idx=None):
        """The baseline: one hq_tile_qr_dev launch per matrix."""
        for i in (range(len(self.sizes)) if idx is None else idx):
            nb, ib = self.plan[i]
            MT, NT = self.tiles[i]
            a = self.A[int(self.a_off[i]):int(self.a_off[i]) + self.elems[i]]
            t = self.TAU[int(self.t_off[i]):int(self.t_off[i]) + self.taus[i]]
            tile_qr_dev(MT, NT, nb, ib, a, t, None)

    def colnorm_error(self):
        """max over matrices and columns of | ||A e_j|| - ||R e_j|| | / ||A e_j||
        (Q orthogonal); zero padding columns are skipped."""
        import torch
        worst = 0.0
        for i, ((MT, NT), (nb, _)) in enumerate(zip(self.tiles, self.plan)):
            lo, hi = int(self.a_off[i]), int(self.a_off[i]) + self.elems[i]
            a = self.pristine[lo:hi].view(NT, MT, nb, nb)
            f = self.A[lo:hi].view(NT, MT, nb, nb)
            tj = torch.arange(NT, device="cuda").view(NT, 1, 1, 1)
            ti = torch.arange(MT, device="cuda").view(1, MT, 1, 1)
            col = torch.arange(nb, device="cuda").view(1, 1, nb, 1)
            row = torch.arange(nb, device="cuda").view(1, 1, 1, nb)
            upper = (ti * nb + row) <= (tj * nb + col)
            an = (a * a).sum(dim=(1, 3)).sqrt()
            rn = (f * f * upper).sum(dim=(1, 3)).sqrt()
            live = an > 0
            worst = max(worst, ((an[live] - rn[live]).abs() / an[live]).max().item())
        return worst


def time_factor(batch, reps=3, per_matrix=False, graph=None):
    """Best-of-reps seconds for factoring the whole batch (restored first):
    one launch per (nb, ib) group, one launch per matrix, or a replay of a
    captured graph."""
    import torch
    best = float("inf")
    for _ in range(reps):
        batch.restore()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if graph is not None:
            graph.replay()
        elif per_matrix:
            batch.factor_per_matrix()
        else:
            batch.factor()
        sync()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def autotune(sizes, candidates=CANDIDATES, sample=192, reps=2, seed=0):
    """Plan for the batch and a report of what was measured."""
    rng = np.random.default_rng(seed)
    report = {"buckets": [], "plans": {}}
    per_bucket = {}
    for b, (lo, hi) in enumerate(BUCKETS):
        idx = [i for i, (m, _) in enumerate(sizes) if lo <= m < hi]
        if not idx:
            continue
        pick = sorted(rng.choice(idx, size=min(sample, len(idx)), replace=False).tolist())
        sub = [sizes[i] for i in pick]
        times = {}
        for c in candidates:
            if c[0] > 2 * max(m for m, _ in sub):
                continue  # a tile far larger than every matrix of the bucket
            bt = Batch(sub, [c] * len(sub), seed0=17)
            times[c] = time_factor(bt, reps)
            del bt
        best = min(times, key=times.get)
        per_bucket[b] = best
        report["buckets"].append({"m_range": [lo, hi], "sample": len(sub),
                                  "seconds": {f"{c[0]}/{c[1]}": round(t, 6) for c, t in times.items()},
                                  "best": f"{best[0]}/{best[1]}"})
    plans = {"per-bucket": [per_bucket[bucket_of(m)] for m, _ in sizes]}
    for c in candidates:
        plans[f"uniform {c[0]}/{c[1]}"] = [c] * len(sizes)
    best_name, best_t = None, float("inf")
    for name, plan in plans.items():
        bt = Batch(sizes, plan, seed0=29)
        t = time_factor(bt, reps)
        del bt
        report["plans"][name] = round(t, 6)
        if t < best_t:
            best_name, best_t = name, t
    report["chosen"] = best_name
    return plans[best_name], report
