"""Batched uneven QR on one GPU: batch storage, the autotuner, and the
measured launch paths.

A batch is a list of (m, n) sizes.  A plan gives each matrix a tile size nb,
inner block ib and panel mapping (0: one task per tile panel QR, one
workgroup per panel; 1: one task per ib block plus one trailing task per
(block, 64-column slab), so a panel spreads over several workgroups).
Matrices with the same (nb, ib, mapping) are factored by ONE persistent launch
of the DAG scheduler (``hq_batched_tile_qr_ex_dev``); a plan with several
groups issues one launch per group on the same stream.

Matrices are FP64, or FP32 (column-major, leading dimension m, factored in
place like a batched sgeqrf): an FP32 batch is widened into the FP64 tile
layout by one layout launch per group, factored in FP64 arithmetic, and the
factor (R and the tile reflectors) rounded back into the FP32 arrays.

The autotuner chooses the plan by measurement:
  1. per size bucket (by m), time every candidate (nb, ib, mapping) on a
     sample of the bucket's matrices, factored as one batch;
  2. time the whole batch under the per-bucket plan and under every
     uniform plan, and keep the fastest.
Choices are measured on the device that runs them; nothing is predicted.
``python -m hybrid3_amd.batched --out FILE`` runs it on the uneven batch of
4000 matrices and writes the choice as a JSON table (``TABLE`` is the one
measured on an MI355X), which ``plan_from_table`` turns into a plan without
measuring.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import (batched_fill_uniform_dev, batched_plan_dev, batched_plan_free, batched_run_dev,
               batched_tile_qr_ex_dev, cm32_to_tile_dev, sync, tile_orgqr_dev, tile_ormqr_dev, tile_qr_dev,
               tile_to_cm32_dev)

CANDIDATES = tuple((nb, ib, mp) for nb in (64, 128, 256) for ib in (16, 32) for mp in (0, 1))
BUCKETS = ((1, 256), (256, 512), (512, 1 << 30))  # half-open ranges of m
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "batched_mi355x.json")


def _key(p):
    """(nb, ib, mapping) of a plan entry given as (nb, ib) or (nb, ib, mapping)."""
    return (int(p[0]), int(p[1]), int(p[2]) if len(p) > 2 else 0)


def _name(c):
    return f"{c[0]}/{c[1]}/{'split' if c[2] else 'tile'}"


def _parse(name):
    nb, ib, mp = name.split("/")
    return (int(nb), int(ib), 1 if mp == "split" else 0)


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
    from seed0 + i (the same values for every plan); with fp32=True those
    values rounded to FP32, kept column-major (ld = m) in ``C32``, which
    factor() overwrites with the FP32 factor."""

    def __init__(self, sizes, plan, seed0=1, fp32=False):
        import torch
        self.sizes = list(sizes)
        self.plan = [_key(p) for p in plan]
        self.fp32 = fp32
        self.tiles = [((m + nb - 1) // nb, (n + nb - 1) // nb) for (m, n), (nb, _, _) in zip(self.sizes, self.plan)]
        elems = [MT * NT * nb * nb for (MT, NT), (nb, _, _) in zip(self.tiles, self.plan)]
        taus = [MT * NT * nb for (MT, NT), (nb, _, _) in zip(self.tiles, self.plan)]
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
            self.groups.setdefault(p, []).append(i)
        for (nb, _, _), idx in self.groups.items():
            batched_fill_uniform_dev(nb, [self.sizes[i][0] for i in idx], [self.sizes[i][1] for i in idx],
                                     [seed0 + i for i in idx], [self.a_ptr[i] for i in idx])
        if fp32:
            cnt = [m * n for m, n in self.sizes]
            self.c_off = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
            self.C32 = torch.empty(int(sum(cnt)), dtype=torch.float32, device="cuda")
            self.c_ptr = [self.C32.data_ptr() + 4 * int(o) for o in self.c_off]
            self._to_cm32()
            self.pristine = self.C32.clone()
        else:
            self.pristine = self.A.clone()

    def _group_args(self, idx):
        return ([self.sizes[i][0] for i in idx], [self.sizes[i][1] for i in idx],
                [self.a_ptr[i] for i in idx], [self.c_ptr[i] for i in idx], [self.sizes[i][0] for i in idx])

    def _to_cm32(self):
        for (nb, _, _), idx in self.groups.items():
            m, n, a, c, ld = self._group_args(idx)
            tile_to_cm32_dev(nb, m, n, a, c, ld)

    def restore(self):
        if self.fp32:
            self.C32.copy_(self.pristine)
        else:
            self.A.copy_(self.pristine)

    def factor(self):
        """One persistent launch per (nb, ib, mapping) group; an FP32 batch
        adds one widening and one rounding layout launch per group."""
        for (nb, ib, mp), idx in self.groups.items():
            if self.fp32:
                m, n, a, c, ld = self._group_args(idx)
                cm32_to_tile_dev(nb, m, n, c, ld, a)
            batched_tile_qr_ex_dev(0, mp, nb, ib, [self.tiles[i][0] for i in idx], [self.tiles[i][1] for i in idx],
                                   [self.a_ptr[i] for i in idx], [self.t_ptr[i] for i in idx],
                                   ncol=[self.sizes[i][1] for i in idx])
            if self.fp32:
                tile_to_cm32_dev(nb, m, n, a, c, ld)

    def capture(self):
        """A hipGraph of one run per (nb, ib, mapping) group of prepared plans
        (hq_batched_plan_ex_dev): replaying it factors the batch in place
        (FP64 batches: the layout launches of an FP32 batch are not captured)."""
        import torch
        if self.fp32:
            raise ValueError("graph capture covers FP64 batches")
        self.release()
        self.plans = []
        for (nb, ib, mp), idx in self.groups.items():
            self.plans.append(batched_plan_dev(nb, ib, [self.tiles[i][0] for i in idx],
                                               [self.tiles[i][1] for i in idx],
                                               [self.a_ptr[i] for i in idx], [self.t_ptr[i] for i in idx],
                                               split_panels=mp, ncol=[self.sizes[i][1] for i in idx]))
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
            nb, ib, _ = self.plan[i]
            MT, NT = self.tiles[i]
            a = self.A[int(self.a_off[i]):int(self.a_off[i]) + self.elems[i]]
            t = self.TAU[int(self.t_off[i]):int(self.t_off[i]) + self.taus[i]]
            tile_qr_dev(MT, NT, nb, ib, a, t, None)

    def _tiles_of(self, src):
        """Tile-layout FP64 copies of the matrices (pristine values or the
        factor): the FP64 tile arrays themselves, or, for an FP32 batch, the
        FP32 column-major arrays widened."""
        import torch
        if not self.fp32:
            return src
        out = torch.empty_like(self.A)
        ptrs = [out.data_ptr() + 8 * int(o) for o in self.a_off]
        for (nb, _, _), idx in self.groups.items():
            cm32_to_tile_dev(nb, [self.sizes[i][0] for i in idx], [self.sizes[i][1] for i in idx],
                             [src.data_ptr() + 4 * int(self.c_off[i]) for i in idx],
                             [self.sizes[i][0] for i in idx], [ptrs[i] for i in idx])
        sync()
        return out

    def colnorm_error(self):
        """max over matrices and columns of | ||A e_j|| - ||R e_j|| | / ||A e_j||
        (Q orthogonal); zero padding columns are skipped.  FP32 batches: R as
        stored in FP32 against the FP32 input."""
        import torch
        pristine = self._tiles_of(self.pristine)
        factor = self._tiles_of(self.C32) if self.fp32 else self.A
        worst = 0.0
        for i, ((MT, NT), (nb, _, _)) in enumerate(zip(self.tiles, self.plan)):
            lo, hi = int(self.a_off[i]), int(self.a_off[i]) + self.elems[i]
            a = pristine[lo:hi].view(NT, MT, nb, nb)
            f = factor[lo:hi].view(NT, MT, nb, nb)
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

    def residuals(self, idx):
        """Per matrix i in idx: ||Q^T A - [R; 0]||_F / ||A||_F (equal to
        ||A - QR||_F / ||A||_F, Q^T applied by the engine's dormqr from the
        stored reflectors) and ||P^T P - I||_F of the dorgqr array P = Q^T E
        (E the first NT*nb columns of the identity), over the m x n matrix.
        FP32 batches use the factor as stored in FP32 (reflectors rounded)."""
        import torch
        pristine = self._tiles_of(self.pristine)
        factor = self._tiles_of(self.C32) if self.fp32 else self.A
        out = []
        for i in idx:
            (m, n), (MT, NT), (nb, ib, _) = self.sizes[i], self.tiles[i], self.plan[i]
            lo, hi = int(self.a_off[i]), int(self.a_off[i]) + self.elems[i]
            tlo, thi = int(self.t_off[i]), int(self.t_off[i]) + self.taus[i]
            ry, tau = factor[lo:hi].clone(), self.TAU[tlo:thi]
            x = pristine[lo:hi].clone()
            tile_ormqr_dev(MT, NT, nb, ib, ry, tau, x, NT)
            q = torch.empty_like(ry)
            tile_orgqr_dev(MT, NT, nb, ib, ry, tau, q)
            sync()

            def dense(t):
                return t.view(NT, MT, nb, nb).permute(1, 3, 0, 2).reshape(MT * nb, NT * nb)
            a, qa, f, p = dense(pristine[lo:hi]), dense(x), dense(ry), dense(q)
            r = torch.triu(f)[:, :n]
            res = (qa[:, :n] - r).norm() / a[:m, :n].norm()
            pp = p[:, :n].T @ p[:, :n]
            orth = (pp - torch.eye(n, dtype=pp.dtype, device=pp.device)).norm()
            out.append((res.item(), orth.item()))
        return out


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


def autotune(sizes, candidates=CANDIDATES, sample=192, reps=2, seed=0, log=None):
    """Plan for the batch and a report of what was measured; log(str), if
    given, receives one line per measurement."""
    rng = np.random.default_rng(seed)
    candidates = [_key(c) for c in candidates]
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
            if log:
                log(f"bucket {lo}..{hi} {_name(c)}: {times[c] * 1e3:.3f} ms")
        best = min(times, key=times.get)
        per_bucket[b] = best
        report["buckets"].append({"m_range": [lo, hi], "sample": len(sub),
                                  "seconds": {_name(c): round(t, 6) for c, t in times.items()},
                                  "best": _name(best)})
    plans = {"per-bucket": [per_bucket[bucket_of(m)] for m, _ in sizes]}
    for c in candidates:
        plans[f"uniform {_name(c)}"] = [c] * len(sizes)
    best_name, best_t = None, float("inf")
    for name, plan in plans.items():
        bt = Batch(sizes, plan, seed0=29)
        t = time_factor(bt, reps)
        del bt
        report["plans"][name] = round(t, 6)
        if log:
            log(f"plan {name}: {t * 1e3:.3f} ms")
        if t < best_t:
            best_name, best_t = name, t
    report["chosen"] = best_name
    return plans[best_name], report


def table_from_report(report, device=None):
    """The JSON tuning table of an autotune report: the per-bucket choices,
    the whole-batch policy chosen, and the times behind them."""
    return {"device": device, "buckets": [{"m_range": b["m_range"], "best": b["best"]} for b in report["buckets"]],
            "chosen": report["chosen"], "measured": report}


def load_table(path=TABLE):
    with open(path) as f:
        return json.load(f)


def plan_from_table(sizes, table):
    """The plan a tuning table gives for `sizes` (no measurement): the
    uniform candidate it chose, or its per-bucket choices."""
    chosen = table["chosen"]
    if chosen.startswith("uniform "):
        c = _parse(chosen[len("uniform "):])
        return [c] * len(sizes)
    best = {}
    for b in table["buckets"]:
        best[tuple(b["m_range"])] = _parse(b["best"])
    plan = []
    for m, _ in sizes:
        lo, hi = BUCKETS[bucket_of(m)]
        plan.append(best[(lo, hi)])
    return plan


def main(argv=None):
    import argparse

    import torch
    ap = argparse.ArgumentParser(description="measure the batched QR plan of the uneven batch and write "
                                             "the tuning table")
    ap.add_argument("--count", type=int, default=4000)
    ap.add_argument("--seed", type=int, default=4000)
    ap.add_argument("--out", default=None, help="JSON table to write (default: print)")
    args = ap.parse_args(argv)
    sizes = uneven_sizes(args.count, seed=args.seed)
    plan, report = autotune(sizes, log=lambda line: print(line, flush=True))
    props = torch.cuda.get_device_properties(0)
    table = table_from_report(report, f"{props.name} ({props.gcnArchName.split(':')[0]}, "
                                      f"{props.multi_processor_count} CUs)")
    table["workload"] = (f"{args.count} matrices, m uniform in [64,1024], n uniform in [32,m], seed {args.seed} "
                         "(python -m hybrid3_amd.batched)")
    text = json.dumps(table, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    print(text, flush=True)


if __name__ == "__main__":
    main()
