#!/usr/bin/env python3
"""Benchmark: plain FP64 flat-tree tile QR on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on):
square 16384 x 16384 FP64, entries uniform[-1,1) from a seeded SplitMix64
generator, tile nb=256, inner block ib=32; the whole factorisation is one
launch of the persistent on-device DAG scheduler.  A "step" is one complete
factorisation of the (restored) matrix, already resident in HBM.

Prints ONE JSON line.  value = (4/3 n^3) / (mean factorisation time), in
TFLOP/s; the timing of each factorisation uses HIP events on the stream the
scheduler kernel is launched on.  The same line carries the batched part of
the metric ("batched": QRs/sec of 4000 uneven FP64 matrices factored by the
autotuned batched launch, beside one launch per matrix).  The CPU baseline is the oracle's OpenMP
restatement of the reference's TileSch dynamic DAG (one team, LAPACK-style
kernels in plain C), timed on a bounded sample on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_MFMA_PEAK = 78.6  # TFLOP/s, MI355X dense FP64 matrix (published)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--nb", type=int, default=256)
    ap.add_argument("--ib", type=int, default=32)
    ap.add_argument("--seed", type=int, default=20151130)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched uneven-QR measurement")
    ap.add_argument("--batch-count", type=int, default=4000)
    ap.add_argument("--retune", action="store_true",
                    help="measure the batched plan (autotuner) instead of reading the committed tuning table")
    ap.add_argument("--no-lsq", action="store_true", help="skip the mixed-precision least-squares part")
    ap.add_argument("--no-batched-lsq", action="store_true", help="skip the batched least-squares part")
    ap.add_argument("--batched-lsq-count", type=int, default=512)
    ap.add_argument("--no-tall", action="store_true", help="skip the 65536 x 4096 QR / dorgqr / dormqr part")
    ap.add_argument("--prec", choices=("fp64", "fp32", "split"), default="fp64",
                    help="precision of the trailing updates of the timed factorisation (the headline is fp64)")
    ap.add_argument("--no-panels", action="store_true", help="skip the whole-chip panel-kernel measurement")
    ap.add_argument("--no-library", action="store_true", help="skip the rocSOLVER / host LAPACK reference lines")
    ap.add_argument("--no-homegrown", action="store_true", help="skip the tile_kernel.h per-call timings")
    ap.add_argument("--no-tile-sizes", action="store_true",
                    help="skip the factorisation with nb = 384 / 512 (generic kernels)")
    ap.add_argument("--only-factor", action="store_true",
                    help="only the timed factorisation and its check (profiling runs)")
    ap.add_argument("--only-panels", action="store_true", help="only the panel-kernel measurement (profiling runs)")
    ap.add_argument("--only-batched", action="store_true",
                    help="only the batched uneven-QR measurement, FP64 and FP32 matrices (profiling runs)")
    ap.add_argument("--only-tall", action="store_true",
                    help="only the 65536 x 4096 QR / dorgqr / dormqr measurement (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only a 2 GiB device copy (calibrates FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--traffic-file",
                    default=os.path.join(ROOT, "profiles", "sched_kernel_traffic.txt"))
    return ap.parse_args()


def note(msg):
    """One progress line on stderr per finished section (stdout holds only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads():
    """Host cores this process may use: its CPU affinity, bounded by OMP_NUM_THREADS when that is set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(omp)) if omp.isdigit() and int(omp) > 0 else n


def cpu_baseline(nb, ib, seed):
    """The reference's CPU path restated by the oracle (TileSch-style dynamic
    DAG over LAPACK-style tile kernels in plain C, OpenMP, every usable host
    core) on an 8192 x 8192 sample of the benchmark's matrix family; beside
    it, the host's LAPACK (numpy.linalg.qr, dgeqrf on multithreaded BLAS) on
    the same sample as a library reference."""
    import numpy as np

    import oracle
    threads = host_threads()
    n = 8192
    t = oracle.time_tile_qr(n // nb, n // nb, nb, ib, seed, threads, 1)
    out = {"value": round(4.0 / 3.0 * n ** 3 / t / 1e12, 5), "unit": "TFLOP/s", "cores": threads,
           "kind": "port",
           "sample": f"{n}x{n} FP64 tile QR nb={nb} ib={ib}, oracle C/OpenMP TileSch-style DAG, "
                     f"{threads} threads, {t:.2f} s"}
    A = oracle.fill_uniform_cm(n, n, seed)
    t0 = time.perf_counter()
    np.linalg.qr(A, mode="r")
    tl = time.perf_counter() - t0
    out["host_lapack_dgeqrf"] = {"value": round(4.0 / 3.0 * n ** 3 / tl / 1e12, 5), "unit": "TFLOP/s",
                                 "sample": f"numpy.linalg.qr(mode='r') of the same {n}x{n} matrix, {tl:.2f} s"}
    return out


def tile_size_bench(n, cases, seed, reps=2):
    """The same n x n FP64 factorisation with other tile sizes (nb, ib): the
    tiles nb = 384 and 512 run the generic kernels (16384 is not a multiple of
    384: edge tiles, hq_tile_qr_mn_dev).  HIP events around each
    factorisation, best of `reps` after one warm-up; rates use 4/3 n^3 of the
    n x n matrix; column-norm check of R against A."""
    import torch

    import hybrid3_amd as hq
    stream = torch.cuda.current_stream()
    out = []
    for nb, ib in cases:
        MT = -(-n // nb)
        pristine = torch.empty(MT * MT * nb * nb, dtype=torch.float64, device="cuda")
        hq.fill_uniform_dev(MT, MT, nb, seed, pristine)
        A = torch.empty_like(pristine)
        TAU = torch.zeros(MT * MT * nb, dtype=torch.float64, device="cuda")
        best = float("inf")
        for _ in range(reps + 1):
            A.copy_(pristine)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            hq.tile_qr_mn_dev(n, n, nb, ib, A, TAU)
            e1.record(stream)
            hq.sync()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) * 1e-3)
        # the call zeroed the padding of A; the pristine copy still holds it
        live = torch.zeros(MT * nb, dtype=torch.bool, device="cuda")
        live[:n] = True
        a = pristine.view(MT, MT, nb, nb)
        f = A.view(MT, MT, nb, nb)
        tj = torch.arange(MT, device="cuda").view(MT, 1, 1, 1)
        ti = torch.arange(MT, device="cuda").view(1, MT, 1, 1)
        col = torch.arange(nb, device="cuda").view(1, 1, nb, 1)
        row = torch.arange(nb, device="cuda").view(1, 1, 1, nb)
        rowok = (ti * nb + row) < n
        upper = (ti * nb + row) <= (tj * nb + col)
        an = (a * a * rowok).sum(dim=(1, 3)).sqrt().reshape(-1)[:n]
        rn = (f * f * upper).sum(dim=(1, 3)).sqrt().reshape(-1)[:n]
        err = ((an - rn).abs() / an).max().item()
        out.append({"nb": nb, "ib": ib, "ms": round(best * 1e3, 3),
                    "tflops": round(4.0 / 3.0 * n ** 3 / best / 1e12, 3),
                    "kernels": "specialised" if nb in (64, 128, 256) else "generic",
                    "max_col_norm_rel_diff": err})
        del pristine, A, TAU
        torch.cuda.empty_cache()
    return out


def rocsolver_reference(n, seed):
    """rocSOLVER's dgeqrf (torch.geqrf) on the same n x n FP64 matrix family,
    resident in HBM, timed with HIP events (best of 3): a library reference
    for the headline."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    A0 = torch.rand(n, n, generator=g, dtype=torch.float64, device="cuda") * 2 - 1
    best = float("inf")
    for _ in range(4):
        A = A0.clone()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        torch.geqrf(A)
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del A0, A
    torch.cuda.empty_cache()
    return {"workload": f"torch.geqrf (rocSOLVER dgeqrf) {n}x{n} FP64, uniform[-1,1)", "ms": round(best, 3),
            "tflops": round(4.0 / 3.0 * n ** 3 / (best * 1e-3) / 1e12, 3)}


def panel_bench(nb=256, ib=32, count=2048, reps=3):
    """The panel kernels at whole-chip load: a batch of `count` single-tile
    matrices (one GEQRT task each) and one of `count` two-tile columns
    (GEQRT, then TSQRT of [R; A2]), each factored in one persistent launch.
    Algorithmic HBM bytes (the tiles read and written once, T and TAU
    written): GEQRT 2*8*nb^2 + 8*ib*nb + 8*nb, TSQRT 3*8*nb^2 + 8*ib*nb +
    8*nb (R triangle read and written as a whole tile).  Rates use the mean
    of `reps` launches (HIP events)."""
    import torch

    import hybrid3_amd as hq
    out = {"nb": nb, "ib": ib, "count": count}
    for name, MT in (("geqrt", 1), ("geqrt_tsqrt", 2)):
        A = torch.empty(count, MT * nb * nb, dtype=torch.float64, device="cuda")
        TAU = torch.zeros(count, MT * nb, dtype=torch.float64, device="cuda")
        pristine = torch.rand(count, MT * nb * nb, dtype=torch.float64, device="cuda") * 2 - 1
        ptrs = [A[i].data_ptr() for i in range(count)]
        tps = [TAU[i].data_ptr() for i in range(count)]
        A.copy_(pristine)
        hq.batched_tile_qr_dev(nb, ib, [MT] * count, [1] * count, ptrs, tps)   # schedule, code load
        hq.sync()
        ms = []
        for _ in range(reps):
            A.copy_(pristine)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            hq.batched_tile_qr_dev(nb, ib, [MT] * count, [1] * count, ptrs, tps)
            e1.record()
            hq.sync()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        t = sum(ms) / len(ms) * 1e-3
        geq = 2 * 8 * nb * nb + 8 * ib * nb + 8 * nb
        tsq = 3 * 8 * nb * nb + 8 * ib * nb + 8 * nb
        byts = count * (geq + (tsq if MT == 2 else 0))
        fl = count * (4.0 / 3.0 * nb ** 3 + (2.0 * nb ** 3 if MT == 2 else 0.0))
        out[name] = {"ms": round(t * 1e3, 3), "algorithmic_GB_per_s": round(byts / t / 1e9, 1),
                     "tflops": round(fl / t / 1e12, 3), "bytes": byts}
        del A, TAU, pristine
        torch.cuda.empty_cache()
    return out


def batched_bench(count, retune=False, check=True):
    """Batched uneven QR: `count` matrices, m ~ U[64,1024], n ~ U[32,m]
    (seed 4000), (nb, ib, panel mapping) plan from the committed tuning table
    (or measured by the autotuner), one persistent launch per group; against
    one hq_tile_qr_dev launch per matrix on the same data.  The same sizes as
    FP32 matrices (widened, factored in FP64, rounded back; the layout
    launches inside the timed region).  Per-matrix ||A - QR|| / ||A|| and
    ||I - Q^T Q|| on every 20th matrix."""
    import os

    from hybrid3_amd import batched as B
    sizes = B.uneven_sizes(count, seed=4000)
    if retune or not os.path.exists(B.TABLE):
        plan, report = B.autotune(sizes)
        source = "autotuner (measured in this run)"
    else:
        table = B.load_table()
        plan, report = B.plan_from_table(sizes, table), table["measured"]
        source = "tuning table " + os.path.relpath(B.TABLE, os.path.dirname(os.path.abspath(__file__)))
    bt = B.Batch(sizes, plan, seed0=1)
    B.time_factor(bt, reps=1)      # schedules built and cached
    t = B.time_factor(bt, reps=3)
    sample = list(range(0, count, 20)) if check else []
    err = bt.colnorm_error() if check else float("nan")
    res = bt.residuals(sample) or [(float("nan"), float("nan"))]
    tg = B.time_factor(bt, reps=3, graph=bt.capture())   # the same runs replayed from a hipGraph
    if check:
        err = max(err, bt.colnorm_error())
    bt.release()
    tp = B.time_factor(bt, reps=1, per_matrix=True) if check else float("nan")
    fl = B.flops(sizes)
    del bt
    b32 = B.Batch(sizes, plan, seed0=1, fp32=True)
    B.time_factor(b32, reps=1)
    t32 = B.time_factor(b32, reps=3)
    err32 = b32.colnorm_error() if check else float("nan")
    res32 = b32.residuals(sample) or [(float("nan"), float("nan"))]
    del b32
    return {
        "workload": f"{count} FP64 matrices, m uniform in [64,1024], n uniform in [32,m], "
                    "uniform[-1,1) entries, seed 4000; zero-padded tile layout",
        "qrs_per_s": round(count / t, 1),
        "ms": round(t * 1e3, 3),
        "tflops_algorithmic": round(fl / t / 1e12, 3),
        "graph_replay_qrs_per_s": round(count / tg, 1),
        "per_matrix_launch_qrs_per_s": round(count / tp, 1),
        "speedup_vs_per_matrix_launch": round(tp / t, 2),
        "launches": len({B._key(p) for p in plan}),
        "plan_source": source,
        "plan": report["chosen"],
        "tuner": report,
        "max_col_norm_rel_diff": err,
        "checked_matrices": len(sample),
        "max_residual": max(r for r, _ in res),
        "max_orthogonality": max(o for _, o in res),
        "fp32": {
            "workload": f"the same {count} sizes as FP32 column-major matrices (the FP64 values rounded), "
                        "widened to the FP64 tile layout, factored in FP64, factor rounded back to FP32",
            "qrs_per_s": round(count / t32, 1),
            "ms": round(t32 * 1e3, 3),
            "tflops_algorithmic": round(fl / t32 / 1e12, 3),
            "max_col_norm_rel_diff": err32,
            "max_residual": max(r for r, _ in res32),
            "max_orthogonality": max(o for _, o in res32),
        },
    }


def homegrown_bench(n=256, reps=5):
    """The homegrown tile_kernel.h calls (host arrays, as the reference's
    callers pass them) on n x n tiles: wall time per call, best of `reps`
    after a warm-up, the copies to and from the device and the stream
    synchronisation included (one workgroup per call)."""
    import ctypes as C

    import numpy as np

    import hybrid3_amd as hq
    L = hq.lib()
    rng = np.random.default_rng(7)

    def P(a):
        return a.ctypes.data_as(C.POINTER(C.c_double))

    def F(a):
        return np.array(a, dtype=np.float64, order="F", copy=True)

    A0 = F(rng.uniform(-1, 1, (n, n)))
    R0 = F(np.triu(rng.uniform(-1, 1, (n, n))))
    B0 = F(rng.uniform(-1, 1, (n, n)))
    TAU, Y, T = np.zeros(n), F(np.zeros((n, n))), F(np.zeros((n, n)))
    calls = {
        "tile_dgeqrt": lambda A: L.tile_dgeqrt(n, n, n, P(A), n, P(TAU), P(Y), P(T), n, None),
        "tile_dtsqrt": lambda A: L.tile_dtsqrt(n, n, n, P(R0.copy(order="F")), n, P(A), n, P(TAU), P(T), n, None),
        "tile_dlarft": lambda A: L.tile_dlarft(n, n, 1, P(Y), n, P(TAU), P(T), n, None),
        "tile_dlarfb": lambda A: L.tile_dlarfb(n, n, P(A), n, P(Y), P(T), n, None),
        "tile_dssrfb": lambda A: L.tile_dssrfb(n, n, P(A), P(B0.copy(order="F")), n, P(Y), P(T), n, None),
    }
    flops = {"tile_dgeqrt": 4.0 / 3.0 * n ** 3, "tile_dtsqrt": 2.0 * n ** 3, "tile_dlarft": n ** 3 / 3.0,
             "tile_dlarfb": 2.0 * n ** 3, "tile_dssrfb": 4.0 * n ** 3}
    out = {"n": n}
    for name, fn in calls.items():
        best = float("inf")
        for _ in range(reps + 1):
            A = A0.copy(order="F")
            t0 = time.perf_counter()
            fn(A)
            best = min(best, time.perf_counter() - t0)
        if hq.lib().hq_last_error() != 0:
            raise RuntimeError(f"{name} failed")
        out[name] = {"ms": round(best * 1e3, 3), "gflops": round(flops[name] / best / 1e9, 1)}
    return out


def tall_bench(m=65536, n=4096, nrhs=64, nb=256, ib=32, seed=20151130, reps=3):
    """Tall FP64 tile QR (m x n), the Q^T array of dorgqr (Q^T applied to the
    first n columns of I), and dormqr applying Q^T to an m x nrhs block (the
    first nrhs columns of one tile column, the rest zeros: the narrow-X entry
    with its default task width; the whole tile column through the plain
    entry beside it), each timed with HIP events, best of `reps`.  Rates use the algorithmic counts
    2mn^2 - 2n^3/3, 4mnk - 2(m+n)k^2 + 4k^3/3 (k = n) and 4 m nrhs k - 2 nrhs k^2.
    Check: Q^T A reproduces R in the top rows and zeros below."""
    import torch

    import hybrid3_amd as hq
    MT, NT = m // nb, n // nb
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    pristine = torch.empty(MT * NT * nb * nb, dtype=torch.float64, device=dev)
    hq.fill_uniform_dev(MT, NT, nb, seed, pristine)
    A = torch.empty_like(pristine)
    TAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device=dev)
    T = torch.zeros(MT * NT * ib * nb, dtype=torch.float64, device=dev)
    Q = torch.empty_like(pristine)
    C = torch.empty(MT * nb * nb, dtype=torch.float64, device=dev)

    def timed(fn, before=None):
        best = float("inf")
        for _ in range(reps + 1):   # the first call builds the schedule
            if before is not None:
                before()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            hq.sync()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) * 1e-3)
        return best

    t_qr = timed(lambda: hq.tile_qr_dev(MT, NT, nb, ib, A, TAU, T), before=lambda: A.copy_(pristine))
    t_q = timed(lambda: hq.tile_orgqr_dev(MT, NT, nb, ib, A, TAU, Q, T=T))
    c0 = pristine[: MT * nb * nb]   # tile column 0 of the original matrix
    t_cw = timed(lambda: hq.tile_ormqr_dev(MT, NT, nb, ib, A, TAU, C, 1, T=T), before=lambda: C.copy_(c0))
    c0n = torch.zeros_like(C).view(MT, nb, nb)   # its first nrhs columns, zeros after them
    c0n[:, :nrhs, :] = c0.view(MT, nb, nb)[:, :nrhs, :]
    t_c = timed(lambda: hq.tile_ormqr_cols_dev(MT, NT, nb, ib, A, TAU, C, 1, nrhs, T=T),
                before=lambda: C.copy_(c0n.view(-1)))
    # Q^T A(:, 0:nrhs) = [R(0:nb, 0:nrhs); 0]: tile (0,0) of the factor, zeros below
    qc = C.view(MT, nb, nb)[:, :nrhs, :]          # [tile row][col][row]
    r00 = A[: nb * nb].view(nb, nb)[:nrhs, :]     # tile (0,0): [col][row]
    upper = torch.tril(torch.ones(nb, nb, dtype=torch.bool, device=dev))[:nrhs, :]   # [col][row]: row <= col
    rmax = r00.abs().max().item()
    top_err = ((qc[0] - r00) * upper).abs().max().item() / rmax
    below = torch.cat([(qc[0] * ~upper).reshape(-1), qc[1:].reshape(-1)]).abs().max().item() / rmax
    pad_zero = C.view(MT, nb, nb)[:, nrhs:, :].abs().max().item() == 0.0
    k = n
    f_qr = 2.0 * m * n * n - 2.0 * n ** 3 / 3.0
    f_q = 4.0 * m * n * k - 2.0 * (m + n) * k * k + 4.0 / 3.0 * k ** 3
    f_c = 4.0 * m * nrhs * k - 2.0 * nrhs * k * k
    out = {"workload": f"tall FP64 tile QR {m}x{n} nb={nb} ib={ib}, dorgqr Q^T array, dormqr on {m}x{nrhs}",
           "qr_ms": round(t_qr * 1e3, 3), "qr_tflops": round(f_qr / t_qr / 1e12, 3),
           "orgqr_ms": round(t_q * 1e3, 3), "orgqr_tflops": round(f_q / t_q / 1e12, 3),
           "ormqr_ms": round(t_c * 1e3, 3), "ormqr_tflops_algorithmic": round(f_c / t_c / 1e12, 3),
           "ormqr_note": f"{nrhs} columns of one tile column, hq_tile_ormqr_cols_dev with width 0 "
                         "(32-column tasks, pipelined chains)",
           "ormqr_whole_tile_ms": round(t_cw * 1e3, 3),
           "check_QtA_top_rel": top_err, "check_QtA_below_rel": below, "check_zero_columns_kept": pad_zero}
    del pristine, A, TAU, T, Q, C, c0n
    torch.cuda.empty_cache()
    return out


def lsq_bench(m=32768, n=8192, kappas=(1e2, 1e6, 1e10), nb=256, ib=32):
    """Least squares on A = U diag(sigma) V^T (U, V from seeded Gaussian QR,
    sigma log-spaced to 1/kappa), b = A x_true + 1e-8 N(0,1): the plain FP64
    tile QR against the mixed ones (FP32 updates; FP16 hi+lo split updates),
    each refined with FP64 residuals by R-preconditioned CGLS.  A mixed
    solve that does not converge (kappa u of its factors beyond 1) falls back
    to the FP64 factorisation, and its time includes both.  A speed-up over
    FP64 is reported only where the solve reached the FP64 solution's
    backward error."""
    import torch

    import hybrid3_amd as hq
    from hybrid3_amd import lsq
    for pv in lsq.PRECISIONS.values():   # schedules and scratch of the full-size factorisations
        hq.tile_qr_prepare((m + nb - 1) // nb, (n + nb - 1) // nb, nb, ib, pv)
    out = []
    for i, kappa in enumerate(kappas):
        A = lsq.svd_matrix(m, n, kappa, seed=300 + i, nb=nb, ib=ib)
        g = torch.Generator(device="cuda").manual_seed(400 + i)
        xt = torch.randn(n, 1, generator=g, dtype=torch.float64, device="cuda")
        b = A @ xt + 1e-8 * torch.randn(m, 1, generator=g, dtype=torch.float64, device="cuda")
        res = {"kappa": kappa}
        if i == 0:
            # full-size buffers and code paths, untimed, with the timed calls'
            # arguments (a one-step warm-up left the first timed solve paying
            # about 30 ms of first-use costs)
            for prec in ("fp64", "fp32", "split"):
                lsq.lstsq(A, b, nb=nb, ib=ib, prec=prec, max_iter=300)
        for prec in ("fp64", "fp32", "split"):
            x, info = lsq.lstsq(A, b, nb=nb, ib=ib, prec=prec, max_iter=300)
            res[prec] = {"factor_ms": round(info["factor_seconds"] * 1e3, 2),
                         "solve_ms": round(info["total_seconds"] * 1e3, 2), "steps": info["steps"],
                         "converged": info["converged"], "fallback_to_fp64": info["fallback"],
                         "backward": info["backward"],
                         "err_vs_x_true": (torch.linalg.norm(x - xt) / torch.linalg.norm(xt)).item()}
        for prec in ("fp32", "split"):
            ok = res[prec]["converged"] and res[prec]["backward"] <= 10 * res["fp64"]["backward"]
            res[f"speedup_{prec}_vs_fp64"] = (round(res["fp64"]["solve_ms"] / res[prec]["solve_ms"], 3)
                                              if ok else None)
        out.append(res)
        note(f"least squares kappa={kappa:g}: fp64 {res['fp64']['solve_ms']} ms, fp32 {res['fp32']['solve_ms']} "
             f"ms ({res['fp32']['converged']}), split {res['split']['solve_ms']} ms ({res['split']['converged']})")
        del A, b
        torch.cuda.empty_cache()
    return {"workload": f"least squares {m}x{n}, A = U diag(sigma) V^T, b = A x_true + 1e-8 N(0,1); "
                        f"tile QR nb={nb} ib={ib} + R-preconditioned CGLS with FP64 residuals",
            "qr_flops": 2.0 * m * n * n - 2.0 * n ** 3 / 3.0, "cases": out}


def batched_lsq_problems(count=512, seed=5000, k=4):
    """count problems m x n, m uniform over the multiples of 256 in [1024,
    8192], n = m/4, kappa log-uniform in [1e1, 1e6], A = U diag(sigma) V^T
    (U, V from seeded Gaussian QR), b = A x_true + 1e-8 N(0,1), k right-hand
    sides; grouped by shape as hybrid3_amd.lsq.batched_lstsq takes them."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    ms = rng.integers(4, 33, count) * 256
    kap = 10.0 ** rng.uniform(1, 6, count)
    groups, xtrue, kappas = {}, {}, {}
    g = torch.Generator(device="cuda").manual_seed(seed)
    for m in sorted(set(ms.tolist())):
        idx = np.nonzero(ms == m)[0]
        n = m // 4
        B = len(idx)
        U = torch.linalg.qr(torch.randn(B, m, n, generator=g, dtype=torch.float64, device="cuda"))[0]
        V = torch.linalg.qr(torch.randn(B, n, n, generator=g, dtype=torch.float64, device="cuda"))[0]
        t = torch.linspace(0, 1, n, dtype=torch.float64, device="cuda")
        lk = torch.from_numpy(np.log10(kap[idx])).cuda()
        sig = 10.0 ** (-lk[:, None] * t[None, :])
        A = torch.bmm(U * sig[:, None, :], V.transpose(1, 2).contiguous())   # (B, m, n)
        del U, V
        xt = torch.randn(B, n, k, generator=g, dtype=torch.float64, device="cuda")
        b = torch.bmm(A, xt) + 1e-8 * torch.randn(B, m, k, generator=g, dtype=torch.float64, device="cuda")
        groups[(m, n)] = (A.transpose(1, 2).contiguous(), b)   # column-major storage
        xtrue[(m, n)] = xt
        kappas[(m, n)] = kap[idx]
        del A
    return groups, xtrue, kappas


def batched_lsq_bench(count=512, nb=256, ib=32):
    """Batched least squares: every problem factored in ONE launch (FP64, or
    FP16 hi+lo split / FP32 trailing updates), refined by R-preconditioned
    CGLS; time to solution and the worst per-problem backward error and
    error against x_true."""
    import torch
    from hybrid3_amd import lsq
    groups, xtrue, _ = batched_lsq_problems(count)
    flops = float(sum(A.shape[0] * (2.0 * m * n * n - 2.0 * n ** 3 / 3.0) for (m, n), (A, _) in groups.items()))
    res = {"workload": f"{count} least-squares problems, m uniform over multiples of 256 in [1024, 8192], n = m/4, "
                       "kappa log-uniform in [1e1, 1e6], 4 right-hand sides, A = U diag(sigma) V^T; one batched "
                       f"tile QR launch (nb={nb}, ib={ib}) + R-preconditioned CGLS with FP64 residuals",
           "qr_flops": flops, "shapes": len(groups)}
    small = {k: (A[:1], b[:1]) for k, (A, b) in list(groups.items())[:2]}
    lsq.batched_lstsq(groups, nb=nb, ib=ib, prec="fp64")   # every shape's code paths, untimed
    for prec in ("fp64", "fp32", "split"):
        lsq.batched_lstsq(small, nb=nb, ib=ib, prec=prec, max_iter=2)   # warm-up
        out, info = lsq.batched_lstsq(groups, nb=nb, ib=ib, prec=prec)
        worst_be, worst_err = 0.0, 0.0
        for sh, (A, b) in groups.items():
            x = out[sh]
            r = b - torch.bmm(A.transpose(1, 2), x)
            atr = torch.linalg.vector_norm(torch.bmm(A, r), dim=1)
            be = atr / (torch.linalg.matrix_norm(A)[:, None] * torch.linalg.vector_norm(r, dim=1))
            worst_be = max(worst_be, be.max().item())
            e = torch.linalg.vector_norm(x - xtrue[sh], dim=1) / torch.linalg.vector_norm(xtrue[sh], dim=1)
            worst_err = max(worst_err, e.max().item())
        res[prec] = {"factor_ms": round(info["factor_seconds"] * 1e3, 2),
                     "solve_ms": round(info["total_seconds"] * 1e3, 2),
                     "factor_tflops": round(flops / info["factor_seconds"] / 1e12, 3),
                     "max_steps": max(info["steps"].values()), "converged": info["converged"],
                     "max_backward": worst_be, "max_err_vs_x_true": worst_err}
        note(f"batched least squares {prec}: {res[prec]['solve_ms']} ms (factor {res[prec]['factor_ms']} ms)")
    for prec in ("fp32", "split"):
        ok = res[prec]["converged"] and res[prec]["max_backward"] <= 10 * res["fp64"]["max_backward"]
        res[f"speedup_{prec}_vs_fp64"] = round(res["fp64"]["solve_ms"] / res[prec]["solve_ms"], 3) if ok else None
    del groups
    torch.cuda.empty_cache()
    return res


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.prec != "fp64" and not args.only_factor:
        print(json.dumps({"error": "--prec selects the profiled factorisation: use it with --only-factor "
                                   "(the headline line is the FP64 factorisation)"}))
        sys.exit(2)
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (no multi-GPU path)"}))
        sys.exit(2)
    import torch

    import hybrid3_amd as hq

    if args.copy_calibration:
        # 2 GiB read + 2 GiB written by one device copy: the reference for
        # the FETCH_SIZE / WRITE_SIZE scale factors
        src = torch.ones(1 << 28, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8}))
        return

    if args.only_panels:
        print(json.dumps(panel_bench(args.nb, args.ib)))
        return
    if args.only_tall:
        print(json.dumps(tall_bench(nb=args.nb, ib=args.ib)))
        return
    if args.only_batched:
        # (--no-check: no per-matrix checks and no per-matrix launches, for
        # profiler runs that should see the batched launches only)
        out = batched_bench(args.batch_count, args.retune, check=not args.no_check)
        out.pop("tuner", None)
        print(json.dumps(out))
        return

    n, nb, ib = args.n, args.nb, args.ib
    factor = {"fp64": hq.tile_qr_dev, "fp32": hq.tile_qr_mixed_dev, "split": hq.tile_qr_split_dev}[args.prec]
    prec_id = {"fp64": 0, "fp32": 1, "split": 2}[args.prec]
    assert n % nb == 0, "n must be a multiple of nb"
    MT = NT = n // nb
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    pristine = torch.empty(MT * NT * nb * nb, dtype=torch.float64, device=dev)
    hq.fill_uniform_dev(MT, NT, nb, args.seed, pristine)
    A = torch.empty_like(pristine)
    TAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device=dev)
    T = torch.zeros(MT * NT * ib * nb, dtype=torch.float64, device=dev)
    ntasks, grid = hq.schedule_info(MT, NT, nb, ib, prec_id)

    def step(ev=None):
        A.copy_(pristine)
        if ev is not None:
            ev[0].record(stream)
        factor(MT, NT, nb, ib, A, TAU, T)
        if ev is not None:
            ev[1].record(stream)

    for _ in range(args.warmup):
        step()
    hq.sync()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(evs[i])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    hq.sync()
    fact_ms = [a.elapsed_time(b) for a, b in evs]
    mean_ms = sum(fact_ms) / len(fact_ms)
    flops = 4.0 / 3.0 * n ** 3
    tflops = flops / (mean_ms * 1e-3) / 1e12
    note(f"tile QR {n}x{n}: {mean_ms:.2f} ms, {tflops:.2f} TFLOP/s")

    # numerical check of the last factorisation (cheap, size-independent):
    # ||A e - Q R e|| via the reflectors is costly at n=16384, so check R's
    # column norms against A's (Q orthogonal => ||R e_j|| = ||A e_j||).
    check = None
    if not args.no_check:
        import numpy as np
        R_diag_ok = bool(torch.isfinite(A).all().item())
        # column norms of A (tile layout) and of R (upper triangle of the factor)
        a = pristine.view(NT, MT, nb, nb)   # [tj][ti][col][row]
        f = A.view(NT, MT, nb, nb)
        tj = torch.arange(NT, device=dev).view(NT, 1, 1, 1)
        ti = torch.arange(MT, device=dev).view(1, MT, 1, 1)
        col = torch.arange(nb, device=dev).view(1, 1, nb, 1)
        row = torch.arange(nb, device=dev).view(1, 1, 1, nb)
        upper = (ti * nb + row) <= (tj * nb + col)
        an = (a * a).sum(dim=(1, 3)).sqrt()
        rn = (f * f * upper).sum(dim=(1, 3)).sqrt()
        colnorm_rel = ((an - rn).abs() / an).max().item()
        check = {"finite": R_diag_ok, "max_col_norm_rel_diff": colnorm_rel}
        if not (R_diag_ok and colnorm_rel < (1e-12 if args.prec == "fp64" else 1e-4)):
            print(json.dumps({"error": "factorisation check failed", "check": check}))
            sys.exit(3)

    # HBM bytes per launch: calibrated FETCH_SIZE + WRITE_SIZE of two
    # rocprofv3 --pmc runs of this same command (profiles/ summary)
    traffic = None
    if os.path.exists(args.traffic_file):
        key = f"{n}_{nb}_{ib}"
        for line in open(args.traffic_file):
            parts = line.split()
            if len(parts) == 2 and parts[0] == key:
                traffic = float(parts[1])

    del pristine, A, TAU, T
    torch.cuda.empty_cache()
    if args.only_factor:
        print(json.dumps({"prec": args.prec, "n": n, "nb": nb, "ib": ib, "ms_factorization": round(mean_ms, 3),
                          "tflops": round(tflops, 4), "check": check}))
        return
    tile_sizes = None if args.no_tile_sizes else tile_size_bench(n, [(384, 32), (512, 32)], args.seed)
    if tile_sizes is not None:
        note("other tile sizes: " + ", ".join(f"nb={c['nb']} ib={c['ib']} {c['tflops']} TFLOP/s" for c in tile_sizes))
        if not all(c["max_col_norm_rel_diff"] < 1e-12 for c in tile_sizes):
            print(json.dumps({"error": "tile-size factorisation check failed", "tile_sizes": tile_sizes}))
            sys.exit(3)
    library = None if args.no_library else rocsolver_reference(n, args.seed)
    if library is not None:
        note(f"rocSOLVER dgeqrf {n}x{n}: {library['ms']} ms, {library['tflops']} TFLOP/s")
    panels = None if args.no_panels else panel_bench(nb, ib)
    if panels is not None:
        note(f"panels: GEQRT {panels['geqrt']['algorithmic_GB_per_s']} GB/s, "
             f"GEQRT+TSQRT {panels['geqrt_tsqrt']['algorithmic_GB_per_s']} GB/s")
    batched = None if args.no_batched else batched_bench(args.batch_count, args.retune)
    if batched is not None:
        note(f"batched: {batched['qrs_per_s']} QRs/s, FP32 matrices {batched['fp32']['qrs_per_s']} QRs/s")
    if batched is not None and not (batched["max_col_norm_rel_diff"] < 1e-12 and batched["max_residual"] < 1e-12
                                    and batched["max_orthogonality"] < 1e-12
                                    and batched["fp32"]["max_col_norm_rel_diff"] < 1e-6
                                    and batched["fp32"]["max_residual"] < 1e-5
                                    and batched["fp32"]["max_orthogonality"] < 1e-5):
        print(json.dumps({"error": "batched factorisation check failed", "batched": batched}))
        sys.exit(3)

    homegrown = None if args.no_homegrown else homegrown_bench()
    if homegrown is not None:
        note("tile_kernel.h calls at 256: " + ", ".join(f"{k} {v['ms']} ms" for k, v in homegrown.items()
                                                          if k != "n"))

    tall = None
    if not args.no_tall:
        tall = tall_bench(nb=nb, ib=ib)
        note(f"tall: QR {tall['qr_tflops']} TFLOP/s, dorgqr {tall['orgqr_tflops']}, dormqr {tall['ormqr_ms']} ms "
             f"(whole tile column {tall['ormqr_whole_tile_ms']} ms)")
        if not (tall["check_QtA_top_rel"] < 1e-12 and tall["check_QtA_below_rel"] < 1e-12
                and tall["check_zero_columns_kept"]):
            print(json.dumps({"error": "tall factorisation check failed", "tall": tall}))
            sys.exit(3)

    lsq_res = None
    if not args.no_lsq:
        try:
            lsq_res = lsq_bench()
        except Exception as e:  # the headline line must still print
            lsq_res = {"error": f"{type(e).__name__}: {e}"}
        note("mixed least squares done")

    blsq = None
    if not args.no_batched_lsq:
        try:
            blsq = batched_lsq_bench(args.batched_lsq_count)
        except Exception as e:  # the headline line must still print
            blsq = {"error": f"{type(e).__name__}: {e}"}

    cpu = None if args.no_cpu_baseline else cpu_baseline(nb, ib, args.seed)
    if cpu is not None:
        note(f"cpu baseline: {cpu['value']} TFLOP/s")
    out = {
        "metric": "QR TFLOP/s FP64 vs mixed+refined at FP64 backward error; batched QRs/sec",
        "value": round(tflops, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round((t1 - t0) * 1e3 / args.steps, 3),
        "ms_factorization": round(mean_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: uniform[-1,1) SplitMix64, seed %d, resident in HBM" % args.seed,
        "config": {"workload": f"plain FP64 flat-tree tile QR, square {n}x{n}, nb={nb}, ib={ib}, "
                               f"persistent on-device DAG scheduler ({ntasks} tasks, {grid} workgroups)",
                   "n": n, "nb": nb, "ib": ib, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 4), "peak": FP64_MFMA_PEAK,
                     "unit": "TFLOP/s", "frac": round(tflops / FP64_MFMA_PEAK, 4), "traffic": traffic,
                     "kernel": "hq::sched_kernel<256,2> (whole factorisation, one launch)",
                     "algorithmic_flops_per_launch": flops},
        "cpu_baseline": cpu,
        "library_reference": library,
        "other_tile_sizes": tile_sizes,
        "homegrown_api": homegrown,
        "panels": panels,
        "check": check,
        "batched": batched,
        "tall": tall,
        "mixed_lsq": lsq_res,
        "batched_lsq": blsq,
        "step_ms": [round(x, 3) for x in fact_ms],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
