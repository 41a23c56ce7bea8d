/*
 * hybrid3_dev.h -- device-pointer entry points and error reporting of the
 * tiled QR engine (libhybrid3.so).
 *
 * These have no counterpart in the reference API (whose functions take host
 * pointers and return void); they are the measured path and the per-kernel
 * boundary used by the parity tests.  Every function returns 0 on success
 * or a nonzero hq error code (also kept in hq_last_error()).  `stream` is a
 * hipStream_t passed as void* (NULL = HIP's default stream); work is
 * queued on it and the call returns without synchronising unless stated.
 * Calls on one device share the library's scratch (claim counter, flags, T /
 * W scratch, batch tables); a call issued on another stream than the
 * previous one waits on the device for the previous call's work, so calls
 * from several streams are ordered rather than racing (prepared batch plans
 * own their scratch and are not ordered this way).
 *
 * ib: any 1 <= ib <= nb.  The panels run with an inner block of at most 32
 * columns; R, V and TAU do not depend on the blocking (results agree to
 * rounding across ib), and a T returned for ib > 32 is rebuilt from V and TAU
 * per ib block (LAPACK dlarft).
 *
 * Tile layout of A / TAU / T (all device memory):
 *   tile (i,j): A   + (i + j*MT)*nb*nb   column-major, ld = nb
 *   TAU  (i,j): TAU + (i + j*MT)*nb
 *   T    (i,j): T   + (i + j*MT)*ib*nb   ld = ib, block b at column b*ib
 */
#ifndef HQ_HYBRID3_DEV_H_
#define HQ_HYBRID3_DEV_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  HQ_OK = 0,
  HQ_ERR_ARG = 1,      /* invalid or unsupported argument */
  HQ_ERR_RANK = 2,     /* my_rank != 0 or Proc_num != 1 */
  HQ_ERR_DEVICE = 3,   /* HIP runtime error */
  HQ_ERR_NOMEM = 4     /* device allocation failed */
};

int hq_last_error(void);
const char *hq_last_error_message(void);

/* Version string of the library build. */
const char *hq_version(void);

/* Flat-tree tile QR of an MT x NT tile grid on the device with the
 * persistent DAG scheduler (one launch).  T may be NULL (the library keeps
 * its own); otherwise it receives every tile's ib x nb T factor. */
int hq_tile_qr_dev(int MT, int NT, int nb, int ib, double *A, double *TAU,
                   double *T, void *stream);

/* Builds and caches the schedule of the factorisation of an MT x NT tile
 * grid with updates in precision prec (0 FP64, 1 FP32, 2 split) and reserves
 * the library scratch it uses (all but T, which a call either passes or
 * reserves itself), so that the calls that follow spend no host time on
 * them.  Optional: the first call of a geometry does the same work. */
int hq_tile_qr_prepare_dev(int MT, int NT, int nb, int ib, int prec);

/* hq_tile_qr_dev for an m x n matrix whose sizes need not be multiples of
 * nb (edge tiles, which the reference's integer division drops,
 * mat_prod.c:24-29): A holds ceil(m/nb) x ceil(n/nb) tiles in the layout
 * above; the entries past row m or column n are set to zero by the call, so
 * the reflectors are zero in the padded rows and the padded columns' taus
 * are zero.  R's leading m x n part is that of the unpadded matrix. */
int hq_tile_qr_mn_dev(int m, int n, int nb, int ib, double *A, double *TAU, double *T,
                      void *stream);

/* Mixed precision (no reference counterpart): the same flat-tree tile QR
 * with GEQRT / TSQRT (reflectors, T factors, R) in FP64 and the trailing
 * updates LARFB / SSRFB computed in FP32 on MFMA with FP32 accumulation
 * (tiles stay FP64 in memory).  R carries FP32-level errors; it is meant as
 * the factorisation inside an iteratively refined least-squares solve.
 * nb in {128, 256}. */
int hq_tile_qr_mixed_dev(int MT, int NT, int nb, int ib, double *A, double *TAU,
                         double *T, void *stream);

/* hq_tile_qr_dev without any device allocation (no counterpart in the
 * reference; for callers that cannot allocate: code under hipGraph capture,
 * serving loops, harnesses that forbid hipMalloc).  The scheduler's claim
 * counter, device epoch, task flags and task list live in a device array of
 * the library of 32 MiB per GPU (held from library load to process end); no
 * hipMalloc, hipFree or torch allocation happens.  T is required (ib x nb per
 * tile, ib <= 32) and LARFB multiplies by T itself, so results agree with
 * hq_tile_qr_dev to rounding (not bit for bit).  Limit: the geometry's task
 * list must fit, e.g. 16384 x 16384 at nb = 256, ib = 32 (399360 tasks, 22 MB);
 * larger ones return HQ_ERR_ARG.  The first call for a geometry copies its task
 * list into the array and synchronises `stream`; later calls with the same
 * geometry only enqueue two kernels (capturable in a hipGraph, the run's
 * epoch being advanced on the device).  All calls on one GPU share the array:
 * calls that overlap on one GPU race.  hq_tile_qr_noalloc_check synchronises
 * `stream` and reports a scheduler time-out of these calls. */
int hq_tile_qr_noalloc_dev(int MT, int NT, int nb, int ib, double *A, double *TAU, double *T,
                           void *stream);
int hq_tile_qr_noalloc_check(void *stream);

/* Split precision (no reference counterpart): as hq_tile_qr_mixed_dev, with
 * the trailing updates LARFB / SSRFB on FP16 MFMA with FP32 accumulation,
 * every product formed as hi*hi + hi*lo + lo*hi of FP16 hi + lo splits (22
 * significant bits), V scaled by 2^10 and each updated column by a power of
 * two chosen from its 2-norm so that every split stays inside FP16's range.
 * Panels (reflectors, T, R) stay FP64.  R carries FP32-level errors.
 * nb in {128, 256}. */
int hq_tile_qr_split_dev(int MT, int NT, int nb, int ib, double *A, double *TAU, double *T,
                         void *stream);

/* Same factorisation, one kernel launch per task in the scheduler's order
 * (stream order replaces the device dependency flags). */
int hq_tile_qr_dev_serial(int MT, int NT, int nb, int ib, double *A,
                          double *TAU, double *T, void *stream);

/* hq_tile_qr_dev with a per-task timeline: trace (device, 4*ntasks int64)
 * receives for task i (in claim order) the device wall-clock ticks at claim,
 * start (dependencies satisfied) and end, and the workgroup index.  The
 * counterpart of the reference's CRAYJ_TIMING / CRAYJ_TIMELINE builds. */
int hq_tile_qr_dev_traced(int MT, int NT, int nb, int ib, double *A,
                          double *TAU, double *T, long long *trace,
                          void *stream);

/* The same timeline for any precision of the trailing updates: prec 0
 * (FP64, = hq_tile_qr_dev_traced), 1 (hq_tile_qr_mixed_dev), 2
 * (hq_tile_qr_split_dev). */
int hq_tile_qr_traced_prec_dev(int MT, int NT, int nb, int ib, int prec, double *A, double *TAU,
                               double *T, long long *trace, void *stream);

/* Task list in claim order: out[5*i..5*i+4] = kind (0 GEQRT, 1 TSQRT,
 * 2 LARFB, 3 SSRFB), step k, row r, column j, slab.  capacity >= 5*ntasks. */
int hq_schedule_tasks(int MT, int NT, int nb, int ib, int *out, int capacity);

/* The same for the factorisation whose trailing updates run in precision
 * prec (0 FP64, 1 FP32, 2 split): reduced-precision LARFB / SSRFB tasks
 * cover 128 columns (slab = 128-column slab), FP64 ones 64. */
int hq_schedule_tasks_prec(int MT, int NT, int nb, int ib, int prec, int *out, int capacity);

/* Host-only: the DAG the device scheduler runs, in claim order (no GPU
 * needed).  XNT < 1: factorisation; XNT >= 1: Q^T application to MT x XNT
 * tiles.  P = number of workers of the list-scheduling simulation that
 * orders the tasks.  tasks[5*i..] as hq_schedule_tasks; deps[6*i..6*i+5] =
 * claim indices of the task's dependencies, -1 padded.  With tasks == NULL
 * only *ntasks is set. */
int hq_dag_host(int MT, int NT, int nb, int ib, int XNT, int P, int *tasks,
                int *deps, int capacity, int *ntasks);
/* Host-only: the task list of a Q^T application to MT x XNT tiles on X's
 * first ncol columns (0: all) with width 16 / 32 (the narrow tasks of
 * hq_tile_ormqr_cols_dev) or 0 (the factorisation's slabs), in claim order
 * for P workers, exactly as the device scheduler receives it: tasks as
 * hq_dag_host; deps[6*i..] = claim indices of the dependencies, and a
 * pipelined predecessor (the previous SSRFB of a narrow chain, waited for
 * sub-block by sub-block inside the task) as -2 - its index; -1 padded. */
int hq_dag_host_apply(int MT, int NT, int nb, int ib, int XNT, int ncol, int width, int P,
                      int *tasks, int *deps, int capacity, int *ntasks);

/* hq_dag_host of the factorisation with updates in precision prec. */
int hq_dag_host_prec(int MT, int NT, int nb, int ib, int prec, int XNT, int P, int *tasks,
                     int *deps, int capacity, int *ntasks);

/* Frequency of the device wall clock used by the trace, in kHz. */
int hq_wall_clock_khz(int *khz);

/* Schedule of the DAG for a geometry: number of tasks and the number of
 * persistent workgroups a launch uses. */
int hq_schedule_info(int MT, int NT, int nb, int ib, int *ntasks, int *grid);
/* hq_schedule_info of the factorisation with updates in precision prec. */
int hq_schedule_info_prec(int MT, int NT, int nb, int ib, int prec, int *ntasks, int *grid);

/* PLASMA-style single-tile kernels on device pointers (1 <= nb <= 512,
 * 1 <= ib <= nb (LARFB / SSRFB at nb other than 128 and 256: ib <= 32),
 * every leading dimension = nb, T leading dimension = ib):
 *   GEQRT = CORE_dgeqrt(nb,nb,ib,A,nb,T,ib,TAU)        (Hybrid_tile_QR.c:985)
 *   TSQRT = CORE_dtsqrt(nb,nb,ib,A1,nb,A2,nb,T,ib,TAU) (Hybrid_tile_QR.c:1135)
 *   LARFB = CORE_dormqr(Left,Trans,nb,nb,nb,ib,V,nb,T,ib,C,nb) (:1276)
 *   SSRFB = CORE_dtsmqr(Left,Trans,nb,nb,nb,nb,nb,ib,A1,nb,A2,nb,V,nb,T,ib) (:1356) */
int hq_dgeqrt_dev(int nb, int ib, double *A, double *T, double *TAU, void *stream);
int hq_dtsqrt_dev(int nb, int ib, double *A1, double *A2, double *T,
                  double *TAU, void *stream);
int hq_dormqr_dev(int nb, int ib, const double *V, const double *T, double *C,
                  void *stream);
int hq_dtsmqr_dev(int nb, int ib, double *A1, double *A2, const double *V,
                  const double *T, void *stream);

/* T factors (ib x nb per tile, layout above) of every reflector tile (r,k),
 * r >= k, rebuilt from RY and TAU (LAPACK dlarft per ib block); the last
 * reflector of each diagonal tile is dropped, as Hybrid_dorgqr.c:83 does. */
int hq_tile_build_t_dev(int MT, int NT, int nb, int ib, const double *RY,
                        const double *TAU, double *T, void *stream);

/* X (MT x XNT tiles) <- Q^T X with Q = H(0,0) H(1,0) ... from the tile QR
 * in RY / TAU (MT x NT tiles); T may be NULL (rebuilt).  The sweep of
 * Hybrid_dormqr.c:64-167 on the device DAG scheduler.  An X of at most 256
 * columns (nb a multiple of 16) runs as the narrow tasks of
 * hq_tile_ormqr_cols_dev (width 0). */
int hq_tile_ormqr_dev(int MT, int NT, int nb, int ib, const double *RY,
                      const double *TAU, const double *T, double *X, int XNT,
                      void *stream);

/* hq_tile_ormqr_dev for an X whose first ncol columns hold entries
 * ((XNT-1)*nb < ncol <= XNT*nb) and whose other columns hold zeros (they
 * still do on return): the C argument of Hybrid_dormqr (Hybrid_dormqr.c:
 * 20-171) when it is a short block of right-hand sides.  Tasks that would
 * update only zero columns are left out.  width: 16 runs the updates as
 * 16-column tasks (one chain per 16 columns along the flat tree, each SSRFB
 * following the previous one 16 reflectors at a time), 32 as 32-column
 * tasks of the same kind (nb <= 256), 64 as the factorisation's update
 * slabs; 0 picks 16 up to 32 columns, 32 up to 256, else 64.  Errors:
 * HQ_ERR_ARG for ncol outside X's last tile column, another width, or 32
 * with nb above 256. */
int hq_tile_ormqr_cols_dev(int MT, int NT, int nb, int ib, const double *RY,
                           const double *TAU, const double *T, double *X,
                           int XNT, int ncol, int width, void *stream);

/* hq_tile_ormqr_cols_dev with the per-task timeline of hq_tile_qr_dev_traced
 * (trace: device, 4*ntasks int64; tasks in claim order). */
int hq_tile_ormqr_cols_traced_dev(int MT, int NT, int nb, int ib,
                                  const double *RY, const double *TAU,
                                  const double *T, double *X, int XNT,
                                  int ncol, int width, long long *trace,
                                  void *stream);

/* Q (MT x NT tiles) <- Q^T (the array Hybrid_dorgqr returns, Hybrid_dorgqr.c:
 * 56-138): identity, then hq_tile_ormqr_dev. */
int hq_tile_orgqr_dev(int MT, int NT, int nb, int ib, const double *RY,
                      const double *TAU, const double *T, double *Q,
                      void *stream);

/* Hybrid_make_Q semantics (Hybrid_make_Q.c:56-237): ones on the diagonal of
 * Q (other entries kept), Q <- Q^T-apply, then RY's reflectors zeroed. */
int hq_tile_make_q_dev(int MT, int NT, int nb, int ib, double *RY,
                       const double *TAU, double *Q, void *stream);

/* A (tile layout, MT x NT tiles of nb) <- uniform[-1,1) from SplitMix64 of
 * (seed, column-major global index). */
int hq_fill_uniform_dev(int MT, int NT, int nb, uint64_t seed, double *A,
                        void *stream);

/* ---- Batches of independent matrices (no counterpart in the reference,
 * which factors one matrix per call; for callers with thousands of small
 * problems).  Arrays of length `count` are host arrays; the pointers in them
 * are device pointers.  A matrix of m x n entries lives in tile layout in
 * ceil(m/nb) x ceil(n/nb) tiles with the entries past row m / column n
 * zero: Householder reflectors of a zero-padded matrix are zero in the
 * padded rows, so R's leading n x n block is that of the unpadded matrix.
 * count <= 65535. */

/* Flat-tree tile QR of every matrix of the batch in ONE persistent launch:
 * matrix i is an MT[i] x NT[i] grid of nb x nb tiles at A[i], TAU[i]
 * (MT[i]*NT[i]*nb doubles), T[i] (MT[i]*NT[i]*ib*nb doubles, or T == NULL /
 * T[i] == NULL for library scratch).  Same results as hq_tile_qr_dev on
 * each matrix. */
int hq_batched_tile_qr_dev(int count, int nb, int ib, const int *MT, const int *NT,
                           double *const *A, double *const *TAU, double *const *T,
                           void *stream);

/* hq_batched_tile_qr_dev with the trailing updates of every matrix in FP64
 * (prec 0), FP32 (1, as hq_tile_qr_mixed_dev) or FP16 hi+lo splits (2, as
 * hq_tile_qr_split_dev); nb >= 128 for prec 1 and 2. */
int hq_batched_tile_qr_prec_dev(int prec, int count, int nb, int ib, const int *MT, const int *NT,
                                double *const *A, double *const *TAU, double *const *T, void *stream);

/* hq_batched_tile_qr_prec_dev with a choice of panel mapping and the
 * matrices' column counts (no counterpart in the reference; the batch
 * autotuner's wave-to-task dimension):
 * split_panels 0 runs each GEQRT / TSQRT of a tile as one task (one
 * workgroup factors the whole nb-column panel and applies it to its own
 * tile), 1 as one task per ib block plus one trailing task per (block,
 * 64-column slab), spreading a panel over several workgroups as the
 * single-matrix entry points do.  Same results either way (same reflectors,
 * same update order per tile).
 * ncol (NULL: none): ncol[i] columns of matrix i hold entries and the rest of
 * its last tile column is zero padding ((NT[i]-1)*nb < ncol[i] <= NT[i]*nb);
 * tasks wholly inside the padding then only write the zeros their outputs
 * hold (TAU, T), and the results are bitwise those of the full run. */
int hq_batched_tile_qr_ex_dev(int prec, int split_panels, int count, int nb, int ib, const int *MT,
                              const int *NT, const int *ncol, double *const *A, double *const *TAU,
                              double *const *T, void *stream);

/* Builds and caches the schedule of a batch (the shapes, nb, ib and update
 * precision of a later hq_batched_tile_qr_prec_dev call) ahead of it;
 * _ex: also the panel mapping of a later hq_batched_tile_qr_ex_dev call. */
int hq_batched_prepare_dev(int prec, int count, int nb, int ib, const int *MT, const int *NT);
int hq_batched_prepare_ex_dev(int prec, int split_panels, int count, int nb, int ib, const int *MT,
                              const int *NT);

/* The same factorisation as a plan prepared once and run many times (no
 * counterpart in the reference; for callers that capture the launch in a
 * hipGraph and replay it, e.g. a serving loop over fixed buffers).
 * hq_batched_plan_dev copies the task list and the per-matrix geometry to the
 * device and allocates the plan's scratch, flags and counters; it records the
 * pointers, which must stay valid while the plan is used.
 * hq_batched_run_dev enqueues a counter reset (which advances the plan's
 * device-side epoch) and the persistent scheduler on `stream`, allocates
 * nothing and copies nothing, so it may be captured and replayed; each run
 * factors the matrices in place at the recorded pointers.  Runs of one plan
 * must not overlap.  hq_batched_plan_check synchronises `stream` and reports
 * a scheduler time-out of the plan; hq_batched_plan_free releases it. */
int hq_batched_plan_dev(int count, int nb, int ib, const int *MT, const int *NT,
                        double *const *A, double *const *TAU, double *const *T,
                        int *plan);
/* hq_batched_plan_dev with the panel mapping and column counts of
 * hq_batched_tile_qr_ex_dev. */
int hq_batched_plan_ex_dev(int split_panels, int count, int nb, int ib, const int *MT, const int *NT,
                           const int *ncol, double *const *A, double *const *TAU, double *const *T, int *plan);
int hq_batched_run_dev(int plan, void *stream);
int hq_batched_plan_check(int plan, void *stream);
int hq_batched_plan_free(int plan);

/* A[i] (tile layout, padding zero) <- m[i] x n[i] uniform[-1,1) values:
 * entry (row, col) = SplitMix64 of (seeds[i], col*m[i] + row), the values of
 * a column-major fill of the unpadded matrix. */
int hq_batched_fill_uniform_dev(int count, int nb, const int *m, const int *n,
                                const uint64_t *seeds, double *const *A, void *stream);

/* Column-major (leading dimension ldc[i]) <-> tile layout conversions
 * (mat_prod.h:57-67 layout, the reference's mat_prod.c:260-281 copies),
 * zero padding written by the conversion into tile layout. */
int hq_cm_to_tile_dev(int count, int nb, const int *m, const int *n, const double *const *C,
                      const int *ldc, double *const *A, void *stream);
int hq_tile_to_cm_dev(int count, int nb, const int *m, const int *n, const double *const *A,
                      double *const *C, const int *ldc, void *stream);
/* The same for FP32 matrices (a batch of FP32 problems stored as FP32): the
 * tile layout the kernels factor is FP64, converted on the way in and out. */
int hq_cm32_to_tile_dev(int count, int nb, const int *m, const int *n, const float *const *C,
                        const int *ldc, double *const *A, void *stream);
int hq_tile_to_cm32_dev(int count, int nb, const int *m, const int *n, const double *const *A,
                        float *const *C, const int *ldc, void *stream);

/* Bound of the scheduler's dependency waits on this device, in polls of
 * about a microsecond (0 = the default 2^22, seconds).  A wait that reaches
 * it raises the device error word; every workgroup then drains without
 * further work and the next hq_sync (or host-pointer call) reports
 * HQ_ERR_DEVICE.  A short bound is a watchdog for callers that prefer an
 * early error to a long stall. */
int hq_set_wait_limit(unsigned spins);

/* Synchronise `stream`; reports a scheduler timeout or a panel wait that
 * gave up on the device (and clears both). */
int hq_sync(void *stream);

/* Release every device buffer the library holds. */
void hq_finalize(void);

#ifdef __cplusplus
}
#endif
#endif
