// hq_internal.h -- types shared by the HIP kernels and the host runtime of
// the tiled QR engine (not part of the public C-ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Task kinds of the flat-tree tile QR DAG (Hybrid_tile_QR.c:960,1123,1254,1332)
// plus the Q-application kinds used by dorgqr / dormqr.
// GEQRT / TSQRT run as one panel task per ib block (field j = block) and
// one trailing task per block and column slab (j = block, slab), or, in
// batches of many matrices, as one task per tile (j = -1).
enum hq_task_kind : int32_t {
  HQ_GEQRT = 0,    // QR of ib block j (all blocks: j = -1) of diagonal tile (k,k)
  HQ_TSQRT = 1,    // QR of ib block j (all blocks: j = -1) of [R(k,k); A(r,k)]
  HQ_LARFB = 2,    // A(k,j) <- Q(k,k)^T A(k,j)            (one column slab)
  HQ_SSRFB = 3,    // [A(k,j); A(r,j)] <- Q(r,k)^T [...]   (one column slab)
  HQ_GEQRT_U = 4,  // block j of GEQRT(k) applied to slab `slab` of tile (k,k)
  HQ_TSQRT_U = 5,  // block j of TSQRT(k,r) applied to slab `slab` of [R(k,k); A(r,k)]
};

#define HQ_MAXDEP 6

// One node of the device DAG.  Dependencies always have smaller indices in
// the task array (the array is a topological order), which is what makes
// in-order claiming by a persistent kernel deadlock-free.
struct hq_task {
  int32_t kind;
  int32_t k;      // elimination step (panel column)
  int32_t r;      // tile row (TSQRT / SSRFB)
  int32_t j;      // tile column (LARFB / SSRFB)
  int32_t slab;   // column slab: hq_upd_slab(nb, prec) columns of tile column j (LARFB / SSRFB),
                  // 64 columns of the panel tile (GEQRT_U / TSQRT_U)
  int32_t mat;    // matrix of a batch (index into the geometry table; 0 otherwise)
  int32_t ndep;
  int32_t dep[HQ_MAXDEP];
};

// Geometry + base pointers of a tile-layout factorization.
//   tile (i,j):  A   + (i + j*MT) * nb*nb      (column-major, ld = nb)
//   TAU  (i,j):  TAU + (i + j*MT) * nb
//   T    (i,j):  T   + (i + j*MT) * ib*nb      (ld = ib; block b at column b*ib)
struct hq_geom {
  double *A;    // reflector tiles (the factorised matrix RY)
  double *TAU;
  double *T;
  int MT, NT, nb, ib;
  double *X;    // matrix updated by LARFB / SSRFB: A itself while factorising,
                // the Q / C array (MT x XNT tiles) while applying Q^T
  int XNT;
  int prec;     // 0: FP64 everywhere; 1: LARFB / SSRFB in FP32; 2: in FP16 hi+lo splits
                // with FP32 accumulation (panels FP64 in both)
  // Optional (FP64 factorisation, nb >= 128, ib % 16 == 0): for each
  // diagonal tile k, W = V T16 at W + k*nb*nb (T16 = the 16x16 diagonal
  // blocks of T) and its unit lower trapezoidal V at VD + k*nb*nb, written by
  // the GEQRT panels and read by LARFB.  nullptr: LARFB multiplies by T
  // itself.
  double *W;
  double *VD;
  // Split-precision factorisation (prec 2): for each reflector tile (i,j),
  // the FP16 hi / lo images of its 16-reflector sub-blocks as the update
  // kernels stage them in LDS, written by the panels; tile (i,j) at
  // VH + (i + j*MT) * hq_vh_tile(nb, ib) halves, sub-block s at + s*64*nb.
  // nullptr otherwise.
  unsigned short *VH;
  // Split-precision factorisation: per matrix column c (NT nb entries) the
  // power-of-two exponent sh of the scale 2^sh applied to the column in the
  // split updates, with 2^sh ||A(:, c)||_2 < 2^12 (the column's 2-norm is
  // invariant under the orthogonal updates, so it bounds every part of the
  // column at every step).  Written by a pass over A before the
  // factorisation; nullptr: each update task finds a scale from its own
  // columns (two sweeps over them).
  const int *CE;
  // Columns of X (of A while factorising) that hold matrix entries: the rest
  // of the last tile column is zero padding (a batch matrix of n columns);
  // 0: every column.  Padding columns are zero and stay zero, and their
  // reflectors are identities (tau = 0, T = 0), so tasks wholly inside the
  // padding only write those zeros (and the diagonal tile's W = 0 and unit
  // V columns) and skip their arithmetic.
  int ncol;
  // Column width of the LARFB / SSRFB tasks when it is not the tile size's
  // default (hq_upd_slab): 16 for the Q^T application to a narrow X, which
  // runs on the generic kernels with one 16-column strip per workgroup.
  int uw;
  // With uw = 16: per-task progress words, (epoch << 32) | (sub-blocks whose
  // X1 rows are written); an SSRFB whose dep[0] is negative (-2 - task) runs
  // each 16-reflector sub-block once that task has written the same rows of
  // X1.  nullptr otherwise.
  unsigned long long *prog;
};

// halves of split images per tile: one image of 64 nb halves per 16-reflector
// sub-block, ceil(nb/ib) blocks of ceil(ib/16) sub-blocks
inline size_t hq_vh_tile(int nb, int ib) { return (size_t)((nb + ib - 1) / ib) * ((ib + 15) / 16) * 64 * nb; }

// One matrix of a batch for the layout kernels: m x n entries in an MT x NT
// grid of nb x nb tiles (tile layout as hq_geom; entries past m or n are 0).
struct hq_bmat {
  double *A;      // tile layout
  double *C;      // column-major m x n, leading dimension ld (conversions)
  float *Cf;      // the same in FP32 (FP32 conversions)
  int m, n, MT, NT, ld;
  uint64_t seed;  // uniform fill: value(row, col) from (seed, col*m + row)
};

#define HQ_THREADS 256
#define HQ_SLAB 64
// Q^T applied to fewer columns than this runs as 16-column update tasks
// (their pipelined chains beat the 64-column slabs up to a 256-column tile)
#define HQ_NARROW_COLS 257

// Column width of one LARFB / SSRFB task.  FP64 and split updates: 64 (one
// 16-column strip per wave).  FP32 updates: 128 (two strips per wave), so
// the MFMA work of a pass -- half that of FP64 per column -- still covers the
// latency of the sub-block loads issued for the next pass, and each V tile is
// read once per 128 columns.  (Two FP32 strips per wave plus the split
// pipeline's operands exceed the register file at nb = 256.)  The panels'
// trailing tasks keep 64-column slabs in every precision.
__host__ __device__ inline constexpr int hq_upd_slab(int nb, int prec) {
  return (prec == 1 ? 128 : HQ_SLAB) < nb ? (prec == 1 ? 128 : HQ_SLAB) : nb;
}
#define HQ_IBMAX 32
// bytes of the device array behind the allocation-free tile QR (claim counter,
// flags and task list of one geometry)
#define HQ_NOALLOC_SCRATCH (32 << 20)
