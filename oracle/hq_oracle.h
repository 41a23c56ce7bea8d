/*
 * hq_oracle.h -- CPU oracle for the tiled Householder QR path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links,
 * loads or calls this code; it is imported by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and
 * there only as the checker (or as the timed CPU baseline).
 *
 * Two families of routines are restated here in plain C:
 *
 *  1. The PLASMA-convention tile kernels that the reference's factorization
 *     calls (Hybrid_tile_QR.c:985,1135,1276,1356): GEQRT = CORE_dgeqrt,
 *     TSQRT = CORE_dtsqrt, LARFB = CORE_dormqr(Left,Trans),
 *     SSRFB = CORE_dtsmqr(Left,Trans).  PLASMA 2.x is a third-party
 *     dependency that is not vendored; its published algorithm (the LAPACK
 *     dgeqr2/dlarft/dlarfb/dlarfg building blocks, blocked by ib) is restated.
 *     Pinned against LAPACK dgeqrt/dtpqrt/dgemqrt/dtpmqrt (golden fixtures
 *     under tests/golden, produced by scipy.linalg.lapack).
 *
 *  2. The homegrown tile_kernel.h API (tile_kernel.c:19-716), whose sign
 *     convention (beta >= 0) and lower-triangular T differ from PLASMA.
 *     Pinned by invariants and closed forms (tests/test_oracle_homegrown.py).
 *
 * plus the tile-layout drivers (tile QR, make_Q, dorgqr, dormqr), the
 * TileSch-style dynamic scheduler used as the CPU baseline, and the input
 * generators.
 *
 * Tile layout (mat_prod.h:57-67): tile (i,j) of an MT x NT grid of nb x nb
 * tiles starts at (i + j*MT)*nb*nb and is column-major inside the tile;
 * TAU of tile (i,j) starts at (i + j*MT)*nb.
 */
#ifndef HQ_ORACLE_H_
#define HQ_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- PLASMA convention (LAPACK dlarfg: beta = -sign(alpha)*||.||) ---- */
void orc_dlarfg(int n, double *alpha, double *x, int incx, double *tau);
void orc_dgeqrt(int m, int n, int ib, double *A, int lda, double *T, int ldt,
                double *tau);
void orc_dtsqrt(int m, int n, int ib, double *A1, int lda1, double *A2,
                int lda2, double *T, int ldt, double *tau);
void orc_dormqr_lt(int m, int n, int k, int ib, const double *V, int ldv,
                   const double *T, int ldt, double *C, int ldc);
void orc_dtsmqr_lt(int m1, int n1, int m2, int n2, int k, int ib, double *A1,
                   int lda1, double *A2, int lda2, const double *V, int ldv,
                   const double *T, int ldt);
void orc_dlarft_fc(int n, int k, const double *V, int ldv, const double *tau,
                   double *T, int ldt);
/* PLASMA-style T of a triangle-on-square reflector block (V = [I; V2]). */
void orc_dtsrft(int m, int n, int ib, const double *V, int ldv,
                const double *tau, double *T, int ldt);

/* ---- homegrown tile_kernel.h convention (beta >= 0, T lower) ---- */
void orc_tile_dgeqrt(int row, int col, int block, double *A, int lda,
                     double *TAU, double *Y, double *T, int ldt, double *work);
void orc_tile_dtsqrt(int row, int col, int block, double *A1, int lda1,
                     double *A2, int lda2, double *TAU, double *T, int ldt,
                     double *work);
void orc_tile_dlarfb(int row, int col, double *A, int lda, const double *Y,
                     const double *T, int ldt, double *work);
void orc_tile_dssrfb(int row, int col, double *A1, double *A2, int lda,
                     const double *Y, const double *T, int ldt, double *work);
void orc_tile_dlarft(int row, int col, int block, const double *Y, int lda,
                     const double *TAU, double *T, int ldt, double *work);
void orc_tile_sub_dgeqr2(int row, int col, double *A, int lda, double *TAU,
                         double *Y, double *work);
void orc_tile_sub_dtsqr2(int row, int col, double *A1, int lda1, double *A2,
                         int lda2, double *TAU, double *work);
void orc_tile_sub_dlarft(int row, int col, const double *Y, int lda,
                         const double *TAU, double *T, int ldt, double *work);
void orc_tile_sub_dlarfb(int row, int Y_col, int A_col, const double *Y,
                         int ldy, const double *T, int ldt, double *A, int lda,
                         double *work);
void orc_tile_sub_dssrfb(int over_row, int under_row, int Y_col, int A_col,
                         const double *Y, int ldy, const double *T, int ldt,
                         double *A1, int lda1, double *A2, int lda2,
                         double *work);

/* ---- tile-layout drivers ---- */
/* Flat-tree tile QR in place (Hybrid_tile_QR semantics, one rank).
 * T (optional, may be NULL) receives the ib x nb T factor of every tile
 * (i,j), i >= j, at (i + j*MT)*ib*nb, leading dimension ib.
 * nthreads <= 1 runs the sequential right-looking order; otherwise the
 * TileSch-style dynamic DAG with that many OpenMP threads. */
int orc_tile_qr(int MT, int NT, int nb, int ib, double *A, double *TAU,
                double *T, int nthreads);
/* Hybrid_make_Q semantics: Q (zeroed by the caller) <- Q^T in tile layout,
 * RY's reflectors are zeroed (RY keeps R). */
void orc_make_Q(int MT, int NT, int nb, double *Q, double *RY,
                const double *TAU);
/* Hybrid_dorgqr semantics: Q <- Q^T in tile layout via the homegrown
 * kernels (RY, TAU unchanged). */
void orc_dorgqr(int MT, int NT, int nb, double *Q, const double *RY,
                const double *TAU);
/* Hybrid_dormqr semantics: C <- Q^T C, C on the same tile grid as RY. */
void orc_dormqr(int MT, int NT, int nb, const double *RY, const double *TAU,
                double *C);

/* ---- generators / layout ---- */
/* Counter-based uniform[-1,1): SplitMix64 of (seed, index). */
double orc_uniform_pm1(uint64_t seed, uint64_t index);
/* Fill an m x n column-major matrix (ld = m) with orc_uniform_pm1. */
void orc_fill_uniform_cm(int m, int n, uint64_t seed, double *A);
/* Reference-compatible generator (mat_prod.c:157-182): glibc rand()/RAND_MAX
 * in tile-layout order, seeded with srand(seed). */
void orc_mat_make_glibc(int MT, int NT, int nb, unsigned int seed, double *A);
/* column-major <-> tile layout */
void orc_cm_to_tile(int MT, int NT, int nb, const double *cm, double *tile);
void orc_tile_to_cm(int MT, int NT, int nb, const double *tile, double *cm);

/* Elapsed wall seconds of a tile QR on nthreads threads (CPU baseline). */
double orc_time_tile_qr(int MT, int NT, int nb, int ib, uint64_t seed,
                        int nthreads, int reps);

#ifdef __cplusplus
}
#endif
#endif
