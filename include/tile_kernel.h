/*
 * tile_kernel.h -- the homegrown tile-kernel API of the reference
 * (tile_kernel.h:28-162), served by HIP kernels on one MI355X.
 *
 * Conventions are those of the reference's tile_kernel.c, not LAPACK's:
 *  - reflectors: beta = +sqrt(alpha^2 + ||x||^2) (R's diagonal >= 0),
 *    v = [1; x/(alpha - beta)], TAU = 2 (alpha-beta)^2 / ((alpha-beta)^2 + ||x||^2);
 *    for a square tile the last column is not reflected (Y(last,last) = 0,
 *    TAU(last) untouched);
 *  - Y is the explicit unit-lower reflector matrix (the caller zeroes its
 *    strictly upper part);
 *  - T is LOWER triangular, T = T_LAPACK^T, and its strictly upper part must
 *    be zero on entry to tile_dlarft.
 * Pointers are HOST pointers; `work` is accepted for signature compatibility
 * (tile_init still allocates it) and is not used by the device path.
 * Square tiles only (row == col), as every reference caller uses them.
 */
#ifndef HQ_TILE_KERNEL_H_
#define HQ_TILE_KERNEL_H_

#ifdef __cplusplus
extern "C" {
#endif

/* tile_kernel.h:28 -- calloc(3*tile_row*tile_col doubles); free() it. */
void *tile_init(char type, int tile_row, int tile_col);

/* tile_kernel.h:31-43 -- QR of a tile, ib-style blocking by `block`. */
void tile_dgeqrt(const int row, const int col, const int block, double *A,
                 const int lda, double *TAU, double *Y, double *T,
                 const int ldt, double *work);

/* tile_kernel.h:46-59 -- QR of [triu(A1); A2]; A2 <- V2. */
void tile_dtsqrt(const int row, const int col, const int block, double *A1,
                 const int lda1, double *A2, const int lda2, double *TAU,
                 double *T, const int ldt, double *work);

/* tile_kernel.h:61-71 -- A <- A - Y (T (tril(Y)^T A)). */
void tile_dlarfb(const int row, const int col, double *A, const int lda,
                 const double *Y, const double *T, const int ldt, double *work);

/* tile_kernel.h:73-84 -- W = A1 + Y^T A2; W = T W; A1 -= W; A2 -= Y W. */
void tile_dssrfb(const int row, const int col, double *A1, double *A2,
                 const int lda, const double *Y, const double *T,
                 const int ldt, double *work);

/* tile_kernel.h:151-162 -- lower T from Y and TAU (block < 2: full
 * recursion; 2 <= block < col: combines the diagonal blocks already in T,
 * with the reference's loop bounds; block >= col: no-op). */
void tile_dlarft(const int row, const int col, const int block,
                 const double *Y, const int lda, const double *TAU, double *T,
                 const int ldt, double *work);

#ifdef __cplusplus
}
#endif
#endif
