/*
 * Hybrid_tile_QR.h -- drop-in entry points of the tiled QR engine.
 *
 * Signatures are those of the reference's Hybrid_tile_QR.h:26-77; the
 * arithmetic runs on one MI355X (gfx950).  Pointers are HOST pointers (the
 * call copies to the device, runs, and copies back), exactly as the
 * reference's callers pass them (main.c:148,175).  my_rank must be 0 and
 * Proc_num 1: the engine runs on one GPU with no ranks; any other value is
 * rejected (hq_last_error() != 0, arrays untouched).  Errors never abort the
 * process; query hq_last_error() / hq_last_error_message() (hybrid3_dev.h).
 *
 * Supported geometry: tile_M == tile_N == nb, 1 <= nb <= 512 (64, 128 and
 * 256 run the specialised MFMA kernels, other sizes the generic ones),
 * 1 <= ib <= nb (main.c:85-88 defaults ib to nb); the matrix is
 * glob_tile_M x glob_tile_N tiles (any glob_tile_M >= 1, glob_tile_N >= 1,
 * tall and wide included).  As in the reference, mat_init's integer
 * division drops rows / columns past the last whole tile; hq_tile_qr_mn_dev
 * (hybrid3_dev.h) factors such matrices with edge tiles.
 * The host-pointer entry points run one at a time per device (a call waits
 * for another thread's call to finish); a device fault (a timed-out wait on
 * the GPU) is reported through hq_last_error().
 */
#ifndef HQ_HYBRID_TILE_QR_H_
#define HQ_HYBRID_TILE_QR_H_

#include "mat_prod.h"
#include "tile_kernel.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces Hybrid_tile_QR (Hybrid_tile_QR.h:26-30, Hybrid_tile_QR.c:307-1589):
 * in-place flat-tree tile QR.  On return the tiles on and above the diagonal
 * hold R, the tiles below hold the TSQRT reflectors V2, the diagonal tiles
 * hold the GEQRT reflectors below their diagonal, and TAU holds the scalar
 * factors (LAPACK dlarfg convention, as PLASMA's CORE_dgeqrt/CORE_dtsqrt). */
void Hybrid_tile_QR(mat_stat_t *status, double *A, double *TAU, int my_rank,
                    int Proc_num, int ib);

/* Replaces Hybrid_make_Q (Hybrid_tile_QR.h:32-37, Hybrid_make_Q.c:56-237):
 * Q (zero on entry except what the caller put there) <- Q^T in tile layout;
 * the reflectors in RY are zeroed so RY keeps R. */
void Hybrid_make_Q(mat_stat_t *status, double *Q, double *RY, double *TAU,
                   int my_rank, int Proc_num, int ib);

/* Replaces Hybrid_dorgqr (Hybrid_tile_QR.h:62-68, Hybrid_dorgqr.c:22-192):
 * Q <- Q^T in tile layout; RY and TAU are not modified. */
void Hybrid_dorgqr(mat_stat_t *status, double *Q, double *RY, double *TAU,
                   int my_rank, int Proc_num);

/* Replaces Hybrid_dormqr (Hybrid_tile_QR.h:71-77, Hybrid_dormqr.c:20-171):
 * C <- Q^T C, C on the same tile grid as RY. */
void Hybrid_dormqr(mat_stat_t *status, double *RY, double *TAU, double *C,
                   int my_rank, int Proc_num);

/* Replaces Hybrid_make_I (Hybrid_tile_QR.h:40-45, Hybrid_make_I.c:20-53):
 * Id <- identity in tile layout. */
void Hybrid_make_I(mat_stat_t *status, double *Id, int my_rank, int Proc_num);

/* Replaces Hybrid_dgemm (Hybrid_tile_QR.h:55-59, Hybrid_dgemm.c:16-70):
 * C <- A^T B on the tile layout (square grids). */
void Hybrid_dgemm(mat_stat_t *status, double *A, double *B, double *C,
                  int my_rank, int Proc_num);

#ifdef __cplusplus
}
#endif
#endif
