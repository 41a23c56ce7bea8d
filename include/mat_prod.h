/*
 * mat_prod.h -- matrix geometry and tile layout of the tiled QR engine.
 *
 * Same type, field meanings and tile-addressing formulas as the reference's
 * mat_prod.h:12-67, so a caller that builds a mat_stat_t and indexes tiles
 * with these macros works unchanged.  Layout: tile (i,j) of the local tile
 * grid starts at (i + j*my_tile_M)*tile_M*tile_N, column-major inside the
 * tile; the TAU vector of tile (i,j) starts at (i + j*my_tile_M)*tile_N.
 */
#ifndef HQ_MAT_PROD_H_
#define HQ_MAT_PROD_H_

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mstatus {
  unsigned long int mat_M;    /* global rows */
  unsigned long int mat_N;    /* global columns */
  unsigned long int my_mat_M; /* rows held by this rank (all of them) */
  unsigned long int my_mat_N; /* columns held by this rank */
  unsigned int tile_M;        /* rows of a tile (nb) */
  unsigned int tile_N;        /* columns of a tile (nb) */
  unsigned int glob_tile_M;   /* tile rows of the whole matrix (MT) */
  unsigned int glob_tile_N;   /* tile columns of the whole matrix (NT) */
  unsigned int my_tile_M;     /* tile rows held by this rank */
  unsigned int my_tile_N;     /* tile columns held by this rank */
} mat_stat_t;

/* Replaces mat_prod.c:15-37.  MT = mat_m / tile_m and NT = mat_n / tile_n
 * (integer division, as the reference); tile columns are dealt cyclically
 * over proc_num ranks.  The engine itself runs with proc_num == 1. */
void mat_init(mat_stat_t *status, long int mat_m, long int mat_n, int tile_m,
              int tile_n, int my_rank, int proc_num);

/* Replaces mat_make (mat_prod.c:157-182): the reference's input generator,
 * glibc rand()/RAND_MAX after srand(seed + my_rank), written in the
 * reference's loop order (tile column, column in tile, tile row, row in
 * tile), so the same seed gives the reference's matrix bit for bit. */
void mat_make(mat_stat_t *status, double *matA, unsigned int seed, int my_rank);

/* Replaces mat_remv (mat_prod.c:184-207): subtracts the matrix mat_make
 * generates with the same seed (A - QR in main.c:253). */
void mat_remv(mat_stat_t *status, double *matA, unsigned int seed, int my_rank);

/* Replaces mat_comp3 (mat_prod.c:115-154): the 1-norm (largest column sum
 * of |E|) of a column-major my_mat_M x my_mat_N matrix; NaN if a column sum
 * is NaN.  One rank: my_rank 0, proc_num 1 (else NaN and hq_last_error). */
double mat_comp3(mat_stat_t *status, double *matE, int my_rank, int proc_num);

/* Replaces mat_dprint (mat_prod.c:260-281): tile layout -> column-major. */
void mat_dprint(mat_stat_t *status, double *out, double *mat);

/* Replaces mat_print (mat_prod.c:209-258): the matrix as text, one matrix
 * row per line, after a "[    0] head" line. */
void mat_print(mat_stat_t *status, FILE *out, double *mat, const char *head_string);

#define mat_mysize(st) ((unsigned long int)((st)->my_mat_M * (st)->my_mat_N))
#define mat_myTAU(st) ((unsigned long int)(((st)->my_mat_M / 2) * ((st)->my_mat_N / 2)))
#define mat_TileSize(st) ((unsigned int)((st)->tile_M * (st)->tile_N))
#define mat_getTileM(st) ((st)->tile_M)
#define mat_getTileN(st) ((st)->tile_N)
#define mat_getM(st) ((st)->mat_M)
#define mat_getN(st) ((st)->mat_N)
#define mat_getMyM(st) ((st)->my_mat_M)
#define mat_getMyN(st) ((st)->my_mat_N)
#define mat_getGTileM(st) ((st)->glob_tile_M)
#define mat_getGTileN(st) ((st)->glob_tile_N)
#define mat_getLTileM(st) ((st)->my_tile_M)
#define mat_getLTileN(st) ((st)->my_tile_N)

/* element (si,sj) of tile (ti,tj) */
#define mat_getPosID(st, ti, tj, si, sj)                                     \
  ((((ti) + (tj) * (st)->my_tile_M) * (st)->tile_M * (st)->tile_N) + (si) + \
   (sj) * (st)->tile_M)
/* first element of tile (i,j) */
#define mat_getMatPos(st, i, j) \
  (((i) + (j) * (st)->my_tile_M) * ((unsigned int)((st)->tile_M * (st)->tile_N)))
/* first TAU entry of tile (i,j) */
#define mat_getVecPos(st, i, j) (((i) + (j) * (st)->my_tile_M) * ((unsigned int)(st)->tile_N))

#ifdef __cplusplus
}
#endif
#endif
