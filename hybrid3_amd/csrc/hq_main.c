/*
 * hq_main.c -- the reference's driver program (main.c:19-277) on the GPU
 * engine: `hybrid3 N NB [IB]` generates the reference's input (mat_make,
 * seed 20151130), factors it with Hybrid_tile_QR, forms Q with Hybrid_make_Q
 * and prints the reference's lines: time and GFLOP/s of the factorisation
 * (main.c:159-163), the make_Q time (:188), ||Q||oo (:197), ||I-Q'Q||oo
 * (:230) and ||A-QR||oo (:258), each the unnormalised 1-norm of mat_comp3,
 * then the time of the norm computation (:265).  One process, no MPI: the
 * node count printed is 1.  Plain C against the library's headers only.
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "Hybrid_tile_QR.h"
#include "hybrid3_dev.h"

#define FLOD (4.0 / 3.0)

static double wtime(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int main(int argc, char *argv[]) {
  const int my_rank = 0, Proc_num = 1;
  if (argc < 3) {
    fprintf(stderr, "This program need two arguments.\n");
    return EXIT_SUCCESS;
  }
  const long matrix_size = atol(argv[1]);
  const long tile_size = atol(argv[2]);
  int ib = (int)tile_size;
  if (argc >= 4) {
    ib = atoi(argv[3]);
    if (ib <= 0) {
      fprintf(stderr, "ERROR: ib is invalid: ib == %d\n", ib);
      return EXIT_FAILURE;
    }
  }
  mat_stat_t st;
  mat_init(&st, matrix_size, matrix_size, (int)tile_size, (int)tile_size, my_rank, Proc_num);
  const unsigned long n = mat_mysize(&st);
  double *A = calloc(n, sizeof(double));
  double *TAU = calloc(mat_myTAU(&st), sizeof(double));
  double *QR = calloc(n, sizeof(double));
  double *O = calloc(n, sizeof(double));
  double *Q = calloc(n, sizeof(double));
  if (!A || !TAU || !QR || !O || !Q) {
    fprintf(stderr, "allocation failed\n");
    return EXIT_FAILURE;
  }
  const unsigned int seed = 20151130;
  mat_make(&st, A, seed, my_rank);

  double t1 = wtime();
  Hybrid_tile_QR(&st, A, TAU, my_rank, Proc_num, ib);
  double t2 = wtime();
  if (hq_last_error() != HQ_OK) return EXIT_FAILURE;
  const double elapsed_qr = t2 - t1;
  const double flops = FLOD * ((double)matrix_size * (double)matrix_size * (double)matrix_size) / elapsed_qr;
  printf("node= %d\tN= %ld\tNB= %ld\tIB= %d\ttime= %-.6f s\tflops= %lf Gflops\n", Proc_num, matrix_size, tile_size,
         ib, elapsed_qr, flops * 1e-9);
  fflush(stdout);

  t1 = wtime();
  Hybrid_make_Q(&st, Q, A, TAU, my_rank, Proc_num, ib);
  t2 = wtime();
  if (hq_last_error() != HQ_OK) return EXIT_FAILURE;
  printf("CRAYJ: Elapsed time for Hybrid_make_Q() == %.6f sec\n", t2 - t1);
  fflush(stdout);
  t1 = wtime();

  double error = mat_comp3(&st, Q, my_rank, Proc_num);
  printf("||Q||oo=\t%e\n", error);
  fflush(stdout);

  for (unsigned long i = 0; i < n; i++) O[i] = Q[i];
  Hybrid_dgemm(&st, Q, O, QR, my_rank, Proc_num); /* Q holds Q^T: QR <- Q Q^T */
  Hybrid_make_I(&st, O, my_rank, Proc_num);
  for (unsigned long i = 0; i < n; i++) QR[i] -= O[i];
  mat_dprint(&st, O, QR);
  error = mat_comp3(&st, O, my_rank, Proc_num);
  printf("||I-Q'Q||oo=\t%e\n", error);

  Hybrid_dgemm(&st, Q, A, QR, my_rank, Proc_num); /* QR <- Q R (A holds R) */
  mat_remv(&st, QR, seed, my_rank);
  mat_dprint(&st, O, QR);
  error = mat_comp3(&st, O, my_rank, Proc_num);
  printf("||A-QR||oo=\t%e\n", error);
  fprintf(stderr, "end ||A-QR||oo\n");
  t2 = wtime();
  printf("CRAYJ: Elapsed time for calculation norms = %f sec.\n", t2 - t1);

  free(Q);
  free(O);
  free(QR);
  free(TAU);
  free(A);
  return hq_last_error() == HQ_OK ? EXIT_SUCCESS : EXIT_FAILURE;
}
