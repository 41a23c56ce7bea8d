// hq_harness.cpp -- the reference's matrix helpers around the tiled QR
// (mat_prod.c:42-298): input generator, its removal, the 1-norm check and the
// tile -> column-major copy, as the reference's main.c uses them
// (main.c:138,193,226-227,253-255).  Host code on host arrays; one rank.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../include/hybrid3_dev.h"
#include "../../include/mat_prod.h"

namespace {

// Walk the local tile grid in the order of mat_prod.c:165-180 and hand each
// position's flat index to f.
template <typename F>
void tile_order(const mat_stat_t *st, F &&f) {
  for (unsigned j = 0; j < st->my_tile_N; j++)
    for (unsigned jj = 0; jj < st->tile_N; jj++)
      for (unsigned i = 0; i < st->my_tile_M; i++)
        for (unsigned ii = 0; ii < st->tile_M; ii++) f(mat_getPosID(st, i, j, ii, jj));
}

}  // namespace

extern "C" {

void mat_make(mat_stat_t *status, double *matA, unsigned int seed, int my_rank) {
  srand(seed + (unsigned)my_rank);
  tile_order(status, [&](unsigned long p) { matA[p] = (double)rand() / (double)RAND_MAX; });
}

void mat_remv(mat_stat_t *status, double *matA, unsigned int seed, int my_rank) {
  srand(seed + (unsigned)my_rank);
  tile_order(status, [&](unsigned long p) { matA[p] -= (double)rand() / (double)RAND_MAX; });
}

double mat_comp3(mat_stat_t *status, double *matE, int my_rank, int proc_num) {
  if (my_rank != 0 || proc_num != 1) {
    fprintf(stderr, "hybrid3: mat_comp3: one rank only (my_rank=%d proc_num=%d)\n", my_rank, proc_num);
    return NAN;
  }
  const unsigned long M = status->my_mat_M, N = status->my_mat_N;
  double e = 0.0;
  for (unsigned long i = 0; i < N; i++) {
    double s = 0.0;
    for (unsigned long j = 0; j < M; j++) s += fabs(matE[i * M + j]);
    if (s > e) e = s;
    if (std::isnan(s)) return s;
  }
  return e;
}

void mat_dprint(mat_stat_t *status, double *out, double *mat) {
  const unsigned tr = status->tile_M, tc = status->tile_N, nr = status->my_tile_M, nc = status->my_tile_N;
  unsigned long c = 0;
  for (unsigned j = 0; j < nc; j++)
    for (unsigned l = 0; l < tc; l++)
      for (unsigned i = 0; i < nr; i++)
        for (unsigned k = 0; k < tr; k++) out[c++] = mat[((unsigned long)i + (unsigned long)nr * j) * tr * tc + k + tr * l];
}

void mat_print(mat_stat_t *status, FILE *out, double *mat, const char *head_string) {
  const unsigned tr = status->tile_M, tc = status->tile_N, nr = status->my_tile_M, nc = status->my_tile_N;
  fprintf(out, "[%5d] %s\n", 0, head_string);
  for (unsigned i = 0; i < nr; i++)
    for (unsigned k = 0; k < tr; k++) {
      for (unsigned j = 0; j < nc; j++)
        for (unsigned l = 0; l < tc; l++)
          fprintf(out, "%lf\t", mat[((unsigned long)i + (unsigned long)nr * j) * tr * tc + k + tr * l]);
      fprintf(out, "\n");
    }
  fprintf(out, "\n");
  fflush(out);
}

}  // extern "C"
