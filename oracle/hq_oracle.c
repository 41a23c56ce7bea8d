/*
 * hq_oracle.c -- CPU oracle for the tiled Householder QR path.
 *
 * TEST INFRASTRUCTURE ONLY (see hq_oracle.h).  Plain C99 + OpenMP.  Every
 * routine cites the reference file:line it restates.  Pinning:
 *   - PLASMA-convention kernels: LAPACK dgeqrt/dtpqrt/dgemqrt/dtpmqrt
 *     golden fixtures (tests/test_oracle_plasma.py).
 *   - homegrown kernels: invariants, closed forms, and the identity
 *     R_homegrown = D * R_LAPACK (tests/test_oracle_homegrown.py).
 *   - drivers: ||A - QR||_F / ||A||_F, ||I - Q^T Q||_F, R vs numpy QR after
 *     row-sign normalisation, schedule independence (tests/test_oracle_drivers.py).
 */
#include "hq_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <time.h>

#define IDX(A, ld, r, c) ((A)[(size_t)(r) + (size_t)(c) * (size_t)(ld)])

static int imin(int a, int b) { return a < b ? a : b; }

/* =========================================================================
 * Small column-major BLAS helpers (plain loops).
 * ========================================================================= */

/* C(m x n) = alpha * A^T B + beta * C, A is k x m, B is k x n. */
static void gemm_tn(int m, int n, int k, double alpha, const double *A,
                    int lda, const double *B, int ldb, double beta, double *C,
                    int ldc) {
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      double s = 0.0;
      const double *a = A + (size_t)i * lda;
      const double *b = B + (size_t)j * ldb;
      for (int p = 0; p < k; p++) s += a[p] * b[p];
      double c = (beta == 0.0) ? 0.0 : beta * IDX(C, ldc, i, j);
      IDX(C, ldc, i, j) = c + alpha * s;
    }
}

/* C(m x n) += alpha * A B, A is m x k, B is k x n. */
static void gemm_nn_acc(int m, int n, int k, double alpha, const double *A,
                        int lda, const double *B, int ldb, double *C,
                        int ldc) {
  for (int j = 0; j < n; j++)
    for (int p = 0; p < k; p++) {
      double b = alpha * IDX(B, ldb, p, j);
      if (b == 0.0) continue;
      const double *a = A + (size_t)p * lda;
      double *c = C + (size_t)j * ldc;
      for (int i = 0; i < m; i++) c[i] += a[i] * b;
    }
}

/* =========================================================================
 * LAPACK building blocks (published algorithms, LAPACK 3.x).
 * ========================================================================= */

static double dlapy2(double x, double y) {
  double xa = fabs(x), ya = fabs(y);
  double w = xa > ya ? xa : ya;
  double z = xa < ya ? xa : ya;
  if (z == 0.0 || w > DBL_MAX) return w;
  double q = z / w;
  return w * sqrt(1.0 + q * q);
}

static double dnrm2(int n, const double *x, int incx) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += x[(size_t)i * incx] * x[(size_t)i * incx];
  return sqrt(s);
}

/* LAPACK dlarfg: H = I - tau v v^T, v = [1; x/(alpha-beta)],
 * beta = -sign(alpha) * ||[alpha; x]||.  Used by PLASMA's dgeqr2/dtsqrt
 * (called at Hybrid_tile_QR.c:985,1135). */
void orc_dlarfg(int n, double *alpha, double *x, int incx, double *tau) {
  if (n <= 1) {
    *tau = 0.0;
    return;
  }
  double xnorm = dnrm2(n - 1, x, incx);
  if (xnorm == 0.0) {
    *tau = 0.0;
    return;
  }
  double beta = -copysign(dlapy2(*alpha, xnorm), *alpha);
  const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
  int knt = 0;
  if (fabs(beta) < safmin) {
    const double rsafmn = 1.0 / safmin;
    do {
      knt++;
      for (int i = 0; i < n - 1; i++) x[(size_t)i * incx] *= rsafmn;
      beta *= rsafmn;
      *alpha *= rsafmn;
    } while (fabs(beta) < safmin && knt < 20);
    xnorm = dnrm2(n - 1, x, incx);
    beta = -copysign(dlapy2(*alpha, xnorm), *alpha);
  }
  *tau = (beta - *alpha) / beta;
  double scal = 1.0 / (*alpha - beta);
  for (int i = 0; i < n - 1; i++) x[(size_t)i * incx] *= scal;
  for (int j = 0; j < knt; j++) beta *= safmin;
  *alpha = beta;
}

/* LAPACK dgeqr2 (unblocked Householder QR of an m x n panel). */
static void dgeqr2(int m, int n, double *A, int lda, double *tau) {
  int k = imin(m, n);
  for (int i = 0; i < k; i++) {
    double *aii = &IDX(A, lda, i, i);
    double *x = (i + 1 < m) ? &IDX(A, lda, i + 1, i) : aii;
    orc_dlarfg(m - i, aii, x, 1, &tau[i]);
    if (i < n - 1 && tau[i] != 0.0) {
      double save = *aii;
      *aii = 1.0;
      const double *v = aii; /* v(0)=1, v(1:) below */
      for (int j = i + 1; j < n; j++) {
        double *c = &IDX(A, lda, i, j);
        double w = 0.0;
        for (int r = 0; r < m - i; r++) w += v[r] * c[r];
        w *= tau[i];
        for (int r = 0; r < m - i; r++) c[r] -= v[r] * w;
      }
      *aii = save;
    }
  }
}

/* LAPACK dlarft('F','C'): upper triangular T of H(0)..H(k-1), V unit lower
 * trapezoidal n x k (implicit unit diagonal, upper part ignored). */
void orc_dlarft_fc(int n, int k, const double *V, int ldv, const double *tau,
                   double *T, int ldt) {
  for (int i = 0; i < k; i++) {
    if (tau[i] == 0.0) {
      for (int j = 0; j <= i; j++) IDX(T, ldt, j, i) = 0.0;
      continue;
    }
    for (int j = 0; j < i; j++) {
      /* v_j^T v_i = V(i,j)*1 + sum_{r>i} V(r,j) V(r,i) */
      double s = IDX(V, ldv, i, j);
      for (int r = i + 1; r < n; r++) s += IDX(V, ldv, r, j) * IDX(V, ldv, r, i);
      IDX(T, ldt, j, i) = -tau[i] * s;
    }
    /* T(0:i,i) = T(0:i,0:i) * T(0:i,i)   (upper triangular, in place) */
    for (int j = 0; j < i; j++) {
      double s = 0.0;
      for (int p = j; p < i; p++) s += IDX(T, ldt, j, p) * IDX(T, ldt, p, i);
      IDX(T, ldt, j, i) = s;
    }
    IDX(T, ldt, i, i) = tau[i];
  }
}

/* LAPACK dlarfb('L','T','F','C'): C <- (I - V T V^T)^T C = (I - V T^T V^T) C,
 * V unit lower trapezoidal m x k, T upper k x k, C m x n. */
static void dlarfb_ltfc(int m, int n, int k, const double *V, int ldv,
                        const double *T, int ldt, double *C, int ldc) {
  if (m <= 0 || n <= 0 || k <= 0) return;
  double *W = (double *)malloc(sizeof(double) * (size_t)k * n);
  /* W = V^T C */
  for (int j = 0; j < n; j++)
    for (int i = 0; i < k; i++) {
      double s = IDX(C, ldc, i, j);
      for (int r = i + 1; r < m; r++) s += IDX(V, ldv, r, i) * IDX(C, ldc, r, j);
      W[i + (size_t)j * k] = s;
    }
  /* W = T^T W  (T^T lower: row i uses T(p,i), p <= i); go bottom-up */
  for (int j = 0; j < n; j++)
    for (int i = k - 1; i >= 0; i--) {
      double s = 0.0;
      for (int p = 0; p <= i; p++) s += IDX(T, ldt, p, i) * W[p + (size_t)j * k];
      W[i + (size_t)j * k] = s;
    }
  /* C -= V W */
  for (int j = 0; j < n; j++)
    for (int i = 0; i < k; i++) {
      double w = W[i + (size_t)j * k];
      if (w == 0.0) continue;
      IDX(C, ldc, i, j) -= w;
      for (int r = i + 1; r < m; r++) IDX(C, ldc, r, j) -= IDX(V, ldv, r, i) * w;
    }
  free(W);
}

/* =========================================================================
 * PLASMA-convention tile kernels.
 * ========================================================================= */

/* GEQRT = CORE_dgeqrt(m,n,ib,A,lda,T,ldt,TAU,work)
 * (called at Hybrid_tile_QR.c:985-990): for each ib-column panel,
 * dgeqr2 + dlarft('F','C') into T(0:sb, i:i+sb) + dlarfb on the rest. */
void orc_dgeqrt(int m, int n, int ib, double *A, int lda, double *T, int ldt,
                double *tau) {
  int k = imin(m, n);
  for (int i = 0; i < k; i += ib) {
    int sb = imin(ib, k - i);
    dgeqr2(m - i, sb, &IDX(A, lda, i, i), lda, &tau[i]);
    orc_dlarft_fc(m - i, sb, &IDX(A, lda, i, i), lda, &tau[i],
                  &IDX(T, ldt, 0, i), ldt);
    if (n > i + sb)
      dlarfb_ltfc(m - i, n - i - sb, sb, &IDX(A, lda, i, i), lda,
                  &IDX(T, ldt, 0, i), ldt, &IDX(A, lda, i, i + sb), lda);
  }
}

/* parfb(Left, Trans, Forward, Columnwise, l=0):
 *   W = A1(0:k, :) + V^T A2 ; W = T^T W ; A1(0:k,:) -= W ; A2 -= V W
 * (the body of CORE_dtsmqr for one ib block). */
static void parfb_lt(int k, int n, int m2, double *A1, int lda1, double *A2,
                     int lda2, const double *V, int ldv, const double *T,
                     int ldt) {
  double *W = (double *)malloc(sizeof(double) * (size_t)k * n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < k; i++) {
      double s = IDX(A1, lda1, i, j);
      const double *v = V + (size_t)i * ldv;
      const double *a = A2 + (size_t)j * lda2;
      for (int r = 0; r < m2; r++) s += v[r] * a[r];
      W[i + (size_t)j * k] = s;
    }
  for (int j = 0; j < n; j++)
    for (int i = k - 1; i >= 0; i--) {
      double s = 0.0;
      for (int p = 0; p <= i; p++) s += IDX(T, ldt, p, i) * W[p + (size_t)j * k];
      W[i + (size_t)j * k] = s;
    }
  for (int j = 0; j < n; j++)
    for (int i = 0; i < k; i++) IDX(A1, lda1, i, j) -= W[i + (size_t)j * k];
  gemm_nn_acc(m2, n, k, -1.0, V, ldv, W, k, A2, lda2);
  free(W);
}

/* SSRFB = CORE_dtsmqr(Left,Trans,m1,n1,m2,n2,k,ib,A1,lda1,A2,lda2,V,ldv,T,ldt)
 * (called at Hybrid_tile_QR.c:1356-1363): forward over ib blocks. */
void orc_dtsmqr_lt(int m1, int n1, int m2, int n2, int k, int ib, double *A1,
                   int lda1, double *A2, int lda2, const double *V, int ldv,
                   const double *T, int ldt) {
  (void)m1;
  (void)n2;
  for (int i = 0; i < k; i += ib) {
    int kb = imin(ib, k - i);
    parfb_lt(kb, n1, m2, &IDX(A1, lda1, i, 0), lda1, A2, lda2,
             V + (size_t)i * ldv, ldv, &IDX(T, ldt, 0, i), ldt);
  }
}

/* TSQRT = CORE_dtsqrt(m,n,ib,A1,lda1,A2,lda2,T,ldt,TAU,work)
 * (called at Hybrid_tile_QR.c:1135-1141): QR of [R1; A2], R1 upper n x n,
 * A2 m x n; A2 <- V2, T(0:sb, ii:ii+sb) per ib block. */
void orc_dtsqrt(int m, int n, int ib, double *A1, int lda1, double *A2,
                int lda2, double *T, int ldt, double *tau) {
  double *work = (double *)malloc(sizeof(double) * (size_t)(ib > 0 ? ib : 1));
  for (int ii = 0; ii < n; ii += ib) {
    int sb = imin(n - ii, ib);
    for (int i = 0; i < sb; i++) {
      int c = ii + i;
      orc_dlarfg(m + 1, &IDX(A1, lda1, c, c), &IDX(A2, lda2, 0, c), 1, &tau[c]);
      if (c + 1 < n) {
        /* apply H(c) to [A1(c, c+1:ii+sb); A2(:, c+1:ii+sb)] */
        double alpha = -tau[c];
        int nr = sb - i - 1;
        for (int j = 0; j < nr; j++) {
          double s = IDX(A1, lda1, c, c + 1 + j);
          for (int r = 0; r < m; r++)
            s += IDX(A2, lda2, r, c + 1 + j) * IDX(A2, lda2, r, c);
          work[j] = s;
        }
        for (int j = 0; j < nr; j++) IDX(A1, lda1, c, c + 1 + j) += alpha * work[j];
        for (int j = 0; j < nr; j++) {
          double w = alpha * work[j];
          for (int r = 0; r < m; r++)
            IDX(A2, lda2, r, c + 1 + j) += IDX(A2, lda2, r, c) * w;
        }
      }
      /* T(0:i, c) = -tau * V2(:, ii:c)^T V2(:, c); then T(0:i,0:i) * that */
      for (int j = 0; j < i; j++) {
        double s = 0.0;
        for (int r = 0; r < m; r++)
          s += IDX(A2, lda2, r, ii + j) * IDX(A2, lda2, r, c);
        IDX(T, ldt, j, c) = -tau[c] * s;
      }
      for (int j = 0; j < i; j++) {
        double s = 0.0;
        for (int p = j; p < i; p++) s += IDX(T, ldt, j, ii + p) * IDX(T, ldt, p, c);
        IDX(T, ldt, j, c) = s;
      }
      IDX(T, ldt, i, c) = tau[c];
    }
    if (n > ii + sb)
      orc_dtsmqr_lt(sb, n - (ii + sb), m, n - (ii + sb), ib, ib,
                    &IDX(A1, lda1, ii, ii + sb), lda1, &IDX(A2, lda2, 0, ii + sb),
                    lda2, &IDX(A2, lda2, 0, ii), lda2, &IDX(T, ldt, 0, ii), ldt);
  }
  free(work);
}

/* LARFB = CORE_dormqr(Left,Trans,m,n,k,ib,V,ldv,T,ldt,C,ldc)
 * (called at Hybrid_tile_QR.c:1276-1283): forward over ib blocks. */
void orc_dormqr_lt(int m, int n, int k, int ib, const double *V, int ldv,
                   const double *T, int ldt, double *C, int ldc) {
  for (int i = 0; i < k; i += ib) {
    int kb = imin(ib, k - i);
    dlarfb_ltfc(m - i, n, kb, &IDX(V, ldv, i, i), ldv, &IDX(T, ldt, 0, i), ldt,
                &IDX(C, ldc, i, 0), ldc);
  }
}

/* T factor of a triangle-on-square block V = [I; V2] (Hybrid_make_Q.c:14-44,
 * "crayj_dtsrft", the T part of CORE_dtsqrt). */
void orc_dtsrft(int m, int n, int ib, const double *V, int ldv,
                const double *tau, double *T, int ldt) {
  for (int ii = 0; ii < n; ii += ib) {
    int sb = imin(n - ii, ib);
    for (int i = 0; i < sb; i++) {
      int c = ii + i;
      for (int j = 0; j < i; j++) {
        double s = 0.0;
        for (int r = 0; r < m; r++) s += IDX(V, ldv, r, ii + j) * IDX(V, ldv, r, c);
        IDX(T, ldt, j, c) = -tau[c] * s;
      }
      for (int j = 0; j < i; j++) {
        double s = 0.0;
        for (int p = j; p < i; p++) s += IDX(T, ldt, j, ii + p) * IDX(T, ldt, p, c);
        IDX(T, ldt, j, c) = s;
      }
      IDX(T, ldt, i, c) = tau[c];
    }
  }
}

/* =========================================================================
 * Homegrown tile_kernel.h convention (tile_kernel.c).  beta = +||.||,
 * v = [1; x/(alpha-beta)], TAU = t*(alpha-beta)^2, T stored lower (= T^T).
 * ========================================================================= */

/* tile_kernel.c:138-218 */
void orc_tile_sub_dgeqr2(int row, int col, double *A, int lda, double *TAU,
                         double *Y, double *work) {
  double *yA = work;
  int i = 0;
  for (i = 0; i < col - 1; i++) {
    /* yA(0:col-i) = A(i+1:row, i)^T A(i+1:row, i:col) */
    gemm_tn(1, col - i, row - i - 1, 1.0, &IDX(A, lda, i + 1, i), lda,
            &IDX(A, lda, i + 1, i), lda, 0.0, yA, 1);
    double a_head = IDX(A, lda, i, i);
    double norm_1 = yA[0];
    double norm = norm_1 + a_head * a_head;
    double b = sqrt(norm);
    double y_head = a_head - b;
    IDX(A, lda, i, i) = y_head;
    double t = 2 / (y_head * y_head + norm_1);
    for (int j = 0; j < col - i - 1; j++) yA[1 + j] += y_head * IDX(A, lda, i, i + 1 + j);
    /* A(i:row, i+1:col) -= t * A(i:row, i) * yA(1:)^T */
    for (int j = 0; j < col - i - 1; j++) {
      double w = -t * yA[1 + j];
      for (int r = 0; r < row - i; r++)
        IDX(A, lda, i + r, i + 1 + j) += IDX(A, lda, i + r, i) * w;
    }
    IDX(A, lda, i, i) = b;
    double rs = 1 / y_head;
    for (int r = 0; r < row - i - 1; r++) IDX(A, lda, i + 1 + r, i) *= rs;
    for (int r = 0; r < row - i - 1; r++) IDX(Y, lda, i + 1 + r, i) = IDX(A, lda, i + 1 + r, i);
    IDX(Y, lda, i, i) = 1;
    TAU[i] = t * y_head * y_head;
  }
  if (row > col) {
    i = col - 1;
    gemm_tn(1, col - i, row - i - 1, 1.0, &IDX(A, lda, i + 1, i), lda,
            &IDX(A, lda, i + 1, i), lda, 0.0, yA, 1);
    double a_head = IDX(A, lda, i, i);
    double norm_1 = yA[0];
    double norm = norm_1 + a_head * a_head;
    double b = sqrt(norm);
    double y_head = a_head - b;
    IDX(A, lda, i, i) = b;
    double rs = 1 / y_head;
    for (int r = 0; r < row - i - 1; r++) IDX(A, lda, i + 1 + r, i) *= rs;
    for (int r = 0; r < row - i - 1; r++) IDX(Y, lda, i + 1 + r, i) = IDX(A, lda, i + 1 + r, i);
    IDX(Y, lda, i, i) = 1;
    double t = 2 / (y_head * y_head + norm_1);
    TAU[i] = t * y_head * y_head;
  } else if (row == col) {
    IDX(Y, lda, i, i) = 0;
  }
}

/* tile_kernel.c:220-262: Tl = T_LAPACK^T (lower) from the explicit Y and TAU.
 * The strictly upper part of T must be zero on entry (the row recursion
 * reads it). */
void orc_tile_sub_dlarft(int row, int col, const double *Y, int lda,
                         const double *TAU, double *T, int ldt, double *work) {
  double *YtY = work; /* col x col, ld col */
  gemm_tn(col, col, row, 1.0, Y, lda, Y, lda, 0.0, YtY, col);
  for (int i = 0; i < col; i++) IDX(T, ldt, i, i) = TAU[i];
  for (int i = 1; i < col; i++) {
    /* RowMajor dgemm: T(i, r) = -T(i,i) * sum_{c<i} T(c, r) * YtY(c, i), r<i */
    double ti = -IDX(T, ldt, i, i);
    for (int r = 0; r < i; r++) {
      double s = 0.0;
      for (int c = 0; c < i; c++) s += IDX(T, ldt, c, r) * YtY[c + (size_t)i * col];
      IDX(T, ldt, i, r) = ti * s;
    }
  }
}

/* tile_kernel.c:264-311:
 *   W = tril(Y)^T A ; W(0:Y_col) = tril(T) W(0:Y_col) ; A -= Y(:,0:Y_col) W.
 * The triangular multiplies ignore Y's / T's strictly upper part; the final
 * product reads Y in full. */
void orc_tile_sub_dlarfb(int row, int Y_col, int A_col, const double *Y,
                         int ldy, const double *T, int ldt, double *A, int lda,
                         double *work) {
  double *W = work; /* row x A_col, ld row */
  for (int j = 0; j < A_col; j++)
    for (int r = 0; r < Y_col; r++) {
      /* (tril(Y)^T A)(r, j) = sum_{c >= r} Y(c, r) A(c, j) */
      double s = 0.0;
      for (int c = r; c < row; c++) s += IDX(Y, ldy, c, r) * IDX(A, lda, c, j);
      W[r + (size_t)j * row] = s;
    }
  for (int j = 0; j < A_col; j++)
    for (int r = Y_col - 1; r >= 0; r--) {
      double s = 0.0;
      for (int c = 0; c <= r; c++) s += IDX(T, ldt, r, c) * W[c + (size_t)j * row];
      W[r + (size_t)j * row] = s;
    }
  gemm_nn_acc(row, A_col, Y_col, -1.0, Y, ldy, W, row, A, lda);
}

/* tile_kernel.c:313-363:
 *   W = A1 ; W(0:Y_col) += Y^T A2 ; W(0:Y_col) = tril(T) W(0:Y_col) ;
 *   A1 -= W (all over_row rows) ; A2 -= Y W(0:Y_col). */
void orc_tile_sub_dssrfb(int over_row, int under_row, int Y_col, int A_col,
                         const double *Y, int ldy, const double *T, int ldt,
                         double *A1, int lda1, double *A2, int lda2,
                         double *work) {
  double *W = work; /* over_row x A_col, ld over_row */
  int ld1 = over_row;
  for (int j = 0; j < A_col; j++)
    for (int r = 0; r < over_row; r++) W[r + (size_t)j * ld1] = IDX(A1, lda1, r, j);
  for (int j = 0; j < A_col; j++)
    for (int r = 0; r < Y_col; r++) {
      double s = 0.0;
      for (int c = 0; c < under_row; c++) s += IDX(Y, ldy, c, r) * IDX(A2, lda2, c, j);
      W[r + (size_t)j * ld1] += s;
    }
  for (int j = 0; j < A_col; j++)
    for (int r = Y_col - 1; r >= 0; r--) {
      double s = 0.0;
      for (int c = 0; c <= r; c++) s += IDX(T, ldt, r, c) * W[c + (size_t)j * ld1];
      W[r + (size_t)j * ld1] = s;
    }
  for (int j = 0; j < A_col; j++)
    for (int r = 0; r < over_row; r++) IDX(A1, lda1, r, j) -= W[r + (size_t)j * ld1];
  gemm_nn_acc(under_row, A_col, Y_col, -1.0, Y, ldy, W, ld1, A2, lda2);
}

/* tile_kernel.c:654-716 */
void orc_tile_sub_dtsqr2(int row, int col, double *A1, int lda1, double *A2,
                         int lda2, double *TAU, double *work) {
  double *dvec = work;
  int i;
  for (i = 0; i < col - 1; i++) {
    gemm_tn(1, col - i, row, 1.0, &IDX(A2, lda2, 0, i), lda2, &IDX(A2, lda2, 0, i),
            lda2, 0.0, dvec, 1);
    double a_head = IDX(A1, lda1, i, i);
    double norm_1 = dvec[0];
    double norm = norm_1 + a_head * a_head;
    double b = sqrt(norm);
    double y_head = a_head - b;
    double t = 2 / (y_head * y_head + norm_1);
    for (int j = 0; j < col - i - 1; j++) dvec[1 + j] += y_head * IDX(A1, lda1, i, i + 1 + j);
    for (int j = 0; j < col - i - 1; j++) IDX(A1, lda1, i, i + 1 + j) += -t * y_head * dvec[1 + j];
    for (int j = 0; j < col - i - 1; j++) {
      double w = -t * dvec[1 + j];
      for (int r = 0; r < row; r++) IDX(A2, lda2, r, i + 1 + j) += IDX(A2, lda2, r, i) * w;
    }
    IDX(A1, lda1, i, i) = b;
    double rs = 1 / y_head;
    for (int r = 0; r < row; r++) IDX(A2, lda2, r, i) *= rs;
    TAU[i] = t * y_head * y_head;
  }
  gemm_tn(1, col - i, row, 1.0, &IDX(A2, lda2, 0, i), lda2, &IDX(A2, lda2, 0, i),
          lda2, 0.0, dvec, 1);
  double a_head = IDX(A1, lda1, i, i);
  double norm_1 = dvec[0];
  double norm = norm_1 + a_head * a_head;
  double b = sqrt(norm);
  double y_head = a_head - b;
  double t = 2 / (y_head * y_head + norm_1);
  IDX(A1, lda1, i, i) = b;
  double rs = 1 / y_head;
  for (int r = 0; r < row; r++) IDX(A2, lda2, r, i) *= rs;
  TAU[i] = t * y_head * y_head;
}

/* tile_kernel.c:28-57 */
void orc_tile_dgeqrt(int row, int col, int block, double *A, int lda,
                     double *TAU, double *Y, double *T, int ldt, double *work) {
  int i;
  for (i = 0; i < col - block; i += block) {
    orc_tile_sub_dgeqr2(row - i, block, &IDX(A, lda, i, i), lda, TAU + i,
                        &IDX(Y, lda, i, i), work);
    orc_tile_sub_dlarft(row - i, block, &IDX(Y, lda, i, i), lda, TAU + i,
                        &IDX(T, ldt, i, i), ldt, work);
    orc_tile_sub_dlarfb(row - i, block, col - i - block, &IDX(Y, lda, i, i), lda,
                        &IDX(T, ldt, i, i), ldt, &IDX(A, lda, i, i + block), lda,
                        work);
  }
  orc_tile_sub_dgeqr2(row - i, col - i, &IDX(A, lda, i, i), lda, TAU + i,
                      &IDX(Y, lda, i, i), work);
  orc_tile_sub_dlarft(row - i, col - i, &IDX(Y, lda, i, i), lda, TAU + i,
                      &IDX(T, ldt, i, i), ldt, work);
}

/* tile_kernel.c:60-104 */
void orc_tile_dtsqrt(int row, int col, int block, double *A1, int lda1,
                     double *A2, int lda2, double *TAU, double *T, int ldt,
                     double *work) {
  int i;
  for (i = 0; i < col - block; i += block) {
    orc_tile_sub_dtsqr2(row, block, &IDX(A1, lda1, i, i), lda1,
                        &IDX(A2, lda2, 0, i), lda2, TAU + i, work);
    orc_tile_sub_dlarft(row, block, &IDX(A2, lda2, 0, i), lda2, TAU + i,
                        &IDX(T, ldt, i, i), ldt, work);
    orc_tile_sub_dssrfb(block, row, block, col - i - block, &IDX(A2, lda2, 0, i),
                        lda2, &IDX(T, ldt, i, i), ldt, &IDX(A1, lda1, i, i + block),
                        lda1, &IDX(A2, lda2, 0, i + block), lda2, work);
  }
  if (col % block > 0) {
    i = block * (col / block);
    orc_tile_sub_dtsqr2(row, col % block, &IDX(A1, lda1, i, i), lda1,
                        &IDX(A2, lda2, 0, i), lda2, TAU + i, work);
    orc_tile_sub_dlarft(row, col % block, &IDX(A2, lda2, 0, i), lda2, TAU + i,
                        &IDX(T, ldt, i, i), ldt, work);
  } else {
    i = col - block;
    orc_tile_sub_dtsqr2(row, block, &IDX(A1, lda1, i, i), lda1,
                        &IDX(A2, lda2, 0, i), lda2, TAU + i, work);
    orc_tile_sub_dlarft(row, block, &IDX(A2, lda2, 0, i), lda2, TAU + i,
                        &IDX(T, ldt, i, i), ldt, work);
  }
}

/* tile_kernel.c:106-119 */
void orc_tile_dlarfb(int row, int col, double *A, int lda, const double *Y,
                     const double *T, int ldt, double *work) {
  orc_tile_sub_dlarfb(row, col, col, Y, lda, T, ldt, A, lda, work);
}

/* tile_kernel.c:121-136 */
void orc_tile_dssrfb(int row, int col, double *A1, double *A2, int lda,
                     const double *Y, const double *T, int ldt, double *work) {
  orc_tile_sub_dssrfb(row, row, col, col, Y, lda, T, ldt, A1, lda, A2, lda, work);
}

/* tile_kernel.c:368-437.  block < 2: full recursion (tile_sub_dlarft).
 * 2 <= block < col: combines the diagonal blocks already in T; the loop
 * bound i < col-block of the reference is kept as written (the last full
 * block is skipped).  block >= col: computes nothing. */
void orc_tile_dlarft(int row, int col, int block, const double *Y, int lda,
                     const double *TAU, double *T, int ldt, double *work) {
  if (block < 2) {
    orc_tile_sub_dlarft(row, col, Y, lda, TAU, T, ldt, work);
    return;
  }
  if (block >= col) return;
  int ldyty = (col / block) * block + block;
  double *YtY = work;                      /* row-major col x col, ld ldyty */
  double *T1YtY = work + (size_t)row * col; /* row-major i x block, ld block */
  /* YtY_rm(r, c) = sum_p Y(p, r) Y(p, c) */
  for (int r = 0; r < col; r++)
    for (int c = 0; c < col; c++) {
      double s = 0.0;
      for (int p = 0; p < row; p++) s += IDX(Y, lda, p, r) * IDX(Y, lda, p, c);
      YtY[(size_t)r * ldyty + c] = s;
    }
  for (int i = block; i < col - block; i += block) {
    /* T1YtY(r, c) = sum_{p<i} T_rm(r, p) YtY_rm(p, i + c), T_rm(r,p)=T(p,r) */
    for (int r = 0; r < i; r++)
      for (int c = 0; c < block; c++) {
        double s = 0.0;
        for (int p = 0; p < i; p++) s += IDX(T, ldt, p, r) * YtY[(size_t)p * ldyty + i + c];
        T1YtY[(size_t)r * block + c] = s;
      }
    /* C_rm(r, c) = T(i + c, r) = -sum_p T1YtY(r, p) T_rm(i + p, i + c) */
    for (int r = 0; r < i; r++)
      for (int c = 0; c < block; c++) {
        double s = 0.0;
        for (int p = 0; p < block; p++)
          s += T1YtY[(size_t)r * block + p] * IDX(T, ldt, i + c, i + p);
        IDX(T, ldt, i + c, r) = -s;
      }
  }
  for (int i = (col / block) * block; i < col; i++) {
    double ti = -IDX(T, ldt, i, i);
    for (int r = 0; r < i; r++) {
      double s = 0.0;
      for (int c = 0; c < i; c++) s += IDX(T, ldt, c, r) * YtY[(size_t)c * ldyty + i];
      IDX(T, ldt, i, r) = ti * s;
    }
  }
}

/* =========================================================================
 * Tile-layout drivers.
 * ========================================================================= */

#define TILE(A, MT, nb, i, j) ((A) + ((size_t)(i) + (size_t)(j) * (MT)) * (size_t)(nb) * (nb))
#define TAUP(TAU, MT, nb, i, j) ((TAU) + ((size_t)(i) + (size_t)(j) * (MT)) * (size_t)(nb))
#define TFAC(T, MT, nb, ib, i, j) ((T) + ((size_t)(i) + (size_t)(j) * (MT)) * (size_t)(ib) * (nb))

/* one task of the flat-tree DAG: (i=step, j=column, k=row), as the
 * reference's TilSch_t (TileSch.h:17-24) and dispatch (Hybrid_tile_QR.c:960,
 * 1123,1254,1332). */
static void run_task(int MT, int nb, int ib, double *A, double *TAU, double *T,
                     double *Tscr, int i, int j, int k) {
  double *Tk = T ? TFAC(T, MT, nb, ib, k, i) : Tscr;
  if (i == j && j == k) {
    orc_dgeqrt(nb, nb, ib, TILE(A, MT, nb, i, i), nb, Tk, ib, TAUP(TAU, MT, nb, i, i));
  } else if (i == j) {
    orc_dtsqrt(nb, nb, ib, TILE(A, MT, nb, i, i), nb, TILE(A, MT, nb, k, i), nb, Tk,
               ib, TAUP(TAU, MT, nb, k, i));
  } else if (i == k) {
    orc_dormqr_lt(nb, nb, nb, ib, TILE(A, MT, nb, i, i), nb, Tk, ib,
                  TILE(A, MT, nb, i, j), nb);
  } else {
    orc_dtsmqr_lt(nb, nb, nb, nb, nb, ib, TILE(A, MT, nb, i, j), nb,
                  TILE(A, MT, nb, k, j), nb, TILE(A, MT, nb, k, i), nb, Tk, ib);
  }
}

#define DEP_MEM 1
#define DEP_LIN 2
#define DEP_YTM 4
#define DEP_END 7

typedef struct { int i, j, k; } otask_t;

/* Dynamic DAG execution with a FIFO of ready tasks and per-task dependency
 * bits, restating TileSch.c:10-61 and Hybrid_tile_QR.c:496-601,905-1461 for
 * one rank.  Two reference defects are not reproduced: one OpenMP team is
 * started (not two stacked parallel regions), and termination counts the
 * completed tasks (the reference stops at the last GEQRT, which never ends
 * for MT > NT). */
static void tile_qr_dynamic(int MT, int NT, int nb, int ib, double *A,
                            double *TAU, double *T, int nthreads) {
  int KT = imin(MT, NT);
  long total = 0;
  for (int i = 0; i < KT; i++) total += (long)(MT - i) * (NT - i);
  unsigned char *pt = (unsigned char *)calloc((size_t)MT * NT * NT, 1);
#define PT(i, j, k) pt[((size_t)(i) * MT * NT + (size_t)(j) * MT + (k))]
  for (int j = 0; j < NT; j++)
    for (int k = 0; k < MT; k++)
      PT(0, j, k) = (j == 0) ? (DEP_MEM | DEP_YTM) : (k == 0) ? (DEP_MEM | DEP_LIN) : DEP_MEM;
  for (int i = 1; i < KT; i++)
    for (int j = i; j < NT; j++)
      for (int k = i; k < MT; k++)
        PT(i, j, k) = (i == j && j == k) ? (DEP_YTM | DEP_LIN)
                      : (j == i)          ? DEP_YTM
                      : (k == i)          ? DEP_LIN
                                          : 0;
  otask_t *fifo = (otask_t *)malloc(sizeof(otask_t) * (size_t)(total + 1));
  long head = 0, tail = 0, done = 0;
  fifo[tail++] = (otask_t){0, 0, 0};

#pragma omp parallel num_threads(nthreads)
  {
    double *Tscr = (double *)malloc(sizeof(double) * (size_t)ib * nb);
    otask_t *ready = (otask_t *)malloc(sizeof(otask_t) * (size_t)(NT + 4));
    for (;;) {
      otask_t t;
      int got = 0;
      long d;
#pragma omp critical(hq_fifo)
      {
        if (head < tail) {
          t = fifo[head++];
          got = 1;
        }
        d = done;
      }
      if (!got) {
        if (d >= total) break;
        continue;
      }
      run_task(MT, nb, ib, A, TAU, T, Tscr, t.i, t.j, t.k);
      int nr = 0;
#pragma omp critical(hq_prog)
      {
        int i = t.i, j = t.j, k = t.k;
#define SETDEP(a, b, c, bit)                                  \
  do {                                                        \
    PT(a, b, c) |= (bit);                                     \
    if (PT(a, b, c) == DEP_END) ready[nr++] = (otask_t){a, b, c}; \
  } while (0)
        if (i == j && j == k) {
          if (k + 1 < MT) SETDEP(i, j, k + 1, DEP_LIN);
          for (int jj = i + 1; jj < NT; jj++) SETDEP(i, jj, k, DEP_YTM);
        } else if (i == j) {
          if (k + 1 < MT) SETDEP(i, j, k + 1, DEP_LIN);
          for (int jj = i + 1; jj < NT; jj++) SETDEP(i, jj, k, DEP_YTM);
        } else if (i == k) {
          if (k + 1 < MT) SETDEP(i, j, k + 1, DEP_LIN);
        } else {
          if (k + 1 < MT) SETDEP(i, j, k + 1, DEP_LIN);
          if (i + 1 < KT) SETDEP(i + 1, j, k, DEP_MEM);
        }
#undef SETDEP
      }
#pragma omp critical(hq_fifo)
      {
        for (int q = 0; q < nr; q++) fifo[tail++] = ready[q];
        done++;
      }
    }
    free(ready);
    free(Tscr);
  }
#undef PT
  free(fifo);
  free(pt);
}

int orc_tile_qr(int MT, int NT, int nb, int ib, double *A, double *TAU,
                double *T, int nthreads) {
  if (MT <= 0 || NT <= 0 || nb <= 0 || ib <= 0) return -1;
  if (nthreads > 1) {
    /* consumers read the producer's T, so every tile keeps its own T (the
     * reference copies (V,T) into a shared ring for the same reason,
     * Hybrid_tile_QR.c:1051-1086) */
    double *Tall = T ? T : (double *)malloc(sizeof(double) * (size_t)MT * NT * ib * nb);
    tile_qr_dynamic(MT, NT, nb, ib, A, TAU, Tall, nthreads);
    if (!T) free(Tall);
    return 0;
  }
  double *Tscr = (double *)malloc(sizeof(double) * (size_t)ib * nb);
  int KT = imin(MT, NT);
  for (int i = 0; i < KT; i++) {
    run_task(MT, nb, ib, A, TAU, T, Tscr, i, i, i);
    for (int j = i + 1; j < NT; j++) run_task(MT, nb, ib, A, TAU, T, Tscr, i, j, i);
    for (int k = i + 1; k < MT; k++) {
      run_task(MT, nb, ib, A, TAU, T, Tscr, i, i, k);
      for (int j = i + 1; j < NT; j++) run_task(MT, nb, ib, A, TAU, T, Tscr, i, j, k);
    }
  }
  free(Tscr);
  return 0;
}

/* Hybrid_make_Q.c:56-237 for one rank: Q <- Q^T (tile layout) using full-nb
 * T factors (LAPACKE_dlarft / crayj_dtsrft with IB = nb) and PLASMA-style
 * updates; zeroes the reflectors in RY.  Q must be zero on entry (the
 * reference only writes the diagonal ones, Hybrid_make_Q.c:91-95). */
void orc_make_Q(int MT, int NT, int nb, double *Q, double *RY,
                const double *TAU) {
  size_t ts = (size_t)nb * nb;
  double *Y = (double *)calloc(ts, sizeof(double));
  double *Tb = (double *)calloc(ts, sizeof(double));
  for (int i = 0; i < NT; i++)
    for (int r = 0; r < nb; r++) IDX(TILE(Q, MT, nb, i, i), nb, r, r) = 1.0;
  for (int gi = 0; gi < NT; gi++) {
    double *D = TILE(RY, MT, nb, gi, gi);
    for (int c = 0; c < nb - 1; c++) {
      for (int r = c + 1; r < nb; r++) IDX(Y, nb, r, c) = IDX(D, nb, r, c);
      IDX(Y, nb, c, c) = 1;
    }
    Y[ts - 1] = 0;
    orc_dlarft_fc(nb, nb, Y, nb, TAUP(TAU, MT, nb, gi, gi), Tb, nb);
    for (int j = 0; j < NT; j++)
      orc_dormqr_lt(nb, nb, nb, nb, Y, nb, Tb, nb, TILE(Q, MT, nb, gi, j), nb);
    for (int c = 0; c < nb; c++)
      for (int r = c + 1; r < nb; r++) IDX(D, nb, r, c) = 0.0;
    for (int k = gi + 1; k < MT; k++) {
      double *V = TILE(RY, MT, nb, k, gi);
      orc_dtsrft(nb, nb, nb, V, nb, TAUP(TAU, MT, nb, k, gi), Tb, nb);
      for (int j = 0; j < NT; j++)
        orc_dtsmqr_lt(nb, nb, nb, nb, nb, nb, TILE(Q, MT, nb, gi, j), nb,
                      TILE(Q, MT, nb, k, j), nb, V, nb, Tb, nb);
      memset(V, 0, ts * sizeof(double));
    }
  }
  free(Y);
  free(Tb);
}

/* Shared sweep of Hybrid_dorgqr.c:56-138 and Hybrid_dormqr.c:64-127 for one
 * rank, with the homegrown kernels (tile_dlarft block=1, tile_dlarfb,
 * tile_dssrfb) applied to every tile column of X. */
static void homegrown_sweep(int MT, int NT, int nb, double *X, const double *RY,
                            const double *TAU) {
  size_t ts = (size_t)nb * nb;
  double *Y = (double *)calloc(ts, sizeof(double));
  double *Tb = (double *)calloc(ts, sizeof(double));
  double *work = (double *)calloc(3 * ts, sizeof(double));
  for (int gi = 0; gi < NT; gi++) {
    const double *D = TILE(RY, MT, nb, gi, gi);
    for (int c = 0; c < nb - 1; c++) {
      for (int r = c + 1; r < nb; r++) IDX(Y, nb, r, c) = IDX(D, nb, r, c);
      IDX(Y, nb, c, c) = 1;
    }
    Y[ts - 1] = 0;
    orc_tile_dlarft(nb, nb, 1, Y, nb, TAUP(TAU, MT, nb, gi, gi), Tb, nb, work);
    for (int j = 0; j < NT; j++)
      orc_tile_dlarfb(nb, nb, TILE(X, MT, nb, gi, j), nb, Y, Tb, nb, work);
    for (int k = gi + 1; k < MT; k++) {
      const double *V = TILE(RY, MT, nb, k, gi);
      orc_tile_dlarft(nb, nb, 1, V, nb, TAUP(TAU, MT, nb, k, gi), Tb, nb, work);
      for (int j = 0; j < NT; j++)
        orc_tile_dssrfb(nb, nb, TILE(X, MT, nb, gi, j), TILE(X, MT, nb, k, j), nb, V,
                        Tb, nb, work);
    }
  }
  free(Y);
  free(Tb);
  free(work);
}

/* Hybrid_dorgqr.c:22-192 (one rank): Q <- I, then the Q^T sweep. */
void orc_dorgqr(int MT, int NT, int nb, double *Q, const double *RY,
                const double *TAU) {
  size_t n = (size_t)MT * NT * nb * nb;
  memset(Q, 0, n * sizeof(double));
  for (int i = 0; i < NT && i < MT; i++)
    for (int r = 0; r < nb; r++) IDX(TILE(Q, MT, nb, i, i), nb, r, r) = 1.0;
  homegrown_sweep(MT, NT, nb, Q, RY, TAU);
}

/* Hybrid_dormqr.c:20-171 (one rank): C <- Q^T C. */
void orc_dormqr(int MT, int NT, int nb, const double *RY, const double *TAU,
                double *C) {
  homegrown_sweep(MT, NT, nb, C, RY, TAU);
}

/* =========================================================================
 * Generators and layout.
 * ========================================================================= */

double orc_uniform_pm1(uint64_t seed, uint64_t index) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (index + 1ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  double u = (double)(z >> 11) * 0x1.0p-53;
  return 2.0 * u - 1.0;
}

void orc_fill_uniform_cm(int m, int n, uint64_t seed, double *A) {
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++)
      A[(size_t)i + (size_t)j * m] = orc_uniform_pm1(seed, (uint64_t)j * m + i);
}

/* mat_prod.c:157-182 (one rank, non-__CHECK_MODE__ branch) */
void orc_mat_make_glibc(int MT, int NT, int nb, unsigned int seed, double *A) {
  srand(seed);
  for (int j = 0; j < NT; j++)
    for (int jj = 0; jj < nb; jj++)
      for (int i = 0; i < MT; i++)
        for (int ii = 0; ii < nb; ii++)
          IDX(TILE(A, MT, nb, i, j), nb, ii, jj) = (double)((double)rand() / (double)RAND_MAX);
}

void orc_cm_to_tile(int MT, int NT, int nb, const double *cm, double *tile) {
  size_t M = (size_t)MT * nb;
  for (int j = 0; j < NT; j++)
    for (int i = 0; i < MT; i++)
      for (int c = 0; c < nb; c++)
        for (int r = 0; r < nb; r++)
          IDX(TILE(tile, MT, nb, i, j), nb, r, c) =
              cm[(size_t)i * nb + r + ((size_t)j * nb + c) * M];
}

void orc_tile_to_cm(int MT, int NT, int nb, const double *tile, double *cm) {
  size_t M = (size_t)MT * nb;
  for (int j = 0; j < NT; j++)
    for (int i = 0; i < MT; i++)
      for (int c = 0; c < nb; c++)
        for (int r = 0; r < nb; r++)
          cm[(size_t)i * nb + r + ((size_t)j * nb + c) * M] =
              IDX(TILE(tile, MT, nb, i, j), nb, r, c);
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

double orc_time_tile_qr(int MT, int NT, int nb, int ib, uint64_t seed,
                        int nthreads, int reps) {
  size_t M = (size_t)MT * nb, N = (size_t)NT * nb;
  double *cm = (double *)malloc(sizeof(double) * M * N);
  double *A = (double *)malloc(sizeof(double) * M * N);
  double *TAU = (double *)calloc(M * N / nb + 1, sizeof(double));
  orc_fill_uniform_cm((int)M, (int)N, seed, cm);
  double best = 1e30;
  for (int r = 0; r < (reps > 0 ? reps : 1); r++) {
    orc_cm_to_tile(MT, NT, nb, cm, A);
    double t0 = now_s();
    orc_tile_qr(MT, NT, nb, ib, A, TAU, NULL, nthreads);
    double t1 = now_s();
    if (t1 - t0 < best) best = t1 - t0;
  }
  free(cm);
  free(A);
  free(TAU);
  return best;
}
