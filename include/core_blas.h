/*
 * core_blas.h -- PLASMA core_blas tile kernels served by the GPU.
 *
 * The reference's factorisation calls PLASMA 2.x core_blas for its four tile
 * kernels (Hybrid_tile_QR.c:985-990 CORE_dgeqrt, :1135-1141 CORE_dtsqrt,
 * :1276-1283 CORE_dormqr, :1356-1363 CORE_dtsmqr; Hybrid_make_Q.c:142,172,222).
 * These entry points keep PLASMA's names, argument lists and return
 * convention, so the reference's own scheduler can drive the HIP kernels task
 * by task (an A/B harness): link libhybrid3.so instead of -lcoreblas.
 *
 * Arithmetic: PLASMA's (LAPACK dlarfg sign convention, ib-blocked T, upper
 * triangular ib x nb blocks with leading dimension LDT).  Pointers are HOST
 * pointers with any leading dimension; each call copies its tiles to the
 * device, runs one kernel and copies the results back, synchronously.  Calls
 * are reentrant (each takes its own stream and scratch).
 *
 * Supported shapes: square nb x nb tiles with 1 <= nb <= 512, K = nb,
 * 1 <= IB <= nb (IB <= 32 for CORE_dormqr / CORE_dtsmqr at nb other than
 * 128 and 256),
 * side = PlasmaLeft, trans = PlasmaTrans.  Anything else returns -i (i = the
 * offending argument, message in hq_last_error_message()); a device failure
 * returns -1000 - (hq error code).  WORK / LDWORK are not used.
 */
#ifndef HQ_CORE_BLAS_H_
#define HQ_CORE_BLAS_H_

#ifdef __cplusplus
extern "C" {
#endif

/* PLASMA's enum values for side and trans (the CBLAS numbering) */
#define HQ_PLASMA_NOTRANS 111
#define HQ_PLASMA_TRANS 112
#define HQ_PLASMA_LEFT 141
#define HQ_PLASMA_RIGHT 142
#ifndef PlasmaLeft
#define PlasmaNoTrans HQ_PLASMA_NOTRANS
#define PlasmaTrans HQ_PLASMA_TRANS
#define PlasmaLeft HQ_PLASMA_LEFT
#define PlasmaRight HQ_PLASMA_RIGHT
#endif

/* QR of the M x N tile A: R above the diagonal, V below, T (IB x N), TAU. */
int CORE_dgeqrt(int M, int N, int IB, double *A, int LDA, double *T, int LDT, double *TAU, double *WORK);

/* QR of [triu(A1); A2] (A1 N x N, A2 M x N): R in A1, V2 in A2, T, TAU. */
int CORE_dtsqrt(int M, int N, int IB, double *A1, int LDA1, double *A2, int LDA2, double *T, int LDT, double *TAU,
                double *WORK);

/* C (M x N) <- Q^T C, Q from the K reflectors in A (unit lower) and T. */
int CORE_dormqr(int side, int trans, int M, int N, int K, int IB, const double *A, int LDA, const double *T, int LDT,
                double *C, int LDC, double *WORK, int LDWORK);

/* [A1; A2] <- Q^T [A1; A2], Q from the K reflectors [I; V] and T. */
int CORE_dtsmqr(int side, int trans, int M1, int N1, int M2, int N2, int K, int IB, double *A1, int LDA1,
                double *A2, int LDA2, const double *V, int LDV, const double *T, int LDT, double *WORK,
                int LDWORK);

#ifdef __cplusplus
}
#endif
#endif
