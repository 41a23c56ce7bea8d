// hq_homegrown.hip -- the reference's homegrown tile-kernel API
// (tile_kernel.h:28-162, tile_kernel.c:19-716) on the GPU, with its own
// conventions: reflectors with beta = +||.||, v = [1; x/(alpha-beta)],
// TAU = t*(alpha-beta)^2, explicit unit-lower Y, and T stored LOWER
// (T_LAPACK^T).  These calls handle one tile each and are not on the
// factorisation's hot path (the reference calls them from Hybrid_dorgqr /
// Hybrid_dormqr only), so each is one 256-thread workgroup working in global
// memory: every loop of the reference becomes a parallel loop between
// barriers, reductions are wave butterflies.  A thread's dot-product loops
// are unrolled (one accumulator, the same summation order), so that their
// loads are in flight together instead of one memory latency per term.
#include "hq_internal.h"

namespace hqh {

constexpr int NT = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

#define IX(A, ld, r, c) ((A)[(size_t)(r) + (size_t)(c) * (size_t)(ld)])

// out[j] = sum_{p<len} x[p] * Y(p, j) for j < n: one wave per output
__device__ __forceinline__ void dots(const double *x, const double *Yc, int ldy, int len, int n,
                                     double *out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j = wave; j < n; j += NT / 64) {
    double s = 0.0;
    for (int p = lane; p < len; p += 64) s += x[p] * IX(Yc, ldy, p, j);
    s = wave_sum(s);
    if (lane == 0) out[j] = s;
  }
}

// tile_kernel.c:138-218 (tile_sub_dgeqr2)
__device__ void sub_dgeqr2(int row, int col, double *A, int lda, double *TAU, double *Y,
                           double *yA, double *sc) {
  const int tid = threadIdx.x;
  int i = 0;
  for (i = 0; i < col - 1; i++) {
    // yA(0:col-i) = A(i+1:row, i)^T A(i+1:row, i:col)
    dots(&IX(A, lda, i + 1, i), &IX(A, lda, i + 1, i), lda, row - i - 1, col - i, yA);
    __syncthreads();
    if (tid == 0) {
      double a_head = IX(A, lda, i, i), norm_1 = yA[0];
      double b = sqrt(norm_1 + a_head * a_head);
      double y_head = a_head - b;
      sc[0] = y_head;
      sc[1] = 2 / (y_head * y_head + norm_1);
      sc[2] = b;
    }
    __syncthreads();
    const double y_head = sc[0], t = sc[1], b = sc[2];
    for (int j = tid; j < col - i - 1; j += NT) yA[1 + j] += y_head * IX(A, lda, i, i + 1 + j);
    __syncthreads();
    // A(i:row, i+1:col) -= t * u yA(1:)^T, u = [y_head; A(i+1:row, i)]
    const int nr = row - i, nc = col - i - 1;
    for (int e = tid; e < nr * nc; e += NT) {
      const int r = e % nr, j = e / nr;
      const double u = r == 0 ? y_head : IX(A, lda, i + r, i);
      IX(A, lda, i + r, i + 1 + j) += u * (-t * yA[1 + j]);
    }
    __syncthreads();
    const double rs = 1 / y_head;
    for (int r = tid; r < row - i - 1; r += NT) {
      const double v = IX(A, lda, i + 1 + r, i) * rs;
      IX(A, lda, i + 1 + r, i) = v;
      IX(Y, lda, i + 1 + r, i) = v;
    }
    if (tid == 0) {
      IX(A, lda, i, i) = b;
      IX(Y, lda, i, i) = 1;
      TAU[i] = t * y_head * y_head;
    }
    __syncthreads();
  }
  if (row > col) {
    i = col - 1;
    dots(&IX(A, lda, i + 1, i), &IX(A, lda, i + 1, i), lda, row - i - 1, 1, yA);
    __syncthreads();
    if (tid == 0) {
      double a_head = IX(A, lda, i, i), norm_1 = yA[0];
      double b = sqrt(norm_1 + a_head * a_head);
      double y_head = a_head - b;
      sc[0] = y_head;
      sc[1] = 2 / (y_head * y_head + norm_1);
      IX(A, lda, i, i) = b;
    }
    __syncthreads();
    const double y_head = sc[0], t = sc[1];
    const double rs = 1 / y_head;
    for (int r = tid; r < row - i - 1; r += NT) {
      const double v = IX(A, lda, i + 1 + r, i) * rs;
      IX(A, lda, i + 1 + r, i) = v;
      IX(Y, lda, i + 1 + r, i) = v;
    }
    if (tid == 0) {
      IX(Y, lda, i, i) = 1;
      TAU[i] = t * y_head * y_head;
    }
    __syncthreads();
  } else if (row == col) {
    if (tid == 0) IX(Y, lda, i, i) = 0;
    __syncthreads();
  }
}

// tile_kernel.c:220-262 (tile_sub_dlarft): YtY = Y^T Y, diag(T) = TAU, then
//   T(i, r) = -T(i,i) * sum_{c<i} T(c, r) YtY(c, i),  r < i   (row recursion)
__device__ void sub_dlarft(int row, int col, const double *Y, int lda, const double *TAU,
                           double *T, int ldt, double *YtY) {
  const int tid = threadIdx.x;
  for (int j = 0; j < col; j++) dots(&IX(Y, lda, 0, j), Y, lda, row, col, YtY + (size_t)j * col);
  for (int i = tid; i < col; i += NT) IX(T, ldt, i, i) = TAU[i];
  __syncthreads();
  for (int i = 1; i < col; i++) {
    const double ti = -IX(T, ldt, i, i);
    for (int r = tid; r < i; r += NT) {
      double s = 0.0;
#pragma unroll 8
      for (int c = 0; c < i; c++) s += IX(T, ldt, c, r) * YtY[c + (size_t)i * col];
      IX(T, ldt, i, r) = ti * s;
    }
    __syncthreads();
  }
}

// tile_kernel.c:264-311 (tile_sub_dlarfb):
//   W = tril(Y)^T A ; W(0:Y_col) = tril(T) W(0:Y_col) ; A -= Y(:, 0:Y_col) W
__device__ void sub_dlarfb(int row, int Y_col, int A_col, const double *Y, int ldy,
                           const double *T, int ldt, double *A, int lda, double *W, double *W2) {
  const int tid = threadIdx.x;
  for (int e = tid; e < Y_col * A_col; e += NT) {
    const int r = e % Y_col, j = e / Y_col;
    double s = 0.0;
#pragma unroll 8
    for (int c = r; c < row; c++) s += IX(Y, ldy, c, r) * IX(A, lda, c, j);
    W[e] = s;
  }
  __syncthreads();
  for (int e = tid; e < Y_col * A_col; e += NT) {
    const int r = e % Y_col, j = e / Y_col;
    double s = 0.0;
#pragma unroll 8
    for (int c = 0; c <= r; c++) s += IX(T, ldt, r, c) * W[c + (size_t)j * Y_col];
    W2[e] = s;
  }
  __syncthreads();
  for (int e = tid; e < row * A_col; e += NT) {
    const int r = e % row, j = e / row;
    double s = 0.0;
#pragma unroll 8
    for (int c = 0; c < Y_col; c++) s += IX(Y, ldy, r, c) * W2[c + (size_t)j * Y_col];
    IX(A, lda, r, j) -= s;
  }
  __syncthreads();
}

// tile_kernel.c:313-363 (tile_sub_dssrfb):
//   W = A1 ; W(0:Y_col) += Y^T A2 ; W(0:Y_col) = tril(T) W(0:Y_col) ;
//   A1 -= W (all over_row rows) ; A2 -= Y W(0:Y_col)
__device__ void sub_dssrfb(int over_row, int under_row, int Y_col, int A_col, const double *Y,
                           int ldy, const double *T, int ldt, double *A1, int lda1, double *A2,
                           int lda2, double *W, double *W2) {
  const int tid = threadIdx.x;
  for (int e = tid; e < over_row * A_col; e += NT) {
    const int r = e % over_row, j = e / over_row;
    double s = IX(A1, lda1, r, j);
    if (r < Y_col) {
      double d = 0.0;
      for (int c = 0; c < under_row; c++) d += IX(Y, ldy, c, r) * IX(A2, lda2, c, j);
      s += d;
    }
    W[e] = s;
  }
  __syncthreads();
  for (int e = tid; e < over_row * A_col; e += NT) {
    const int r = e % over_row, j = e / over_row;
    double s = W[e];
    if (r < Y_col) {
      s = 0.0;
#pragma unroll 8
      for (int c = 0; c <= r; c++) s += IX(T, ldt, r, c) * W[c + (size_t)j * over_row];
    }
    W2[e] = s;
  }
  __syncthreads();
  for (int e = tid; e < over_row * A_col; e += NT) {
    const int r = e % over_row, j = e / over_row;
    IX(A1, lda1, r, j) -= W2[e];
  }
  for (int e = tid; e < under_row * A_col; e += NT) {
    const int r = e % under_row, j = e / under_row;
    double s = 0.0;
#pragma unroll 8
    for (int c = 0; c < Y_col; c++) s += IX(Y, ldy, r, c) * W2[c + (size_t)j * over_row];
    IX(A2, lda2, r, j) -= s;
  }
  __syncthreads();
}

// tile_kernel.c:654-716 (tile_sub_dtsqr2)
__device__ void sub_dtsqr2(int row, int col, double *A1, int lda1, double *A2, int lda2,
                           double *TAU, double *dv, double *sc) {
  const int tid = threadIdx.x;
  int i;
  for (i = 0; i < col - 1; i++) {
    dots(&IX(A2, lda2, 0, i), &IX(A2, lda2, 0, i), lda2, row, col - i, dv);
    __syncthreads();
    if (tid == 0) {
      double a_head = IX(A1, lda1, i, i), norm_1 = dv[0];
      double b = sqrt(norm_1 + a_head * a_head);
      double y_head = a_head - b;
      sc[0] = y_head;
      sc[1] = 2 / (y_head * y_head + norm_1);
      sc[2] = b;
    }
    __syncthreads();
    const double y_head = sc[0], t = sc[1], b = sc[2];
    for (int j = tid; j < col - i - 1; j += NT) {
      const double d = dv[1 + j] + y_head * IX(A1, lda1, i, i + 1 + j);
      dv[1 + j] = d;
      IX(A1, lda1, i, i + 1 + j) += -t * y_head * d;
    }
    __syncthreads();
    const int nc = col - i - 1;
    for (int e = tid; e < row * nc; e += NT) {
      const int r = e % row, j = e / row;
      IX(A2, lda2, r, i + 1 + j) += IX(A2, lda2, r, i) * (-t * dv[1 + j]);
    }
    __syncthreads();
    const double rs = 1 / y_head;
    for (int r = tid; r < row; r += NT) IX(A2, lda2, r, i) *= rs;
    if (tid == 0) {
      IX(A1, lda1, i, i) = b;
      TAU[i] = t * y_head * y_head;
    }
    __syncthreads();
  }
  dots(&IX(A2, lda2, 0, i), &IX(A2, lda2, 0, i), lda2, row, 1, dv);
  __syncthreads();
  if (tid == 0) {
    double a_head = IX(A1, lda1, i, i), norm_1 = dv[0];
    double b = sqrt(norm_1 + a_head * a_head);
    double y_head = a_head - b;
    sc[0] = y_head;
    sc[1] = 2 / (y_head * y_head + norm_1);
    IX(A1, lda1, i, i) = b;
  }
  __syncthreads();
  const double y_head = sc[0], t = sc[1];
  const double rs = 1 / y_head;
  for (int r = tid; r < row; r += NT) IX(A2, lda2, r, i) *= rs;
  if (tid == 0) TAU[i] = t * y_head * y_head;
  __syncthreads();
}

// tile_kernel.c:28-57 (tile_dgeqrt)
__global__ __launch_bounds__(NT) void dgeqrt_kernel(int row, int col, int block, double *A, int lda,
                                                   double *TAU, double *Y, double *T, int ldt,
                                                   double *work) {
  __shared__ double vec[4096 + 8];
  double *sc = vec + 4096;
  double *W = work, *W2 = work + (size_t)row * col;
  int i;
  for (i = 0; i < col - block; i += block) {
    sub_dgeqr2(row - i, block, &IX(A, lda, i, i), lda, TAU + i, &IX(Y, lda, i, i), vec, sc);
    sub_dlarft(row - i, block, &IX(Y, lda, i, i), lda, TAU + i, &IX(T, ldt, i, i), ldt, W);
    __syncthreads();
    sub_dlarfb(row - i, block, col - i - block, &IX(Y, lda, i, i), lda, &IX(T, ldt, i, i), ldt,
               &IX(A, lda, i, i + block), lda, W, W2);
  }
  sub_dgeqr2(row - i, col - i, &IX(A, lda, i, i), lda, TAU + i, &IX(Y, lda, i, i), vec, sc);
  sub_dlarft(row - i, col - i, &IX(Y, lda, i, i), lda, TAU + i, &IX(T, ldt, i, i), ldt, W);
}

// tile_kernel.c:60-104 (tile_dtsqrt)
__global__ __launch_bounds__(NT) void dtsqrt_kernel(int row, int col, int block, double *A1,
                                                   int lda1, double *A2, int lda2, double *TAU,
                                                   double *T, int ldt, double *work) {
  __shared__ double vec[4096 + 8];
  double *sc = vec + 4096;
  double *W = work, *W2 = work + (size_t)(row > col ? row : col) * col;
  int i;
  for (i = 0; i < col - block; i += block) {
    sub_dtsqr2(row, block, &IX(A1, lda1, i, i), lda1, &IX(A2, lda2, 0, i), lda2, TAU + i, vec, sc);
    sub_dlarft(row, block, &IX(A2, lda2, 0, i), lda2, TAU + i, &IX(T, ldt, i, i), ldt, W);
    __syncthreads();
    sub_dssrfb(block, row, block, col - i - block, &IX(A2, lda2, 0, i), lda2, &IX(T, ldt, i, i),
               ldt, &IX(A1, lda1, i, i + block), lda1, &IX(A2, lda2, 0, i + block), lda2, W, W2);
  }
  if (col % block > 0) {
    i = block * (col / block);
    sub_dtsqr2(row, col % block, &IX(A1, lda1, i, i), lda1, &IX(A2, lda2, 0, i), lda2, TAU + i,
               vec, sc);
    sub_dlarft(row, col % block, &IX(A2, lda2, 0, i), lda2, TAU + i, &IX(T, ldt, i, i), ldt, W);
  } else {
    i = col - block;
    sub_dtsqr2(row, block, &IX(A1, lda1, i, i), lda1, &IX(A2, lda2, 0, i), lda2, TAU + i, vec, sc);
    sub_dlarft(row, block, &IX(A2, lda2, 0, i), lda2, TAU + i, &IX(T, ldt, i, i), ldt, W);
  }
}

// tile_kernel.c:106-119 (tile_dlarfb)
__global__ __launch_bounds__(NT) void dlarfb_kernel(int row, int col, double *A, int lda,
                                                   const double *Y, const double *T, int ldt,
                                                   double *work) {
  sub_dlarfb(row, col, col, Y, lda, T, ldt, A, lda, work, work + (size_t)row * col);
}


// tile_kernel.c:368-437 (tile_dlarft): block < 2 -> full recursion;
// 2 <= block < col -> combine the diagonal blocks already in T with the
// reference's loop bounds; block >= col -> nothing.
__global__ __launch_bounds__(NT) void dlarft_kernel(int row, int col, int block, const double *Y,
                                                   int lda, const double *TAU, double *T, int ldt,
                                                   double *work) {
  const int tid = threadIdx.x;
  if (block < 2) {
    sub_dlarft(row, col, Y, lda, TAU, T, ldt, work);
    return;
  }
  if (block >= col) return;
  const int ldyty = (col / block) * block + block;
  double *YtY = work;                       // row-major col x col, ld ldyty
  double *T1YtY = work + (size_t)row * col;  // row-major i x block
  for (int e = tid; e < col * col; e += NT) {
    const int r = e / col, c = e % col;
    double s = 0.0;
#pragma unroll 8
    for (int p = 0; p < row; p++) s += IX(Y, lda, p, r) * IX(Y, lda, p, c);
    YtY[(size_t)r * ldyty + c] = s;
  }
  __syncthreads();
  for (int i = block; i < col - block; i += block) {
    for (int e = tid; e < i * block; e += NT) {
      const int r = e / block, c = e % block;
      double s = 0.0;
#pragma unroll 8
      for (int p = 0; p < i; p++) s += IX(T, ldt, p, r) * YtY[(size_t)p * ldyty + i + c];
      T1YtY[(size_t)r * block + c] = s;
    }
    __syncthreads();
    for (int e = tid; e < i * block; e += NT) {
      const int r = e / block, c = e % block;
      double s = 0.0;
#pragma unroll 8
      for (int p = 0; p < block; p++) s += T1YtY[(size_t)r * block + p] * IX(T, ldt, i + c, i + p);
      IX(T, ldt, i + c, r) = -s;
    }
    __syncthreads();
  }
  for (int i = (col / block) * block; i < col; i++) {
    const double ti = -IX(T, ldt, i, i);
    for (int r = tid; r < i; r += NT) {
      double s = 0.0;
#pragma unroll 8
      for (int c = 0; c < i; c++) s += IX(T, ldt, c, r) * YtY[(size_t)c * ldyty + i];
      IX(T, ldt, i, r) = ti * s;
    }
    __syncthreads();
  }
}

// ---- Hybrid_make_I / Hybrid_dgemm (verification helpers of main.c) ----

// C(i,j) = sum_k A(k,i)^T B(k,j) on MT x NT tile grids of nb x nb tiles
// (Hybrid_dgemm.c:38-67).  One workgroup per (tile of C, 64 x 64 block).
__global__ __launch_bounds__(NT) void tile_gemm_tn_kernel(const double *A, const double *B,
                                                         double *Cm, int MT, int NT_, int nb) {
  __shared__ double As[16][65], Bs[16][65];
  const int ti = blockIdx.z % MT, tj = blockIdx.z / MT;
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // 16 x 16 threads, 4 x 4 each
  const size_t ts = (size_t)nb * nb;
  double acc[4][4] = {};
  for (int k = 0; k < MT; k++) {
    const double *At = A + ((size_t)k + (size_t)ti * MT) * ts;  // A(k, ti)
    const double *Bt = B + ((size_t)k + (size_t)tj * MT) * ts;  // B(k, tj)
    for (int p0 = 0; p0 < nb; p0 += 16) {
      for (int e = threadIdx.x; e < 16 * 64; e += NT) {
        const int pp = e % 16, q = e / 16;
        const bool pin = p0 + pp < nb;
        As[pp][q] = (pin && r0 + q < nb) ? At[(size_t)(r0 + q) * nb + p0 + pp] : 0.0;
        Bs[pp][q] = (pin && c0 + q < nb) ? Bt[(size_t)(c0 + q) * nb + p0 + pp] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int pp = 0; pp < 16; pp++)
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
          for (int b = 0; b < 4; b++) acc[a][b] += As[pp][ty * 4 + a] * Bs[pp][tx * 4 + b];
      __syncthreads();
    }
  }
  double *Ct = Cm + ((size_t)ti + (size_t)tj * MT) * ts;
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) {
      const int r = r0 + ty * 4 + a, c = c0 + tx * 4 + b;
      if (r < nb && c < nb) Ct[(size_t)c * nb + r] = acc[a][b];
    }
}

#undef IX
}  // namespace hqh

hipError_t hqh_launch_dgeqrt(int row, int col, int block, double *A, int lda, double *TAU,
                             double *Y, double *T, int ldt, double *work, hipStream_t st) {
  hipLaunchKernelGGL(hqh::dgeqrt_kernel, dim3(1), dim3(hqh::NT), 0, st, row, col, block, A, lda,
                     TAU, Y, T, ldt, work);
  return hipGetLastError();
}
hipError_t hqh_launch_dtsqrt(int row, int col, int block, double *A1, int lda1, double *A2,
                             int lda2, double *TAU, double *T, int ldt, double *work,
                             hipStream_t st) {
  hipLaunchKernelGGL(hqh::dtsqrt_kernel, dim3(1), dim3(hqh::NT), 0, st, row, col, block, A1, lda1,
                     A2, lda2, TAU, T, ldt, work);
  return hipGetLastError();
}
hipError_t hqh_launch_dlarfb(int row, int col, double *A, int lda, const double *Y,
                             const double *T, int ldt, double *work, hipStream_t st) {
  hipLaunchKernelGGL(hqh::dlarfb_kernel, dim3(1), dim3(hqh::NT), 0, st, row, col, A, lda, Y, T,
                     ldt, work);
  return hipGetLastError();
}
hipError_t hqh_launch_dlarft(int row, int col, int block, const double *Y, int lda,
                             const double *TAU, double *T, int ldt, double *work, hipStream_t st) {
  hipLaunchKernelGGL(hqh::dlarft_kernel, dim3(1), dim3(hqh::NT), 0, st, row, col, block, Y, lda,
                     TAU, T, ldt, work);
  return hipGetLastError();
}
hipError_t hqh_launch_tile_gemm_tn(const double *A, const double *B, double *Cm, int MT, int NT,
                                   int nb, hipStream_t st) {
  const int nblk = (nb + 63) / 64;
  hipLaunchKernelGGL(hqh::tile_gemm_tn_kernel, dim3(nblk, nblk, MT * NT), dim3(hqh::NT), 0, st, A,
                     B, Cm, MT, NT, nb);
  return hipGetLastError();
}
