// hq_factor.hip -- FP64 tile kernels of the flat-tree tiled Householder QR
// (GEQRT, TSQRT, LARFB, SSRFB in the LAPACK/PLASMA convention used by
// Hybrid_tile_QR.c:985,1135,1276,1356) and the persistent on-device DAG
// scheduler that replaces the TileSch FIFO (TileSch.c, Hybrid_tile_QR.c:905-1461).
//
// Execution model (gfx950, wave64):
//  * one workgroup = 256 threads = 4 waves; one task per workgroup at a time;
//  * LARFB / SSRFB work on a 64-column slab of the target tile: each wave owns
//    16 columns and keeps the whole nb x 16 strip of the updated tile in MFMA
//    accumulator layout (v_mfma_f64_16x16x4f64) for all ib blocks of the task;
//    the V block and T^T of the current ib block are staged in LDS;
//  * the reduction index of W = V^T A2 and of A2 -= V W is permuted so that the
//    accumulator layout of A2 / W doubles as the B operand of the next MFMA, so
//    no data moves between the three products;
//  * GEQRT / TSQRT factor one ib-column panel at a time in LDS (one barrier per
//    column, FP64 throughout), build T in registers and update the rest of the
//    tile with the same MFMA block routine.
#include "hq_internal.h"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace hq {

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// v_mfma_f64_16x16x4f64 lane layout (gfx950): A(i,k) at lane i + 16k,
// B(k,j) at lane j + 16k, D(i,j) at lane j + 16*(i%4), register i/4.
// Feeding A-operand row i from matrix row arow(i) = 4*(i%4) + i/4 makes
// register v of lane group g hold matrix row 4g + v, so an accumulator
// strip of a column is four consecutive doubles per lane (32-byte loads),
// and the same registers serve as the B operand with k index g, register s
// <-> row 4g + s.
__device__ __forceinline__ int arow(int i) { return 4 * (i & 3) + (i >> 2); }

// Butterfly sum: every lane ends with the bitwise-identical total (each level
// adds the same two operands, only in swapped order).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double dlapy2(double x, double y) {
  double xa = fabs(x), ya = fabs(y);
  double w = xa > ya ? xa : ya;
  double z = xa < ya ? xa : ya;
  if (z == 0.0) return w;
  double q = z / w;
  return w * sqrt(1.0 + q * q);
}

template <int NB>
struct Smem {
  static constexpr int LDV = NB + 1;        // odd stride: spreads LDS banks
  static constexpr int LDT = HQ_IBMAX + 1;
  double V[HQ_IBMAX * LDV];   // V block / panel, column-major V[c*LDV + r]
  double TT[HQ_IBMAX * LDT];  // T^T of the block, TT[c*LDT + p] = T(p,c)
  double A1[HQ_IBMAX * LDT];  // TSQRT triangle block, A1[c*LDT + p] = R(p,c)
  double G[HQ_IBMAX * LDT];   // Gram V^T V, G[p*LDT + c]
  double tau[HQ_IBMAX];
  int task;
};

// --------------------------------------------------------------------------
// accumulator-layout strip load/store: lane (g = lane/16, n = lane%16) holds
// rows 16*rb + 4*g + {0..3} of column `col`.
// --------------------------------------------------------------------------
template <int NB>
__device__ __forceinline__ void load_strip(d4 (&acc)[NB / 16], const double *tile, int col,
                                           bool ok, int rb0) {
  const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int rb = 0; rb < NB / 16; rb++) {
    if (ok && rb >= rb0)
      acc[rb] = *reinterpret_cast<const d4 *>(tile + (size_t)col * NB + 16 * rb + 4 * g);
    else
      acc[rb] = d4{0.0, 0.0, 0.0, 0.0};
  }
}

template <int NB>
__device__ __forceinline__ void store_strip(const d4 (&acc)[NB / 16], double *tile, int col,
                                            bool ok, int rb0) {
  const int g = (threadIdx.x & 63) >> 4;
  if (!ok) return;
#pragma unroll
  for (int rb = 0; rb < NB / 16; rb++)
    if (rb >= rb0) *reinterpret_cast<d4 *>(tile + (size_t)col * NB + 16 * rb + 4 * g) = acc[rb];
}

// --------------------------------------------------------------------------
// Apply one block reflector (V, T from LDS) from the left, transposed:
//   W  = [A1 rows] + V^T X        (HAS_A1: SSRFB / TSQRT trailing; else LARFB)
//   W  = T^T W
//   A1 -= W ;  X -= V W
// X is the strip held in `acc` (rows >= 16*rb0 take part).  WB = IBP/16.
// --------------------------------------------------------------------------
template <int NB, int WB, bool HAS_A1>
__device__ __forceinline__ void apply_block(d4 (&acc)[NB / 16], int rb0, const Smem<NB> &S,
                                            double *a1col, int kb, bool ok) {
  constexpr int LDV = Smem<NB>::LDV, LDT = Smem<NB>::LDT;
  const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
  const int pm = arow(m);
  d4 Wa[WB], Wb[WB], A1v[WB];
#pragma unroll
  for (int wb = 0; wb < WB; wb++) {
    A1v[wb] = d4{0.0, 0.0, 0.0, 0.0};
    if (HAS_A1 && ok) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int row = 16 * wb + 4 * g + v;
        if (row < kb) A1v[wb][v] = a1col[row];
      }
    }
    Wa[wb] = A1v[wb];
    Wb[wb] = d4{0.0, 0.0, 0.0, 0.0};
  }
  // W = A1 + V^T X: the k index of MFMA (rb, s) at lane group g is row 16rb+4g+s.
#pragma unroll
  for (int rb = 0; rb < NB / 16; rb++) {
    if (rb < rb0) continue;
#pragma unroll
    for (int s = 0; s < 4; s++) {
#pragma unroll
      for (int wb = 0; wb < WB; wb++) {
        double a = S.V[(16 * wb + pm) * LDV + 16 * rb + 4 * g + s];
        if (rb & 1)
          Wb[wb] = mfma(a, acc[rb][s], Wb[wb]);
        else
          Wa[wb] = mfma(a, acc[rb][s], Wa[wb]);
      }
    }
    // keep the operand loads of one row block next to its MFMAs (stops the
    // scheduler from hoisting every LDS load of the loop and spilling)
    __builtin_amdgcn_sched_group_barrier(0x100, 4 * WB, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 4 * WB, 0);
  }
  d4 W[WB], W2[WB];
#pragma unroll
  for (int wb = 0; wb < WB; wb++) {
    W[wb] = Wa[wb] + Wb[wb];
    W2[wb] = d4{0.0, 0.0, 0.0, 0.0};
  }
  // W2 = T^T W (T^T lower triangular)
#pragma unroll
  for (int wi = 0; wi < WB; wi++)
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int wo = wi; wo < WB; wo++) {
        double a = S.TT[(16 * wo + pm) * LDT + 16 * wi + 4 * g + s];
        W2[wo] = mfma(a, W[wi][s], W2[wo]);
      }
  if (HAS_A1 && ok) {
#pragma unroll
    for (int wb = 0; wb < WB; wb++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int row = 16 * wb + 4 * g + v;
        if (row < kb) a1col[row] = A1v[wb][v] - W2[wb][v];
      }
  }
#pragma unroll
  for (int wb = 0; wb < WB; wb++) W2[wb] = -W2[wb];
  // X -= V W: k index of MFMA (wb, s) at lane group g is W row 16wb+4g+s.
#pragma unroll
  for (int rb = 0; rb < NB / 16; rb++) {
    if (rb < rb0) continue;
#pragma unroll
    for (int wb = 0; wb < WB; wb++)
#pragma unroll
      for (int s = 0; s < 4; s++) {
        double a = S.V[(16 * wb + 4 * g + s) * LDV + 16 * rb + pm];
        acc[rb] = mfma(a, W2[wb][s], acc[rb]);
      }
    __builtin_amdgcn_sched_group_barrier(0x100, 4 * WB, 1);
    __builtin_amdgcn_sched_group_barrier(0x008, 4 * WB, 1);
  }
}

// --------------------------------------------------------------------------
// Panel factorisation in LDS.
//  GEQRT (TS=false): panel columns 0..sb-1 = tile columns ii.., rows ii..NB-1;
//    column c's reflector has its unit element at row d = ii+c.
//  TSQRT (TS=true): triangle block S.A1 (sb x sb) on top of S.V (NB rows);
//    reflector c = [e_c ; V(:,c)].
// Each lane always handles rows lane + 64q, so row data never crosses lanes
// except through the butterfly sums.  One barrier per column: the wave that
// owns column c+1 (c+1 = wave mod 4) updates it first and generates the next
// reflector before the barrier.
// --------------------------------------------------------------------------
template <int NB, bool TS>
__device__ __forceinline__ void gen_reflector(Smem<NB> &S, int c, int d, double (&x)[NB / 64]) {
  constexpr int LDV = Smem<NB>::LDV, LDT = Smem<NB>::LDT;
  const int lane = threadIdx.x & 63;
  double ss = 0.0, xd = 0.0;
#pragma unroll
  for (int q = 0; q < NB / 64; q++) {
    int r = lane + 64 * q;
    if (r > d) ss += x[q] * x[q];
    if (!TS && r == d) xd = x[q];
  }
  ss = wave_sum(ss);
  double alpha = TS ? S.A1[c * LDT + c] : wave_sum(xd);
  double xnorm = sqrt(ss);
  double tau, beta, scal;
  if (xnorm == 0.0) {
    tau = 0.0;
    beta = alpha;
    scal = 1.0;
  } else {
    beta = -copysign(dlapy2(alpha, xnorm), alpha);
    tau = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
  }
#pragma unroll
  for (int q = 0; q < NB / 64; q++) {
    int r = lane + 64 * q;
    if (r > d) {
      x[q] *= scal;
      S.V[c * LDV + r] = x[q];
    }
    if (!TS && r == d) S.V[c * LDV + r] = beta;
  }
  if (lane == 0) {
    S.tau[c] = tau;
    if (TS) S.A1[c * LDT + c] = beta;
  }
}

template <int NB, bool TS>
__device__ __forceinline__ void panel_factor(Smem<NB> &S, int ii, int sb) {
  constexpr int LDV = Smem<NB>::LDV, LDT = Smem<NB>::LDT;
  constexpr int NQ = NB / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (wave == 0) {
    double x[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) x[q] = S.V[lane + 64 * q];
    gen_reflector<NB, TS>(S, 0, TS ? -1 : ii, x);
  }
  __syncthreads();
  for (int c = 0; c < sb; c++) {
    const double tau = S.tau[c];
    const int d = TS ? -1 : ii + c;
    double v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      int r = lane + 64 * q;
      double val = S.V[c * LDV + r];
      v[q] = TS ? val : (r > d ? val : (r == d ? 1.0 : 0.0));
    }
    for (int j = c + 1 + ((wave - c - 1) & 3); j < sb; j += 4) {
      double x[NQ];
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        x[q] = S.V[j * LDV + lane + 64 * q];
        s += v[q] * x[q];
      }
      s = wave_sum(s);
      double top = TS ? S.A1[j * LDT + c] : 0.0;
      double w = tau * (s + top);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        x[q] -= v[q] * w;
        S.V[j * LDV + lane + 64 * q] = x[q];
      }
      if (TS && lane == 0) S.A1[j * LDT + c] = top - w;
      if (j == c + 1) gen_reflector<NB, TS>(S, c + 1, TS ? -1 : ii + c + 1, x);
    }
    __syncthreads();
  }
}

// G = V^T V (upper block triangle) with MFMA; rows < 16*rb0 are zero in V.
template <int NB, int WB>
__device__ __forceinline__ void gram(Smem<NB> &S, int rb0) {
  constexpr int LDV = Smem<NB>::LDV, LDT = Smem<NB>::LDT;
  const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4, wave = threadIdx.x >> 6;
  const int pm = arow(m);
  constexpr int NBLK = WB * (WB + 1) / 2;
  if (wave >= NBLK) return;
  int pb = 0, cb = 0, t = wave;
  for (pb = 0; pb < WB; pb++) {
    if (t < WB - pb) {
      cb = pb + t;
      break;
    }
    t -= WB - pb;
  }
  d4 acc = d4{0.0, 0.0, 0.0, 0.0};
  for (int rb = rb0; rb < NB / 16; rb++)
#pragma unroll
    for (int s = 0; s < 4; s++) {
      double a = S.V[(16 * pb + pm) * LDV + 16 * rb + 4 * g + s];
      double b = S.V[(16 * cb + m) * LDV + 16 * rb + 4 * g + s];
      acc = mfma(a, b, acc);
    }
#pragma unroll
  for (int v = 0; v < 4; v++) S.G[(16 * pb + 4 * g + v) * LDT + 16 * cb + m] = acc[v];
}

// T (upper, sb x sb) from G and tau, LAPACK dlarft('F','C') recursion
//   T(0:c, c) = -tau_c T(0:c, 0:c) G(0:c, c),  T(c,c) = tau_c.
// Lane p of wave 0 owns row p of T, kept in S.TT (TT[c*LDT + p] = T(p,c)), so
// every value a lane reads it wrote itself.  Writes T to global (ld = ib,
// block at column ii), and tau to TAU.
template <int NB>
__device__ __forceinline__ void build_t(Smem<NB> &S, int sb, double *Tg, int ib, double *taug) {
  constexpr int LDT = Smem<NB>::LDT;
  const int p = threadIdx.x;  // wave 0 only
  if (p < HQ_IBMAX) {
    for (int c = 0; c < sb; c++) {
      double t = 0.0;
#pragma unroll 8
      for (int q = 0; q < c; q++) t += S.TT[q * LDT + p] * S.G[q * LDT + c];
      const double tc = S.tau[c];
      S.TT[c * LDT + p] = (p < c) ? -tc * t : (p == c ? tc : 0.0);
    }
    for (int c = sb; c < HQ_IBMAX; c++) S.TT[c * LDT + p] = 0.0;
    if (p < ib)
      for (int c = 0; c < sb; c++) Tg[(size_t)c * ib + p] = S.TT[c * LDT + p];
  }
  if (p < sb) taug[p] = S.tau[p];
}

// --------------------------------------------------------------------------
// Tasks.
// --------------------------------------------------------------------------

// GEQRT = CORE_dgeqrt(nb, nb, ib, A, nb, T, ib, TAU)
template <int NB, int WB>
__device__ __forceinline__ void task_geqrt(Smem<NB> &S, double *A, double *T, double *TAU, int ib) {
  constexpr int LDV = Smem<NB>::LDV;
  constexpr int IBP = 16 * WB;
  const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, wave = tid >> 6;
  for (int ii = 0; ii < NB; ii += ib) {
    const int sb = min(ib, NB - ii);
    for (int e = tid; e < IBP * NB; e += HQ_THREADS) {
      int c = e / NB, r = e - c * NB;
      S.V[c * LDV + r] = (c < sb && r >= ii) ? A[(size_t)(ii + c) * NB + r] : 0.0;
    }
    __syncthreads();
    panel_factor<NB, false>(S, ii, sb);
    // write R / V back, then make the LDS copy unit lower trapezoidal
    for (int e = tid; e < sb * NB; e += HQ_THREADS) {
      int c = e / NB, r = e - c * NB;
      if (r >= ii) {
        A[(size_t)(ii + c) * NB + r] = S.V[c * LDV + r];
        if (r <= ii + c) S.V[c * LDV + r] = (r == ii + c) ? 1.0 : 0.0;
      }
    }
    __syncthreads();
    gram<NB, WB>(S, ii / 16);
    __syncthreads();
    if (wave == 0) build_t<NB>(S, sb, T + (size_t)ii * ib, ib, TAU + ii);
    __syncthreads();
    const int j0 = ii + sb;
    const int ngrp = (NB - j0 + 15) / 16;
    for (int grp = wave; grp < ngrp; grp += 4) {
      int col = j0 + 16 * grp + m;
      bool ok = col < NB;
      d4 acc[NB / 16];
      load_strip<NB>(acc, A, col, ok, ii / 16);
      apply_block<NB, WB, false>(acc, ii / 16, S, nullptr, sb, ok);
      store_strip<NB>(acc, A, col, ok, ii / 16);
    }
    __syncthreads();
  }
}

// TSQRT = CORE_dtsqrt(nb, nb, ib, A1, nb, A2, nb, T, ib, TAU)
template <int NB, int WB>
__device__ __forceinline__ void task_tsqrt(Smem<NB> &S, double *A1, double *A2, double *T, double *TAU, int ib) {
  constexpr int LDV = Smem<NB>::LDV, LDT = Smem<NB>::LDT;
  constexpr int IBP = 16 * WB;
  const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, wave = tid >> 6;
  for (int ii = 0; ii < NB; ii += ib) {
    const int sb = min(ib, NB - ii);
    for (int e = tid; e < IBP * NB; e += HQ_THREADS) {
      int c = e / NB, r = e - c * NB;
      S.V[c * LDV + r] = (c < sb) ? A2[(size_t)(ii + c) * NB + r] : 0.0;
    }
    for (int e = tid; e < IBP * IBP; e += HQ_THREADS) {
      int c = e / IBP, p = e - c * IBP;
      S.A1[c * LDT + p] = (c < sb && p <= c) ? A1[(size_t)(ii + c) * NB + ii + p] : 0.0;
    }
    __syncthreads();
    panel_factor<NB, true>(S, ii, sb);
    for (int e = tid; e < sb * NB; e += HQ_THREADS) {
      int c = e / NB, r = e - c * NB;
      A2[(size_t)(ii + c) * NB + r] = S.V[c * LDV + r];
    }
    for (int e = tid; e < sb * sb; e += HQ_THREADS) {
      int c = e / sb, p = e - c * sb;
      if (p <= c) A1[(size_t)(ii + c) * NB + ii + p] = S.A1[c * LDT + p];
    }
    __syncthreads();
    gram<NB, WB>(S, 0);
    __syncthreads();
    if (wave == 0) build_t<NB>(S, sb, T + (size_t)ii * ib, ib, TAU + ii);
    __syncthreads();
    const int j0 = ii + sb;
    const int ngrp = (NB - j0 + 15) / 16;
    for (int grp = wave; grp < ngrp; grp += 4) {
      int col = j0 + 16 * grp + m;
      bool ok = col < NB;
      d4 acc[NB / 16];
      load_strip<NB>(acc, A2, col, ok, 0);
      apply_block<NB, WB, true>(acc, 0, S, A1 + (size_t)col * NB + ii, sb, ok);
      store_strip<NB>(acc, A2, col, ok, 0);
    }
    __syncthreads();
  }
}

// stage block b of a reflector tile: unit lower (LARFB) or full (SSRFB), and T^T
template <int NB, int WB, bool UNIT>
__device__ __forceinline__ void stage_block(Smem<NB> &S, const double *V, const double *T, int ib,
                                            int i0, int kb) {
  constexpr int LDV = Smem<NB>::LDV, LDT = Smem<NB>::LDT;
  constexpr int IBP = 16 * WB;
  const int tid = threadIdx.x;
  for (int e = tid; e < IBP * NB; e += HQ_THREADS) {
    int c = e / NB, r = e - c * NB;
    double val = 0.0;
    if (c < kb) {
      if (UNIT)
        val = (r > i0 + c) ? V[(size_t)(i0 + c) * NB + r] : (r == i0 + c ? 1.0 : 0.0);
      else
        val = V[(size_t)(i0 + c) * NB + r];
    }
    S.V[c * LDV + r] = val;
  }
  for (int e = tid; e < IBP * IBP; e += HQ_THREADS) {
    int c = e / IBP, p = e - c * IBP;
    S.TT[c * LDT + p] = (c < kb && p <= c) ? T[(size_t)(i0 + c) * ib + p] : 0.0;
  }
}

// LARFB = CORE_dormqr(Left, Trans, nb, nb, nb, ib, V, nb, T, ib, C, nb), one slab
template <int NB, int WB>
__device__ __forceinline__ void task_larfb(Smem<NB> &S, const double *V, const double *T, double *C, int ib,
                           int slab) {
  const int lane = threadIdx.x & 63, m = lane & 15, wave = threadIdx.x >> 6;
  const int col = slab * HQ_SLAB + 16 * wave + m;
  const bool ok = col < NB;
  d4 acc[NB / 16];
  load_strip<NB>(acc, C, col, ok, 0);
  for (int i0 = 0; i0 < NB; i0 += ib) {
    const int kb = min(ib, NB - i0);
    stage_block<NB, WB, true>(S, V, T, ib, i0, kb);
    __syncthreads();
    apply_block<NB, WB, false>(acc, i0 / 16, S, nullptr, kb, ok);
    __syncthreads();
  }
  store_strip<NB>(acc, C, col, ok, 0);
}

// SSRFB = CORE_dtsmqr(Left, Trans, nb, nb, nb, nb, nb, ib, A1, nb, A2, nb, V, nb, T, ib), one slab
template <int NB, int WB>
__device__ __forceinline__ void task_ssrfb(Smem<NB> &S, double *A1, double *A2, const double *V, const double *T,
                           int ib, int slab) {
  const int lane = threadIdx.x & 63, m = lane & 15, wave = threadIdx.x >> 6;
  const int col = slab * HQ_SLAB + 16 * wave + m;
  const bool ok = col < NB;
  d4 acc[NB / 16];
  load_strip<NB>(acc, A2, col, ok, 0);
  for (int i0 = 0; i0 < NB; i0 += ib) {
    const int kb = min(ib, NB - i0);
    stage_block<NB, WB, false>(S, V, T, ib, i0, kb);
    __syncthreads();
    apply_block<NB, WB, true>(acc, 0, S, A1 + (size_t)col * NB + i0, kb, ok);
    __syncthreads();
  }
  store_strip<NB>(acc, A2, col, ok, 0);
}

__device__ __forceinline__ double *tile_ptr(const hq_geom &G, int i, int j) {
  return G.A + ((size_t)i + (size_t)j * G.MT) * (size_t)G.nb * G.nb;
}
__device__ __forceinline__ double *t_ptr(const hq_geom &G, int i, int j) {
  return G.T + ((size_t)i + (size_t)j * G.MT) * (size_t)G.ib * G.nb;
}
__device__ __forceinline__ double *tau_ptr(const hq_geom &G, int i, int j) {
  return G.TAU + ((size_t)i + (size_t)j * G.MT) * (size_t)G.nb;
}

template <int NB, int WB>
__device__ __forceinline__ void run_task(Smem<NB> &S, const hq_geom &G, const hq_task &t) {
  switch (t.kind) {
    case HQ_GEQRT:
      task_geqrt<NB, WB>(S, tile_ptr(G, t.k, t.k), t_ptr(G, t.k, t.k), tau_ptr(G, t.k, t.k), G.ib);
      break;
    case HQ_TSQRT:
      task_tsqrt<NB, WB>(S, tile_ptr(G, t.k, t.k), tile_ptr(G, t.r, t.k), t_ptr(G, t.r, t.k),
                         tau_ptr(G, t.r, t.k), G.ib);
      break;
    case HQ_LARFB:
      task_larfb<NB, WB>(S, tile_ptr(G, t.k, t.k), t_ptr(G, t.k, t.k), tile_ptr(G, t.k, t.j), G.ib,
                         t.slab);
      break;
    case HQ_SSRFB:
      task_ssrfb<NB, WB>(S, tile_ptr(G, t.k, t.j), tile_ptr(G, t.r, t.j), tile_ptr(G, t.r, t.k),
                         t_ptr(G, t.r, t.k), G.ib, t.slab);
      break;
    default:
      break;
  }
}

// Persistent DAG scheduler.  Workgroups claim task indices in order from one
// atomic counter; a claimed task waits (thread 0, with s_sleep back-off) until
// every dependency's flag equals this run's epoch, runs, then publishes its own
// flag with release semantics.  Dependencies have smaller indices and were
// claimed by running workgroups, so every wait ends.
template <int NB, int WB>
__global__ __launch_bounds__(HQ_THREADS) void sched_kernel(hq_geom G, const hq_task *__restrict__ tasks,
                                                          int ntasks, int *flags, int epoch,
                                                          int *counter) {
  __shared__ Smem<NB> S;
  int done = -1;  // task finished by this workgroup in the previous iteration
  for (;;) {
    // Thread 0 publishes the previous task and claims the next one in ONE
    // divergent region.  (Two separate thread-0 regions at the end and the
    // start of the loop body get merged across the back edge by the CFG
    // structurizer into an inner loop that the other lanes never leave.)
    if (threadIdx.x == 0) {
      if (done >= 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&flags[done], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      int t = atomicAdd(counter, 1);
      // counter[1] != 0: a wait timed out somewhere; drain without work
      if (__hip_atomic_load(&counter[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
        t = ntasks;
      if (t < ntasks) {
        const hq_task &tk = tasks[t];
        for (int i = 0; i < tk.ndep && t < ntasks; i++) {
          const int d = tk.dep[i];
          unsigned spins = 0;
          while (t < ntasks &&
                 __hip_atomic_load(&flags[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > (1u << 22)) {  // seconds: a dependency never completed
              __hip_atomic_store(&counter[1], 1 + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              t = ntasks;
            }
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      S.task = t;
    }
    __syncthreads();
    // readfirstlane: the task index, the loop exit and every barrier after it
    // are wave-uniform for the compiler
    const int t = __builtin_amdgcn_readfirstlane(S.task);
    if (t >= ntasks) break;
    const hq_task tk = tasks[t];
    run_task<NB, WB>(S, G, tk);
    __syncthreads();
    done = t;
  }
}

// One task per launch (stream order provides the dependencies).
template <int NB, int WB>
__global__ __launch_bounds__(HQ_THREADS) void task_kernel(hq_geom G, hq_task tk) {
  __shared__ Smem<NB> S;
  run_task<NB, WB>(S, G, tk);
}

// Standalone tile kernels on explicit tile pointers (ld = nb, T ld = ib).
template <int NB, int WB>
__global__ __launch_bounds__(HQ_THREADS) void geqrt_kernel(double *A, double *T, double *TAU, int ib) {
  __shared__ Smem<NB> S;
  task_geqrt<NB, WB>(S, A, T, TAU, ib);
}
template <int NB, int WB>
__global__ __launch_bounds__(HQ_THREADS) void tsqrt_kernel(double *A1, double *A2, double *T,
                                                          double *TAU, int ib) {
  __shared__ Smem<NB> S;
  task_tsqrt<NB, WB>(S, A1, A2, T, TAU, ib);
}
template <int NB, int WB>
__global__ __launch_bounds__(HQ_THREADS) void larfb_kernel(const double *V, const double *T, double *C,
                                                          int ib) {
  __shared__ Smem<NB> S;
  task_larfb<NB, WB>(S, V, T, C, ib, blockIdx.x);
}
template <int NB, int WB>
__global__ __launch_bounds__(HQ_THREADS) void ssrfb_kernel(double *A1, double *A2, const double *V,
                                                          const double *T, int ib) {
  __shared__ Smem<NB> S;
  task_ssrfb<NB, WB>(S, A1, A2, V, T, ib, blockIdx.x);
}

__global__ void reset_counter_kernel(int *counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

// SplitMix64 of (seed, global column-major index) -> uniform [-1, 1), written
// in tile layout.  Bit-identical to the CPU generator used by the tests.
__device__ __forceinline__ double uniform_pm1(uint64_t seed, uint64_t index) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (index + 1ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  double u = (double)(z >> 11) * 0x1.0p-53;
  return 2.0 * u - 1.0;
}

__global__ void fill_uniform_tile_kernel(double *A, int MT, int NT, int nb, uint64_t seed) {
  const size_t ts = (size_t)nb * nb;
  const size_t total = ts * (size_t)MT * NT;
  const size_t M = (size_t)MT * nb;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    size_t t = e / ts, w = e - t * ts;
    size_t ti = t % MT, tj = t / MT;
    size_t c = w / nb, r = w - c * nb;
    size_t row = ti * nb + r, col = tj * nb + c;
    A[e] = uniform_pm1(seed, col * M + row);
  }
}

}  // namespace hq

// ---------------------------------------------------------------------------
// Launch wrappers used by the host runtime (hq_runtime.cpp).
// ---------------------------------------------------------------------------
#include <type_traits>

template <typename Fn>
static hipError_t hq_dispatch(int nb, int ib, Fn &&fn) {
  auto go = [&](auto nbc) -> hipError_t {
    if (ib <= 16)
      fn(nbc, std::integral_constant<int, 1>{});
    else
      fn(nbc, std::integral_constant<int, 2>{});
    return hipGetLastError();
  };
  switch (nb) {
    case 64: return go(std::integral_constant<int, 64>{});
    case 128: return go(std::integral_constant<int, 128>{});
    case 256: return go(std::integral_constant<int, 256>{});
    default: return hipErrorInvalidValue;
  }
}

#define HQ_NBWB                                 \
  constexpr int NB = decltype(nbc)::value;      \
  constexpr int WB = decltype(wbc)::value;

hipError_t hq_launch_sched(const hq_geom &G, const hq_task *tasks, int ntasks, int *flags, int epoch,
                           int *counter, int grid, hipStream_t st) {
  hipLaunchKernelGGL(hq::reset_counter_kernel, dim3(1), dim3(64), 0, st, counter);
  return hq_dispatch(G.nb, G.ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    hipLaunchKernelGGL((hq::sched_kernel<NB, WB>), dim3(grid), dim3(HQ_THREADS), 0, st, G, tasks,
                       ntasks, flags, epoch, counter);
  });
}

hipError_t hq_launch_task(const hq_geom &G, const hq_task &tk, hipStream_t st) {
  return hq_dispatch(G.nb, G.ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    hipLaunchKernelGGL((hq::task_kernel<NB, WB>), dim3(1), dim3(HQ_THREADS), 0, st, G, tk);
  });
}

hipError_t hq_launch_geqrt(int nb, int ib, double *A, double *T, double *TAU, hipStream_t st) {
  return hq_dispatch(nb, ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    hipLaunchKernelGGL((hq::geqrt_kernel<NB, WB>), dim3(1), dim3(HQ_THREADS), 0, st, A, T, TAU, ib);
  });
}
hipError_t hq_launch_tsqrt(int nb, int ib, double *A1, double *A2, double *T, double *TAU,
                           hipStream_t st) {
  return hq_dispatch(nb, ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    hipLaunchKernelGGL((hq::tsqrt_kernel<NB, WB>), dim3(1), dim3(HQ_THREADS), 0, st, A1, A2, T, TAU,
                       ib);
  });
}
hipError_t hq_launch_larfb(int nb, int ib, const double *V, const double *T, double *C,
                           hipStream_t st) {
  return hq_dispatch(nb, ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    hipLaunchKernelGGL((hq::larfb_kernel<NB, WB>), dim3((NB + HQ_SLAB - 1) / HQ_SLAB),
                       dim3(HQ_THREADS), 0, st, V, T, C, ib);
  });
}
hipError_t hq_launch_ssrfb(int nb, int ib, double *A1, double *A2, const double *V, const double *T,
                           hipStream_t st) {
  return hq_dispatch(nb, ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    hipLaunchKernelGGL((hq::ssrfb_kernel<NB, WB>), dim3((NB + HQ_SLAB - 1) / HQ_SLAB),
                       dim3(HQ_THREADS), 0, st, A1, A2, V, T, ib);
  });
}

hipError_t hq_launch_fill_uniform(double *A, int MT, int NT, int nb, uint64_t seed, hipStream_t st) {
  size_t total = (size_t)MT * NT * nb * nb;
  size_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(hq::fill_uniform_tile_kernel, dim3((unsigned)blocks), dim3(256), 0, st, A, MT,
                     NT, nb, seed);
  return hipGetLastError();
}
