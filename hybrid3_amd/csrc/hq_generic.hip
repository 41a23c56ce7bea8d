// hq_generic.hip -- tile kernels and DAG scheduler for tile sizes without a
// register-resident specialisation: any nb up to 512 other than 64, 128 and
// 256 (e.g. nb = 384 or 512, or an nb that is not a multiple of 16).
//
// The task set, the DAG and the results are those of the specialised kernels
// (flat-tree GEQRT / TSQRT / LARFB / SSRFB, PLASMA's T layout with ld = ib,
// Hybrid_tile_QR.c:960-1363); only the data movement differs.  At these sizes
// the nb x 16 strip of a 64-column slab does not fit a wave's registers next
// to its operands (512 rows x 16 columns is 128 doubles per lane), so:
//   * a panel block (<= 32 columns of <= 512 rows) is factored in LDS: one
//     pass per column computes the column's squared norm and its dot products
//     with every column right of it at once (8 row groups per column), and one
//     pass applies the reflector;
//   * a block reflector is applied to a slab of <= 64 columns by two sweeps
//     over 64-row chunks: W = V^T X (X chunk and V chunk staged in LDS, each
//     thread 4 reflectors of one column), then W2 = T^T W in LDS, then
//     X -= V W2 (each thread one row and 8 columns); each sweep keeps the
//     next chunk's loads in flight.  Workgroups of 8 waves (two per SIMD).
//     The slab makes three passes over memory per ib block instead of
//     staying in registers, so these sizes run well below the specialised
//     ones (16384^2: about 9 TFLOP/s at nb = 384 / 512 against 46.7 at 256).
// FP64 everywhere (reduced-precision updates exist for nb = 128 and 256).
#include <hip/hip_runtime.h>

#include "hq_internal.h"

namespace hqg {

constexpr int NBMAX = 512;
constexpr int IBM = HQ_IBMAX;  // reflectors per block (the run's ib <= 32)
constexpr int LDP = NBMAX + 1; // panel column stride
constexpr int LDT = IBM + 1;   // T / Gram / triangle column stride
constexpr int CH = 64;         // rows per chunk of a block application
constexpr int LDX = CH + 1;    // X chunk column stride
constexpr int LDVC = IBM + 2;  // V chunk row stride (16-byte aligned rows)
constexpr int NT = 512;        // threads per workgroup: 8 waves, two per SIMD, to keep more loads in flight
constexpr int NW = NT / 64;    // waves
constexpr int PR = IBM / NW;   // reflectors per thread in W = V^T X
constexpr int QC = HQ_SLAB / NW; // columns per thread in X -= V W2
constexpr int RPT = (NBMAX + NT - 1) / NT;  // panel rows per thread
static_assert(PR * NW == IBM && QC * NW == HQ_SLAB && PR % 2 == 0 && QC % 2 == 0, "thread mappings");
static_assert(CH == 64, "chunk rows are the lanes of a wave");

struct Smem {
  union {
    double P[IBM * LDP];  // panel block: column c at P + c*LDP, row i = panel row i
    struct {
      alignas(16) double V[CH * LDVC];     // V chunk: V(i, p) at V[i*LDVC + p]
      alignas(16) double X[HQ_SLAB * LDX]; // X chunk: X(i, q) at X[q*LDX + i]
      alignas(16) double W[IBM * HQ_SLAB]; // W / W2: W(p, q) at W[p*HQ_SLAB + q]
    } u;
  };
  double R[IBM * LDT];   // TSQRT: triangle block of R(k,k), R(p, c) at R[c*LDT + p]
  double G[IBM * LDT];   // Gram v_p^T v_c at G[c*LDT + p]
  double TT[IBM * LDT];  // T(p, c) at TT[c*LDT + p]
  double tau[IBM];
  double red[NW * IBM];  // per-wave partial sums of a column step
  int task;
  hq_task next;
};

__device__ __forceinline__ double *tile_ptr(const hq_geom &G, int i, int j) {
  return G.A + ((size_t)i + (size_t)j * G.MT) * (size_t)G.nb * G.nb;
}
__device__ __forceinline__ double *xtile_ptr(const hq_geom &G, int i, int j) {
  return G.X + ((size_t)i + (size_t)j * G.MT) * (size_t)G.nb * G.nb;
}
__device__ __forceinline__ double *t_ptr(const hq_geom &G, int i, int j) {
  return G.T + ((size_t)i + (size_t)j * G.MT) * (size_t)G.ib * G.nb;
}
__device__ __forceinline__ double *tau_ptr(const hq_geom &G, int i, int j) {
  return G.TAU + ((size_t)i + (size_t)j * G.MT) * (size_t)G.nb;
}

// T of the block from the Gram matrix in S.G and tau in S.tau (dlarft,
// forward, columnwise): T(c, c) = tau_c, T(0:c, c) = -tau_c T(0:c, 0:c)
// G(0:c, c).  Wave 0, lane p owns row p.  T (ld ib, rows below the diagonal
// zero) and tau to global.
__device__ void build_t(Smem &S, int kb, double *Tb, int ib, double *taug) {
  const int p = threadIdx.x;
  if (p >= 64) return;
  for (int c = 0; c < IBM; c++) {
    double val = 0.0;
    if (c < kb && p < c) {
      double s = 0.0;
      for (int q = p; q < c; q++) s += S.TT[q * LDT + p] * S.G[c * LDT + q];
      val = -S.tau[c] * s;
    } else if (c < kb && p == c) {
      val = S.tau[c];
    }
    if (p < IBM) S.TT[c * LDT + p] = val;
  }
  if (p < ib)
    for (int c = 0; c < kb; c++) Tb[(size_t)c * ib + p] = p < IBM ? S.TT[c * LDT + p] : 0.0;
  if (taug && p < kb) taug[p] = S.tau[p];
}

// Householder QR of the block of columns ii .. ii+kb-1 (kb = min(ib, nb-ii)):
//   GEQRT (!TS): rows ii.. of tile A (dgeqr2; R on and above the diagonal,
//                V below it);
//   TSQRT (TS):  [R(ii:ii+kb, ii:ii+kb) of the diagonal tile A1; tile A],
//                the TSQRT reflectors (identity on top, V in A).
// Writes the block back, its T block (T + ii*ib, ld ib) and TAU(ii..).
// Leaves V (rows of the panel) in S.P and T in S.TT.
template <bool TS>
__device__ void panel(Smem &S, double *A, double *A1, double *T, double *TAU, int nb, int ib, int ii) {
  const int tid = threadIdx.x;
  const int kb = min(ib, nb - ii);
  const int r0 = TS ? 0 : ii;  // first tile row of the panel
  const int h = nb - r0;       // panel rows
  for (int e = tid; e < kb * h; e += NT) {
    const int c = e / h, i = e - c * h;
    S.P[c * LDP + i] = A[(size_t)(ii + c) * nb + r0 + i];
  }
  if (TS)
    for (int e = tid; e < kb * kb; e += NT) {
      const int c = e / kb, p = e - c * kb;
      S.R[c * LDT + p] = p <= c ? A1[(size_t)(ii + c) * nb + ii + p] : 0.0;
    }
  __syncthreads();
  for (int c = 0; c < kb; c++) {
    const int i0 = TS ? 0 : c + 1;  // rows of the reflector's tail
    {
      // squared norm of the tail (slot j = c) and its dot products with the
      // tails of the columns right of it (slots j > c): NT/32 row groups per
      // slot
      const int j = tid & 31, rg = tid >> 5;
      double s = 0.0;
      if (j >= c && j < kb) {
        const double *x = S.P + c * LDP, *y = S.P + j * LDP;
        for (int i = i0 + rg; i < h; i += NT / 32) s += x[i] * y[i];
      }
      s += __shfl_xor(s, 32, 64);
      if ((tid & 63) < 32) S.red[(tid >> 6) * IBM + j] = s;
    }
    __syncthreads();
    // dlarfg: beta = -sign(alpha) ||(alpha, x)||, tau = (beta - alpha) / beta,
    // v = x / (alpha - beta); a zero tail gives tau = 0 and leaves x as it is
    double sigma = 0.0;
#pragma unroll
    for (int g = 0; g < NW; g++) sigma += S.red[g * IBM + c];
    const double alpha = TS ? S.R[c * LDT + c] : S.P[c * LDP + c];
    double tau = 0.0, beta = alpha, scale = 1.0;
    if (sigma != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
      tau = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    // the tail rows: v_i = x_i * scale, then y_j -= tau v_i w_j with
    // w_j = top_j + v^T y_j = top_j + scale d_j (top = row c of the panel)
    double v[RPT];
    int rows[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      rows[r] = i0 + tid + r * NT;
      v[r] = rows[r] < h ? S.P[c * LDP + rows[r]] * scale : 0.0;
    }
    if (tau != 0.0) {
      for (int j = c + 1; j < kb; j++) {
        double d = 0.0;
#pragma unroll
        for (int g = 0; g < NW; g++) d += S.red[g * IBM + j];
        const double top = TS ? S.R[j * LDT + c] : S.P[j * LDP + c];
        const double tw = tau * (top + scale * d);
#pragma unroll
        for (int r = 0; r < RPT; r++)
          if (rows[r] < h) S.P[j * LDP + rows[r]] -= v[r] * tw;
      }
    }
#pragma unroll
    for (int r = 0; r < RPT; r++)
      if (rows[r] < h) S.P[c * LDP + rows[r]] = v[r];
    __syncthreads();  // every thread has read row c (top) before it changes
    // row c: R(c, j) -= tau w_j right of the column, R(c, c) = beta
    if (tid < kb && tid >= c) {
      const int j = tid;
      double *top = TS ? &S.R[j * LDT + c] : &S.P[j * LDP + c];
      if (j == c) {
        *top = beta;
      } else if (tau != 0.0) {
        double d = 0.0;
#pragma unroll
        for (int g = 0; g < NW; g++) d += S.red[g * IBM + j];
        *top -= tau * (*top + scale * d);
      }
    }
    if (tid == 0) S.tau[c] = tau;
    __syncthreads();
  }
  // write back
  for (int e = tid; e < kb * h; e += NT) {
    const int c = e / h, i = e - c * h;
    A[(size_t)(ii + c) * nb + r0 + i] = S.P[c * LDP + i];
  }
  if (TS)
    for (int e = tid; e < kb * kb; e += NT) {
      const int c = e / kb, p = e - c * kb;
      if (p <= c) A1[(size_t)(ii + c) * nb + ii + p] = S.R[c * LDT + p];
    }
  // Gram G(p, c) = v_p^T v_c, p < c: one wave per pair.  GEQRT: v_p is 1 at
  // panel row p and v_c is 1 at row c, zero above.
  {
    const int lane = tid & 63, wave = tid >> 6;
    const int npair = kb * (kb - 1) / 2;
    for (int e = wave; e < npair; e += NT / 64) {
      int c = 1;
      while ((c + 1) * c / 2 <= e) c++;
      const int p = e - c * (c - 1) / 2;
      const double *vp = S.P + p * LDP, *vc = S.P + c * LDP;
      double s = 0.0;
      for (int i = (TS ? 0 : c + 1) + lane; i < h; i += 64) s += vp[i] * vc[i];
      for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
      if (lane == 0) S.G[c * LDT + p] = TS ? s : s + vp[c];
    }
  }
  __syncthreads();
  build_t(S, kb, T + (size_t)ii * ib, ib, TAU + ii);
  __syncthreads();
}

// Entry (row, p) of the block's reflector matrix V (rows of the panel, i.e.
// tile row r0 + row): the stored tile value, or the unit lower trapezoid's 1
// and 0 for GEQRT / LARFB.
template <bool TS>
__device__ __forceinline__ double v_at(const double *Vt, int nb, int ii, int r0, int p, int row) {
  if (TS || row > p) return Vt[(size_t)(ii + p) * nb + r0 + row];
  return row == p ? 1.0 : 0.0;
}

// X <- H^T X for the block reflector H = I - V T V^T of the block at column ii
// (kb reflectors of tile Vt, T block Tb with ld ib), on X's columns
// c0 .. c1-1 (c1 - c0 <= 64):
//   !TS: V unit lower trapezoidal (rows ii.. of Vt), X rows ii.. of tile X;
//   TS:  [X1 rows ii..ii+kb; X] with V = [I; Vt columns], X all rows.
// Each sweep over the 64-row chunks keeps the next chunk's global loads in
// flight (registers) while the current chunk is computed from LDS.
template <bool TS>
__device__ void apply_block(Smem &S, const double *Vt, const double *Tb, int ib, double *X, double *X1, int nb,
                            int ii, int kb, int c0, int c1) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r0 = TS ? 0 : ii;
  const int h = nb - r0;
  const int w = c1 - c0;
  for (int e = tid; e < IBM * IBM; e += NT) {
    const int c = e / IBM, p = e - c * IBM;
    S.TT[c * LDT + p] = (c < kb && p <= c) ? Tb[(size_t)c * ib + p] : 0.0;
  }
  // chunk loads: V(row, p) for p = wave + NW k (k < PR) and X(row, qq) for
  // qq = wave + NW k (k < QC), row = b0 + lane; zero outside the block
  double vreg[PR], xreg[QC];
  auto load_v = [&](int b0) {
    const int row = b0 + lane;
#pragma unroll
    for (int k = 0; k < PR; k++) {
      const int p = wave + NW * k;
      vreg[k] = (p < kb && row < h) ? v_at<TS>(Vt, nb, ii, r0, p, row) : 0.0;
    }
  };
  auto load_x = [&](int b0) {
    const int row = b0 + lane;
#pragma unroll
    for (int k = 0; k < QC; k++) {
      const int qq = wave + NW * k;
      xreg[k] = (qq < w && row < h) ? X[(size_t)(c0 + qq) * nb + r0 + row] : 0.0;
    }
  };
  // W = V^T X (+ X1): thread (q, pg) sums reflectors PR pg .. PR pg + PR-1 of
  // column q
  const int q = lane, pg = wave;
  double acc[PR];
#pragma unroll
  for (int pp = 0; pp < PR; pp++) acc[pp] = 0.0;
  load_v(0);
  load_x(0);
  for (int b0 = 0; b0 < h; b0 += CH) {
    __syncthreads();  // the previous chunk has been read
#pragma unroll
    for (int k = 0; k < PR; k++) S.u.V[lane * LDVC + wave + NW * k] = vreg[k];
#pragma unroll
    for (int k = 0; k < QC; k++) S.u.X[(wave + NW * k) * LDX + lane] = xreg[k];
    __syncthreads();
    if (b0 + CH < h) {
      load_v(b0 + CH);
      load_x(b0 + CH);
    }
    // rows past the block were staged as zeros: a fixed trip count
#pragma unroll 8
    for (int i = 0; i < CH; i++) {
      const double x = S.u.X[q * LDX + i];
      const double2 *vr = reinterpret_cast<const double2 *>(S.u.V + i * LDVC + PR * pg);
#pragma unroll
      for (int pp = 0; pp < PR / 2; pp++) {
        const double2 v2 = vr[pp];
        acc[2 * pp] += v2.x * x;
        acc[2 * pp + 1] += v2.y * x;
      }
    }
  }
  if (TS && q < w) {
#pragma unroll
    for (int pp = 0; pp < PR; pp++) {
      const int p = PR * pg + pp;
      if (p < kb) acc[pp] += X1[(size_t)(c0 + q) * nb + ii + p];
    }
  }
  __syncthreads();  // the last chunk has been read
#pragma unroll
  for (int pp = 0; pp < PR; pp++) S.u.W[(PR * pg + pp) * HQ_SLAB + q] = acc[pp];
  __syncthreads();
  // W2 = T^T W (W2(p, q) = sum_{r <= p} T(r, p) W(r, q)); TS: X1 -= W2
  double w2[PR];
#pragma unroll
  for (int pp = 0; pp < PR; pp++) {
    const int p = PR * pg + pp;
    double s = 0.0;
    for (int r = 0; r <= p; r++) s += S.TT[p * LDT + r] * S.u.W[r * HQ_SLAB + q];
    w2[pp] = s;
  }
  __syncthreads();
#pragma unroll
  for (int pp = 0; pp < PR; pp++) {
    const int p = PR * pg + pp;
    S.u.W[p * HQ_SLAB + q] = w2[pp];
    if (TS && q < w && p < kb) X1[(size_t)(c0 + q) * nb + ii + p] -= w2[pp];
  }
  // X -= V W2: thread (row i of the chunk, QC-column group qg) with its QC
  // entries of X in registers (independent FMA chains); the V chunk staged
  // column-major in LDS (lane-consecutive reads), W2 read as broadcast
  // 16-byte LDS loads.  The next chunk's V and X loads are in flight during
  // the arithmetic.
  const int i = lane, qg = wave;
  const int nq = min(QC, w - QC * qg);  // this thread's live columns (may be <= 0)
  double *VT = S.u.X;                   // V chunk, V(i, p) at VT[p*LDX + i]
  double xn[QC];
  auto load_rows = [&](int b0) {
    const int row = b0 + i;
#pragma unroll
    for (int qq = 0; qq < QC; qq++)
      xn[qq] = (qq < nq && row < h) ? X[(size_t)(c0 + QC * qg + qq) * nb + r0 + row] : 0.0;
  };
  load_v(0);
  load_rows(0);
  for (int b0 = 0; b0 < h; b0 += CH) {
    __syncthreads();  // W2 written / the previous chunk's V read
#pragma unroll
    for (int k = 0; k < PR; k++) VT[(wave + NW * k) * LDX + lane] = vreg[k];
    double x[QC];
#pragma unroll
    for (int qq = 0; qq < QC; qq++) x[qq] = xn[qq];
    __syncthreads();
    const bool live = b0 + i < h && nq > 0;
    if (b0 + CH < h) {
      load_v(b0 + CH);
      load_rows(b0 + CH);
    }
#pragma unroll 4
    for (int p = 0; p < IBM; p++) {
      const double v = VT[p * LDX + i];
      const double2 *wr = reinterpret_cast<const double2 *>(S.u.W + p * HQ_SLAB + QC * qg);
#pragma unroll
      for (int qq = 0; qq < QC / 2; qq++) {
        const double2 w2v = wr[qq];
        x[2 * qq] -= v * w2v.x;
        x[2 * qq + 1] -= v * w2v.y;
      }
    }
    if (live) {
      double *xrow = X + (size_t)(c0 + QC * qg) * nb + r0 + b0 + i;
#pragma unroll
      for (int qq = 0; qq < QC; qq++)
        if (qq < nq) xrow[(size_t)qq * nb] = x[qq];
    }
  }
  __syncthreads();
}

// The block's reflectors applied to every column right of it in the tile
// up to column `live` (whole-tile GEQRT / TSQRT), 64 columns at a time.
template <bool TS>
__device__ void panel_rest(Smem &S, double *A, double *A1, const double *T, int nb, int ib, int ii, int live) {
  const int kb = min(ib, nb - ii);
  for (int c0 = ii + kb; c0 < live; c0 += HQ_SLAB)
    apply_block<TS>(S, A, T + (size_t)ii * ib, ib, A, A1, nb, ii, kb, c0, min(live, c0 + HQ_SLAB));
}

// Columns of tile column j that hold matrix entries (hq_geom::ncol).
__device__ __forceinline__ int live_cols(const hq_geom &G, int j) {
  return G.ncol > 0 ? min(G.nb, max(0, G.ncol - j * G.nb)) : G.nb;
}

// TAU = 0 and T = 0 for the panel blocks from column ii on, wholly inside a
// batch matrix's zero padding (their reflectors are identities).
__device__ void pad_panel_blocks(double *T, double *TAU, int nb, int ib, int ii) {
  for (int e = threadIdx.x; e < (nb - ii) * ib; e += NT) T[(size_t)ii * ib + e] = 0.0;
  for (int c = ii + threadIdx.x; c < nb; c += NT) TAU[c] = 0.0;
}

// One task of the DAG (the task kinds and fields of hq_task).  Tasks wholly
// inside a batch matrix's zero padding write only the zeros their outputs
// hold, or nothing; updates stop at the last column that holds entries.
__device__ void run_task(Smem &S, const hq_geom &G, const hq_task &t) {
  const int nb = G.nb, ib = G.ib;
  switch (t.kind) {
    case HQ_GEQRT: {
      double *A = tile_ptr(G, t.k, t.k), *T = t_ptr(G, t.k, t.k), *TAU = tau_ptr(G, t.k, t.k);
      const int live = live_cols(G, t.k);
      if (t.j >= 0) {
        if (t.j * ib >= live)
          pad_panel_blocks(T, TAU, nb, ib, t.j * ib);
        else
          panel<false>(S, A, nullptr, T, TAU, nb, ib, t.j * ib);
      } else {
        int ii = 0;
        for (; ii < live; ii += ib) {
          panel<false>(S, A, nullptr, T, TAU, nb, ib, ii);
          panel_rest<false>(S, A, nullptr, T, nb, ib, ii, live);
        }
        if (ii < nb) pad_panel_blocks(T, TAU, nb, ib, ii);
      }
      break;
    }
    case HQ_TSQRT: {
      double *A1 = tile_ptr(G, t.k, t.k), *A2 = tile_ptr(G, t.r, t.k);
      double *T = t_ptr(G, t.r, t.k), *TAU = tau_ptr(G, t.r, t.k);
      const int live = live_cols(G, t.k);
      if (t.j >= 0) {
        if (t.j * ib >= live)
          pad_panel_blocks(T, TAU, nb, ib, t.j * ib);
        else
          panel<true>(S, A2, A1, T, TAU, nb, ib, t.j * ib);
      } else {
        int ii = 0;
        for (; ii < live; ii += ib) {
          panel<true>(S, A2, A1, T, TAU, nb, ib, ii);
          panel_rest<true>(S, A2, A1, T, nb, ib, ii, live);
        }
        if (ii < nb) pad_panel_blocks(T, TAU, nb, ib, ii);
      }
      break;
    }
    case HQ_GEQRT_U: {
      double *A = tile_ptr(G, t.k, t.k);
      const int ii = t.j * ib, kb = min(ib, nb - ii), live = live_cols(G, t.k);
      const int c0 = max(t.slab * HQ_SLAB, ii + kb), c1 = min(live, (t.slab + 1) * HQ_SLAB);
      if (ii < live && c0 < c1)
        apply_block<false>(S, A, t_ptr(G, t.k, t.k) + (size_t)ii * ib, ib, A, nullptr, nb, ii, kb, c0, c1);
      break;
    }
    case HQ_TSQRT_U: {
      double *A1 = tile_ptr(G, t.k, t.k), *A2 = tile_ptr(G, t.r, t.k);
      const int ii = t.j * ib, kb = min(ib, nb - ii), live = live_cols(G, t.k);
      const int c0 = max(t.slab * HQ_SLAB, ii + kb), c1 = min(live, (t.slab + 1) * HQ_SLAB);
      if (ii < live && c0 < c1)
        apply_block<true>(S, A2, t_ptr(G, t.r, t.k) + (size_t)ii * ib, ib, A2, A1, nb, ii, kb, c0, c1);
      break;
    }
    case HQ_LARFB: {
      const double *V = tile_ptr(G, t.k, t.k), *T = t_ptr(G, t.k, t.k);
      double *X = xtile_ptr(G, t.k, t.j);
      const int uw = hq_upd_slab(nb, 0), c0 = t.slab * uw, c1 = min(live_cols(G, t.j), c0 + uw);
      if (c0 < c1)
        for (int ii = 0; ii < nb; ii += ib)
          apply_block<false>(S, V, T + (size_t)ii * ib, ib, X, nullptr, nb, ii, min(ib, nb - ii), c0, c1);
      break;
    }
    case HQ_SSRFB: {
      const double *V = tile_ptr(G, t.r, t.k), *T = t_ptr(G, t.r, t.k);
      double *X1 = xtile_ptr(G, t.k, t.j), *X2 = xtile_ptr(G, t.r, t.j);
      const int uw = hq_upd_slab(nb, 0), c0 = t.slab * uw, c1 = min(live_cols(G, t.j), c0 + uw);
      if (c0 < c1)
        for (int ii = 0; ii < nb; ii += ib)
          apply_block<true>(S, V, T + (size_t)ii * ib, ib, X2, X1, nb, ii, min(ib, nb - ii), c0, c1);
      break;
    }
    default:
      break;
  }
}

// Persistent DAG scheduler: the claim / wait / publish protocol of the
// specialised scheduler kernel (tasks claimed in order from one atomic
// counter, dependency flags equal to this run's epoch, release before
// publishing, acquire after the wait, a bounded wait that reports the task in
// counter[1] and drains the grid), without its look-ahead claim.
__global__ __launch_bounds__(NT) void sched_kernel(hq_geom G0, const hq_geom *__restrict__ geoms,
                                                   const hq_task *__restrict__ tasks, int ntasks, int *flags,
                                                   int epoch_arg, int *counter, long long *trace,
                                                   const int *epoch_dev, unsigned wait_limit) {
  __shared__ Smem S;
  const int epoch = epoch_dev ? __hip_atomic_load(epoch_dev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : epoch_arg;
  int done = -1;
  long long t_claim = 0;
  for (;;) {
    // publish the previous task and claim the next one in one thread-0 region
    if (threadIdx.x == 0) {
      if (done >= 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_store(&flags[done], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (trace) trace[4 * done + 2] = wall_clock64();
      }
      int t = atomicAdd(counter, 1);
      if (trace) t_claim = wall_clock64();
      if (__hip_atomic_load(&counter[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) t = ntasks;
      if (t < ntasks) {
        const hq_task &tk = tasks[t];
        S.next = tk;
        const int nd = tk.ndep;
        unsigned spins = 0;
        for (;;) {
          int f[HQ_MAXDEP];
#pragma unroll
          for (int i = 0; i < HQ_MAXDEP; i++)
            f[i] = i < nd ? __hip_atomic_load(&flags[tk.dep[i]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                          : epoch;
          bool ready = true;
#pragma unroll
          for (int i = 0; i < HQ_MAXDEP; i++) ready = ready && f[i] == epoch;
          if (ready) break;
          __builtin_amdgcn_s_sleep(2);
          if (++spins > wait_limit) {
            __hip_atomic_store(&counter[1], 1 + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            t = ntasks;
            break;
          }
        }
        if (t < ntasks) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      S.task = t;
    }
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(S.task);
    if (t >= ntasks) break;
    hq_task tk;
    tk.kind = __builtin_amdgcn_readfirstlane(S.next.kind);
    tk.k = __builtin_amdgcn_readfirstlane(S.next.k);
    tk.r = __builtin_amdgcn_readfirstlane(S.next.r);
    tk.j = __builtin_amdgcn_readfirstlane(S.next.j);
    tk.slab = __builtin_amdgcn_readfirstlane(S.next.slab);
    tk.mat = __builtin_amdgcn_readfirstlane(S.next.mat);
    tk.ndep = 0;
    if (trace && threadIdx.x == 0) {
      trace[4 * t + 0] = t_claim;
      trace[4 * t + 1] = wall_clock64();
      trace[4 * t + 3] = blockIdx.x;
    }
    run_task(S, geoms ? geoms[tk.mat] : G0, tk);
    // every wave's stores have completed before the flag can be published
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    done = t;
  }
  if (trace && threadIdx.x == 0 && done >= 0) trace[4 * done + 2] = wall_clock64();
}

// One task per launch (stream order provides the dependencies).
__global__ __launch_bounds__(NT) void task_kernel(hq_geom G, hq_task tk) {
  __shared__ Smem S;
  run_task(S, G, tk);
}

// Standalone tile kernels on explicit tile pointers (ld = nb, T ld = ib <= 32).
__global__ __launch_bounds__(NT) void geqrt_kernel(double *A, double *T, double *TAU, int nb, int ib) {
  __shared__ Smem S;
  for (int ii = 0; ii < nb; ii += ib) {
    panel<false>(S, A, nullptr, T, TAU, nb, ib, ii);
    panel_rest<false>(S, A, nullptr, T, nb, ib, ii, nb);
  }
}
__global__ __launch_bounds__(NT) void tsqrt_kernel(double *A1, double *A2, double *T, double *TAU, int nb, int ib) {
  __shared__ Smem S;
  for (int ii = 0; ii < nb; ii += ib) {
    panel<true>(S, A2, A1, T, TAU, nb, ib, ii);
    panel_rest<true>(S, A2, A1, T, nb, ib, ii, nb);
  }
}
// grid.x = 64-column slab of C / of [A1; A2]
__global__ __launch_bounds__(NT) void larfb_kernel(const double *V, const double *T, double *C, int nb, int ib) {
  __shared__ Smem S;
  const int c0 = blockIdx.x * HQ_SLAB, c1 = min(nb, c0 + HQ_SLAB);
  for (int ii = 0; ii < nb; ii += ib)
    apply_block<false>(S, V, T + (size_t)ii * ib, ib, C, nullptr, nb, ii, min(ib, nb - ii), c0, c1);
}
__global__ __launch_bounds__(NT) void ssrfb_kernel(double *A1, double *A2, const double *V, const double *T, int nb,
                                                   int ib) {
  __shared__ Smem S;
  const int c0 = blockIdx.x * HQ_SLAB, c1 = min(nb, c0 + HQ_SLAB);
  for (int ii = 0; ii < nb; ii += ib)
    apply_block<true>(S, V, T + (size_t)ii * ib, ib, A2, A1, nb, ii, min(ib, nb - ii), c0, c1);
}

// T of one block of kb <= ib reflectors (any ib <= nb) rebuilt from the
// reflector tile V (ld nb; diag: unit lower with the diagonal at row i0 + c,
// else the V of a TSQRT tile under an identity) and tau: Gram entries one
// wave per pair into T's strict upper part, then dlarft column by column
// (column c of the Gram matrix read into LDS before the column is
// overwritten); rows below the diagonal written as zero.  drop_last: the
// last reflector of a diagonal tile (a 1-row column) counts with tau = 0.
__device__ void tbuild_body(const double *V, const double *tau, double *Tb, int nb, int ib, int i0, int kb,
                            bool diag, bool drop_last) {
  __shared__ double gcol[NBMAX];
  __shared__ double tv[NBMAX];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int c = tid; c < kb; c += NT) {
    double t = tau[c];
    if (drop_last && diag && i0 + c == nb - 1) t = 0.0;
    tv[c] = t;
  }
  const int npair = kb * (kb - 1) / 2;
  for (int e = wave; e < npair; e += NT / 64) {
    int c = 1;
    while ((c + 1) * c / 2 <= e) c++;
    const int p = e - c * (c - 1) / 2;
    const double *vp = V + (size_t)(i0 + p) * nb, *vc = V + (size_t)(i0 + c) * nb;
    double s = 0.0;
    for (int row = lane; row < nb; row += 64) {
      double a = vp[row], b = vc[row];
      if (diag) {
        const int dp = i0 + p, dc = i0 + c;
        a = row > dp ? a : (row == dp ? 1.0 : 0.0);
        b = row > dc ? b : (row == dc ? 1.0 : 0.0);
      }
      s += a * b;
    }
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) Tb[(size_t)c * ib + p] = s;
  }
  __syncthreads();
  for (int c = 0; c < kb; c++) {
    for (int q = tid; q < c; q += NT) gcol[q] = Tb[(size_t)c * ib + q];
    __syncthreads();
    const double tc = tv[c];
    for (int p = tid; p < ib; p += NT) {
      double val = 0.0;
      if (p < c) {
        double s = 0.0;
        for (int q = p; q < c; q++) s += Tb[(size_t)q * ib + p] * gcol[q];
        val = -tc * s;
      } else if (p == c) {
        val = tc;
      }
      Tb[(size_t)c * ib + p] = val;
    }
    __syncthreads();
  }
}

// Every reflector tile of a tile grid (grid.x = ib block, grid.y = entry of
// the lower tile list (r, k)).
__global__ __launch_bounds__(NT) void tbuild_kernel(hq_geom G, const int2 *tiles, int drop_last) {
  const int2 rk = tiles[blockIdx.y];
  const int r = rk.x, k = rk.y, ib = G.ib, i0 = blockIdx.x * ib;
  if (i0 >= G.nb) return;
  tbuild_body(tile_ptr(G, r, k), tau_ptr(G, r, k) + i0, t_ptr(G, r, k) + (size_t)i0 * ib, G.nb, ib, i0,
              min(ib, G.nb - i0), r == k, drop_last != 0);
}
__global__ __launch_bounds__(NT) void tbuild_tile_kernel(const double *V, const double *tau, double *T, int nb,
                                                         int ib, int diag) {
  const int i0 = blockIdx.x * ib;
  if (i0 >= nb) return;
  tbuild_body(V, tau + i0, T + (size_t)i0 * ib, nb, ib, i0, min(ib, nb - i0), diag != 0, false);
}

}  // namespace hqg

// ---------------------------------------------------------------------------
// Launch wrappers (called by the nb dispatch of the specialised launchers).
// ---------------------------------------------------------------------------
hipError_t hqg_launch_sched(const hq_geom &G, const hq_geom *geoms, const hq_task *tasks, int ntasks, int *flags,
                            int epoch, int *counter, int grid, long long *trace, hipStream_t st, int *epoch_dev,
                            unsigned wait_limit) {
  if (G.nb > hqg::NBMAX || G.ib > hqg::IBM || G.prec != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(hqg::sched_kernel, dim3(grid), dim3(hqg::NT), 0, st, G, geoms, tasks, ntasks, flags, epoch,
                     counter, trace, (const int *)epoch_dev, wait_limit);
  return hipGetLastError();
}
hipError_t hqg_launch_task(const hq_geom &G, const hq_task &tk, hipStream_t st) {
  if (G.nb > hqg::NBMAX || G.ib > hqg::IBM) return hipErrorInvalidValue;
  hipLaunchKernelGGL(hqg::task_kernel, dim3(1), dim3(hqg::NT), 0, st, G, tk);
  return hipGetLastError();
}
hipError_t hqg_launch_geqrt(int nb, int ib, double *A, double *T, double *TAU, hipStream_t st) {
  if (nb > hqg::NBMAX || ib > hqg::IBM) return hipErrorInvalidValue;
  hipLaunchKernelGGL(hqg::geqrt_kernel, dim3(1), dim3(hqg::NT), 0, st, A, T, TAU, nb, ib);
  return hipGetLastError();
}
hipError_t hqg_launch_tsqrt(int nb, int ib, double *A1, double *A2, double *T, double *TAU, hipStream_t st) {
  if (nb > hqg::NBMAX || ib > hqg::IBM) return hipErrorInvalidValue;
  hipLaunchKernelGGL(hqg::tsqrt_kernel, dim3(1), dim3(hqg::NT), 0, st, A1, A2, T, TAU, nb, ib);
  return hipGetLastError();
}
hipError_t hqg_launch_larfb(int nb, int ib, const double *V, const double *T, double *C, hipStream_t st) {
  if (nb > hqg::NBMAX || ib > hqg::IBM) return hipErrorInvalidValue;
  hipLaunchKernelGGL(hqg::larfb_kernel, dim3((nb + HQ_SLAB - 1) / HQ_SLAB), dim3(hqg::NT), 0, st, V, T, C, nb, ib);
  return hipGetLastError();
}
hipError_t hqg_launch_ssrfb(int nb, int ib, double *A1, double *A2, const double *V, const double *T,
                            hipStream_t st) {
  if (nb > hqg::NBMAX || ib > hqg::IBM) return hipErrorInvalidValue;
  hipLaunchKernelGGL(hqg::ssrfb_kernel, dim3((nb + HQ_SLAB - 1) / HQ_SLAB), dim3(hqg::NT), 0, st, A1, A2, V, T, nb,
                     ib);
  return hipGetLastError();
}
hipError_t hqg_launch_tbuild(const hq_geom &G, const int2 *tiles, int ntiles, int drop_last, hipStream_t st) {
  if (G.nb > hqg::NBMAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(hqg::tbuild_kernel, dim3((G.nb + G.ib - 1) / G.ib, ntiles), dim3(hqg::NT), 0, st, G, tiles,
                     drop_last);
  return hipGetLastError();
}
hipError_t hqg_launch_tbuild_tile(int nb, int ib, const double *V, const double *tau, double *T, int diag,
                                  hipStream_t st) {
  if (nb > hqg::NBMAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(hqg::tbuild_tile_kernel, dim3((nb + ib - 1) / ib), dim3(hqg::NT), 0, st, V, tau, T, nb, ib,
                     diag);
  return hipGetLastError();
}
