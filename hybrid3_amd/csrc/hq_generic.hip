// hq_generic.hip -- tile kernels and DAG scheduler for tile sizes without a
// register-resident specialisation: any nb up to 512 other than 64, 128 and
// 256 (e.g. nb = 384 or 512, or an nb that is not a multiple of 16).
//
// The task set, the DAG and the results are those of the specialised kernels
// (flat-tree GEQRT / TSQRT / LARFB / SSRFB, PLASMA's T layout with ld = ib,
// Hybrid_tile_QR.c:960-1363); only the data movement differs.  At these sizes
// the nb x 16 strip of a 64-column slab does not fit one wave's registers next
// to its operands (512 rows x 16 columns is 128 doubles per lane), and the
// specialised kernels' LDS layout does not fit either, so:
//   * a panel block (<= 32 columns of <= 512 rows) is factored in LDS: one
//     pass per column computes the column's squared norm and its dot products
//     with every column right of it at once (16 row groups per column), and one
//     pass applies the reflector;
//   * a block reflector is applied on the matrix cores with the slab held in
//     registers for the whole task, two waves per 16-column strip (one per
//     row half, 64 doubles per lane each) whose partial V^T X products are
//     exchanged in LDS once per 16-reflector sub-block (update_strips).
//     Workgroups of 8 waves, two per SIMD.
// FP64 everywhere (reduced-precision updates exist for nb = 128 and 256).
#include <hip/hip_runtime.h>

#include "hq_internal.h"

namespace hqg {

constexpr int NBMAX = 512;
constexpr int IBM = HQ_IBMAX;  // reflectors per block (the run's ib <= 32)
constexpr int LDP = NBMAX + 1; // panel column stride
constexpr int LDT = IBM + 1;   // T / Gram / triangle column stride
constexpr int NT = 512;        // threads per workgroup: 8 waves, two per SIMD
constexpr int NW = NT / 64;    // waves
constexpr int RPT = (NBMAX + NT - 1) / NT;  // panel rows per thread
constexpr int LDV2 = NBMAX + 2;  // column stride of a staged 16-reflector sub-block of V

struct Smem {
  union {
    double P[IBM * LDP];  // panel block: column c at P + c*LDP, row i = panel row i
    struct {
      alignas(32) double Vs[16 * LDV2];   // sub-block of V: V(r, c) at Vs[c*LDV2 + r]
      alignas(32) double Wx[NW * 256];   // partial W of each (row group, strip), MFMA layout
    } m;
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

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// v_mfma_f64_16x16x4f64 lane layout (gfx950): A(i,k) at lane i + 16k,
// B(k,j) at lane j + 16k, D(i,j) at lane j + 16*(i%4), register i/4.  An
// A-operand row i fed from matrix row arow(i) = 4*(i%4) + i/4 makes register
// v of lane group g hold matrix row 4g + v of the output, so a column strip
// in accumulator form is four consecutive rows per lane, and the same
// registers are the B operand of the next product with k index g.
__device__ __forceinline__ int arow(int i) { return 4 * (i & 3) + (i >> 2); }

// Rows of the strip a wave owns: one of R row groups of the tile, in whole
// 16-row blocks.
template <int R>
__device__ __forceinline__ int group_rows(int nb) { return (((nb + R - 1) / R) + 15) & ~15; }

// X <- H_b^T X for the blocks b = b0 .. b1-1 in order (H_b = I - V_b T_b
// V_b^T, V_b the reflectors at columns b*ib .. of tile Vt, T_b at
// T + b*ib*ib with ld ib), on X's columns c0 .. c1-1 (c1 - c0 <= 16 SN):
//   !TS: V_b unit lower trapezoidal (diagonal at row b*ib + c), X the tile;
//   TS:  [X1 rows of the block; X] with V_b = [I; Vt columns].
// Register-resident, on the matrix cores: wave w owns the 16-column strip
// w % SN of the slab over row group w / SN (of NW / SN groups), as MFMA
// accumulators for the whole call.  Each block is applied as sub-blocks of
// 16 reflectors with the 16x16 diagonal blocks of T_b (the compact WY form
// of a product of two blocks), V of a sub-block staged in LDS (the next
// one's loads in flight in registers meanwhile):
//   W = V^T X: each wave its row group; the groups' partials exchanged in
//     LDS and summed in group order (every wave of a strip holds the same W);
//   W2 = T^T W (+ X1 rows first, TS; then X1 -= W2 by row group 0);
//   X -= V W2 on the wave's row blocks.
// SN = 4 (64-column slabs, two row halves) for the factorisation; SN = 1
// (one 16-column strip over 8 row groups) for a narrow X, whose chain of
// tasks along the flat tree then runs on a fraction of the work each.
// Its pointers are cast to the global address space (the compiler keeps
// this function out of line, where they would be generic: a flat access also
// counts against the LDS counter, so every wait for an LDS read would wait
// for the V loads in flight as well).
typedef __attribute__((address_space(1))) double gdouble;
template <bool TS, int SN = 4>
__device__ void update_strips(Smem &S, const double *Vt_, const double *T_, int ib, double *X_, double *X1_, int nb,
                              int b0, int b1, int c0, int c1) {
  const gdouble *Vt = (const gdouble *)Vt_, *T = (const gdouble *)T_;
  gdouble *X = (gdouble *)X_, *X1 = (gdouble *)X1_;
  constexpr int RG = NW / SN;              // row groups
  constexpr int RBM = NBMAX / 16 / RG;     // 16-row blocks of a group at most
  const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int s = wave % SN, h = wave / SN;
  const int RH = group_rows<RG>(nb), rbase = h * RH;
  const int nrb = max(0, min(RH, nb - rbase) + 15) >> 4;  // 16-row blocks of this group
  const int col = c0 + 16 * s + m;
  const bool colok = col < c1;
  // !TS: rows above the first block's diagonal are neither read nor written
  // (V is zero there; they hold R rows of earlier blocks, which tasks of the
  // same step may be updating meanwhile)
  const int row_lo = TS ? 0 : b0 * ib;
  // the strip, rows rbase + 16 rb + 4g + v at acc[rb][v]
  d4 acc[RBM];
#pragma unroll
  for (int rb = 0; rb < RBM; rb++) {
    acc[rb] = d4{0.0, 0.0, 0.0, 0.0};
    if (rb < nrb && colok) {
      const int r = rbase + 16 * rb + 4 * g;
      const gdouble *xp = X + (size_t)col * nb + r;
      if (r >= row_lo && r + 3 < nb) {
        acc[rb] = *reinterpret_cast<const __attribute__((address_space(1))) d4 *>(xp);
      } else {
#pragma unroll
        for (int v = 0; v < 4; v++)
          if (r + v >= row_lo && r + v < nb) acc[rb][v] = xp[v];
      }
    }
  }
  // sub-blocks: (block b, first reflector q0 of the block) in order
  const int nsb_per = (ib + 15) / 16;
  const int nsb = (b1 - b0) * nsb_per;
  // V of a sub-block: 16 columns of LDV2 rows (zero past nb), NT threads x
  // VPT doubles
  constexpr int VPT = (16 * LDV2 + NT - 1) / NT;
  double vpre[VPT];
  auto load_v = [&](int sbi) {
    const int b = b0 + sbi / nsb_per, q0 = (sbi % nsb_per) * 16;
    const int ii = b * ib, kb = min(ib, nb - ii);
#pragma unroll
    for (int k = 0; k < VPT; k++) {
      const int e = tid + k * NT, c = e / LDV2, r = e - c * LDV2;
      const int q = q0 + c, d = ii + q;  // reflector index in the block, its diagonal row
      double val = 0.0;
      if (e < 16 * LDV2 && q < kb && r < nb) {
        if (TS || r > d)
          val = Vt[(size_t)d * nb + r];
        else
          val = r == d ? 1.0 : 0.0;
      }
      vpre[k] = val;
    }
  };
  if (nsb > 0) load_v(0);
  for (int sbi = 0; sbi < nsb; sbi++) {
    const int b = b0 + sbi / nsb_per, q0 = (sbi % nsb_per) * 16;
    const int ii = b * ib, kb = min(ib, nb - ii);
    const int nq = min(16, kb - q0);  // reflectors of this sub-block
    if (nq <= 0) {                    // past the last (narrower) block's end; uniform
      if (sbi + 1 < nsb) load_v(sbi + 1);
      continue;
    }
    __syncthreads();                  // the previous sub-block's V and W are read
#pragma unroll
    for (int k = 0; k < VPT; k++)
      if (tid + k * NT < 16 * LDV2) S.m.Vs[tid + k * NT] = vpre[k];
    for (int e = tid; e < 16 * 16; e += NT) {
      const int c = e >> 4, p = e & 15;
      S.TT[c * LDT + p] = (c < nq && p <= c) ? T[(size_t)ii * ib + (size_t)(q0 + c) * ib + q0 + p] : 0.0;
    }
    __syncthreads();
    if (sbi + 1 < nsb) load_v(sbi + 1);  // in flight during the products below
    const double *Vs = S.m.Vs;
    // W partial over this half's rows: A(i, k) = V(row_k, i), B(k, j) = X(row_k, j)
    d4 wp = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      if (rb < nrb) {
#pragma unroll
        for (int v = 0; v < 4; v++) wp = mfma(Vs[m * LDV2 + rbase + 16 * rb + 4 * g + v], acc[rb][v], wp);
      }
    }
    // TS: the block's rows of X1 (read before the exchange barrier; the upper
    // half writes them after it)
    d4 x1 = d4{0.0, 0.0, 0.0, 0.0};
    if (TS && colok) {
#pragma unroll
      for (int v = 0; v < 4; v++)
        if (4 * v + g < nq) x1[v] = X1[(size_t)col * nb + ii + q0 + 4 * v + g];
    }
    *reinterpret_cast<d4 *>(S.m.Wx + (h * SN + s) * 256 + 4 * lane) = wp;
    __syncthreads();
    d4 w;
    if constexpr (RG == 2) {
      // own + other: the same sum in both waves (addition commutes)
      w = wp + *reinterpret_cast<const d4 *>(S.m.Wx + ((1 - h) * SN + s) * 256 + 4 * lane);
    } else {
      w = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < RG; q++) w += *reinterpret_cast<const d4 *>(S.m.Wx + (q * SN + s) * 256 + 4 * lane);
    }
    if (TS) w += x1;
    // W2 = T^T W: A(i, k) = T(k, i), B = W (D layout = B layout per k step)
    d4 w2 = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 4; kk++) w2 = mfma(S.TT[m * LDT + 4 * kk + g], w[kk], w2);
    if (TS && h == 0 && colok) {
#pragma unroll
      for (int v = 0; v < 4; v++)
        if (4 * v + g < nq) X1[(size_t)col * nb + ii + q0 + 4 * v + g] = x1[v] - w2[v];
    }
    const d4 nw2 = -w2;
    // X -= V W2 on this group's row blocks: A(i, k) = V(rbase + 16 rb + arow(i), k)
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      if (rb < nrb) {
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
          acc[rb] = mfma(Vs[(4 * kk + g) * LDV2 + rbase + 16 * rb + arow(m)], nw2[kk], acc[rb]);
      }
    }
  }
  // store the strip
#pragma unroll
  for (int rb = 0; rb < RBM; rb++) {
    if (rb < nrb && colok) {
      const int r = rbase + 16 * rb + 4 * g;
      gdouble *xp = X + (size_t)col * nb + r;
      if (r >= row_lo && r + 3 < nb) {
        *reinterpret_cast<__attribute__((address_space(1))) d4 *>(xp) = acc[rb];
      } else {
#pragma unroll
        for (int v = 0; v < 4; v++)
          if (r + v >= row_lo && r + v < nb) xp[v] = acc[rb][v];
      }
    }
  }
  __syncthreads();
}

// X <- H^T X for all blocks of tile Vt (as update_strips with b0 = 0, b1 =
// the block count) on one 16-column strip of X (c1 - c0 <= 16): wave w owns
// row group w of NW.  No two waves need the same rows of V, so each wave
// reads its own rows of each 16-reflector sub-block (and the sub-block's
// 16 x 16 block of T) with whole-line loads -- lane l: reflector l / 4, four
// consecutive rows -- PD sub-blocks ahead in registers, and rearranges them
// into the MFMA operand layouts through an LDS area of its own (no barrier:
// a wave's LDS accesses complete in order).  The groups' partial W are summed
// in group order through LDS, double-buffered, one barrier per sub-block.
// RB: 16-row blocks a group holds at most.  Out of line, so that the
// apply-only kernel allocates registers for each of its two bodies on its
// own; its pointers are then generic, and are cast to the global address
// space (flat loads would also count against the LDS counter, so every wait
// for an LDS read would wait for the loads in flight as well).
constexpr int NLD = 17;  // column stride of a wave's LDS copy of V / T (odd: fewer bank conflicts)
// Sub-block pipelining along a chain of SSRFBs on one strip: sub-block s of
// SSRFB(k, r) reads and writes only rows 16s .. 16s+15 of X1, so it can run
// once SSRFB(k, r-1) has written those rows.  prog: this task's progress
// word (published after each sub-block once its X1 rows are stored), wait:
// the predecessor's (or null), both (epoch << 32) | sub-blocks done.  The X1
// rows pass between the two tasks as agent-scope atomic stores and loads, so
// the word needs neither an L2 write-back nor an invalidation.  A wait is
// bounded by wait_limit sleeps; past it the task index goes to err (the
// scheduler's time-out word) and the task runs on.
struct Pipe {
  unsigned long long *prog;
  const unsigned long long *wait;
  unsigned long long epoch;
  unsigned wait_limit;
  int *err;
  int self;
};

template <bool TS, int RB, int PD, int NS = 1>
__device__ __noinline__ void update_narrow(Smem &S, const double *Vt_, const double *T_, int ib, double *X_,
                                           double *X1_, int nb, int c0, int c1, Pipe pp) {
  const gdouble *Vt = (const gdouble *)Vt_, *T = (const gdouble *)T_;
  gdouble *X = (gdouble *)X_, *X1 = (gdouble *)X1_;
  const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4;
  const int h = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int RH = group_rows<NW>(nb), rbase = h * RH;
  const int nrb = max(0, min(RH, nb - rbase) + 15) >> 4;
  // NS strips of 16 columns: strip q is column c0 + 16 q + m of this lane
  int col[NS];
  bool colok[NS];
#pragma unroll
  for (int q = 0; q < NS; q++) {
    col[q] = c0 + 16 * q + m;
    colok[q] = col[q] < c1;
  }
  // LDS: partial W (2 x NW x NS x 256), then per wave RB blocks of V (16 x
  // NLD each) and the T block (16 x NLD)
  static_assert(2 * NW * NS * 256 + NW * (RB + 1) * 16 * NLD <= sizeof(S.P) / sizeof(double), "LDS areas");
  double *Wx = S.P;
  double *Vw = S.P + 2 * NW * NS * 256 + h * (RB + 1) * 16 * NLD, *Tw = Vw + RB * 16 * NLD;
  // the strips, rows rbase + 16 rb + 4g + v at acc[q][rb][v]
  d4 acc[NS][RB];
#pragma unroll
  for (int q = 0; q < NS; q++)
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
      acc[q][rb] = d4{0.0, 0.0, 0.0, 0.0};
      if (rb < nrb && colok[q]) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = rbase + 16 * rb + 4 * g + v;
          if (r < nb) acc[q][rb][v] = X[(size_t)col[q] * nb + r];
        }
      }
    }
  const int nsb_per = (ib + 15) / 16, nsb = ((nb + ib - 1) / ib) * nsb_per;
  // loaded layout: lane l holds reflector lc = l / 4, rows 4 (l % 4) + e
  const int lc = lane >> 2, lr = 4 * (lane & 3);
  double vr[PD][RB][4], tr[PD][4];
  auto load = [&](int p, int sbi) {
    const bool live = sbi < nsb;
    const int b = sbi / nsb_per, q0 = (sbi % nsb_per) * 16, ii = b * ib, kb = min(ib, nb - ii);
    const int q = q0 + lc, d = ii + q;  // reflector of this lane, its diagonal row
    const bool qok = live && q < kb;
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
      const int r0 = rbase + 16 * rb + lr;
      const bool whole = qok && rb < nrb && r0 + 3 < nb && (nb & 3) == 0;
      d4 x = d4{0.0, 0.0, 0.0, 0.0};
      if (whole) {
        x = *reinterpret_cast<const __attribute__((address_space(1))) d4 *>(Vt + (size_t)d * nb + r0);
      } else if (qok && rb < nrb) {
#pragma unroll
        for (int e = 0; e < 4; e++)
          if (r0 + e < nb) x[e] = Vt[(size_t)d * nb + r0 + e];
      }
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int r = r0 + e;
        // !TS: unit lower trapezoidal (R above the diagonal is not V)
        vr[p][rb][e] = TS || r > d ? x[e] : (r == d && qok ? 1.0 : 0.0);
      }
    }
    // T(q0 + lr + e, q0 + lc): upper triangle of the sub-block's diagonal block
    const int nq = min(16, kb - q0);
    d4 t = d4{0.0, 0.0, 0.0, 0.0};
    if (live && lc < nq) {
      const gdouble *tp = T + (size_t)ii * ib + (size_t)(q0 + lc) * ib + q0 + lr;
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (lr + e <= lc) t[e] = tp[e];
    }
#pragma unroll
    for (int e = 0; e < 4; e++) tr[p][e] = t[e];
  };
  auto apply = [&](int p, int sbi) {
    const int b = sbi / nsb_per, q0 = (sbi % nsb_per) * 16, ii = b * ib, kb = min(ib, nb - ii);
    const int nq = min(16, kb - q0);
    if (nq <= 0) return;  // past the last (narrower) block's end; uniform
    // this wave's copy: V(row 16 rb + r, reflector c) at Vw[rb * 16 * NLD + c * NLD + r]
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
      for (int e = 0; e < 4; e++) Vw[rb * 16 * NLD + lc * NLD + lr + e] = vr[p][rb][e];
#pragma unroll
    for (int e = 0; e < 4; e++) Tw[lc * NLD + lr + e] = tr[p][e];
    // W partials over this group's rows: A(i, k) = V(row 4k + v, i) (read
    // once for all strips), B = the strip
    d4 wp[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) wp[q] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
      if (rb < nrb) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const double a = Vw[rb * 16 * NLD + m * NLD + 4 * g + v];
#pragma unroll
          for (int q = 0; q < NS; q++) wp[q] = mfma(a, acc[q][rb][v], wp[q]);
        }
      }
    double *Wb = Wx + (sbi & 1) * (NW * NS * 256);
#pragma unroll
    for (int q = 0; q < NS; q++) *reinterpret_cast<d4 *>(Wb + (h * NS + q) * 256 + 4 * lane) = wp[q];
    // TS, pipelined: the predecessor's rows of this sub-block written
    if (TS && pp.wait && tid == 0) {
      const unsigned long long need = (pp.epoch << 32) | (unsigned long long)(sbi + 1);
      unsigned spins = 0;
      while (__hip_atomic_load(pp.wait, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > pp.wait_limit || __hip_atomic_load(pp.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          __hip_atomic_store(pp.err, 1 + pp.self, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
    __syncthreads();
    // TS: the sub-block's rows of X1 (group 0 writes them below).  With a
    // pipelined chain they are read and written as agent-scope atomics
    // (coherent across the XCDs' L2s without write-back or invalidation),
    // which is all that passes between the two tasks
    d4 x1[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) {
      x1[q] = d4{0.0, 0.0, 0.0, 0.0};
      if (TS && colok[q]) {
#pragma unroll
        for (int v = 0; v < 4; v++)
          if (4 * v + g < nq) {
            double *a = (double *)&X1[(size_t)col[q] * nb + ii + q0 + 4 * v + g];
            x1[q][v] = pp.prog ? __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *a;
          }
      }
    }
    d4 w[NS], w2[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) {
      w[q] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int qg = 0; qg < NW; qg++) w[q] += *reinterpret_cast<const d4 *>(Wb + (qg * NS + q) * 256 + 4 * lane);
      if (TS) w[q] += x1[q];
      w2[q] = d4{0.0, 0.0, 0.0, 0.0};
    }
    // W2 = T^T W: A(i, k) = T(k, i)
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const double t = Tw[m * NLD + 4 * kk + g];
#pragma unroll
      for (int q = 0; q < NS; q++) w2[q] = mfma(t, w[q][kk], w2[q]);
    }
    if (TS && h == 0) {
#pragma unroll
      for (int q = 0; q < NS; q++)
        if (colok[q]) {
#pragma unroll
          for (int v = 0; v < 4; v++)
            if (4 * v + g < nq) {
              double *a = (double *)&X1[(size_t)col[q] * nb + ii + q0 + 4 * v + g];
              if (pp.prog)
                __hip_atomic_store(a, x1[q][v] - w2[q][v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              else
                *a = x1[q][v] - w2[q][v];
            }
        }
    }
    if (TS && pp.prog && h == 0) {
      // this sub-block's X1 rows (atomic stores) complete before the word
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_waitcnt(0);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      if (lane == 0)
        __hip_atomic_store(pp.prog, (pp.epoch << 32) | (unsigned long long)(sbi + 1), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    // X -= V W2: A(i, k) = V(row arow(i) of block rb, reflector k), read once
    // for all strips
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
      if (rb < nrb) {
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
          const double a = Vw[rb * 16 * NLD + (4 * kk + g) * NLD + arow(m)];
#pragma unroll
          for (int q = 0; q < NS; q++) acc[q][rb] = mfma(a, -w2[q][kk], acc[q][rb]);
        }
      }
  };
#pragma unroll
  for (int p = 0; p < PD; p++) load(p, p);
  for (int s0 = 0; s0 < nsb; s0 += PD) {
#pragma unroll
    for (int p = 0; p < PD; p++) {
      if (s0 + p < nsb) {
        apply(p, s0 + p);
        load(p, s0 + p + PD);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NS; q++)
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
      if (rb < nrb && colok[q]) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = rbase + 16 * rb + 4 * g + v;
          if (r < nb) X[(size_t)col[q] * nb + r] = acc[q][rb][v];
        }
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
    update_strips<TS>(S, A, T, ib, A, A1, nb, ii / ib, ii / ib + 1, c0, min(live, c0 + HQ_SLAB));
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
// NRB > 0: an application of Q^T only (LARFB / SSRFB on 16-column slabs,
// hq_geom::uw; update_narrow with NRB row blocks per group), a kernel of its
// own.
template <int NRB = 0, int NS = 1>
__device__ void run_task(Smem &S, const hq_geom &G, const hq_task &t, int self = -1, int epoch = 0,
                         unsigned wait_limit = 0, int *err = nullptr) {
  constexpr bool NARROW = NRB > 0;
  constexpr int PD = NS > 1 ? 3 : NRB <= 2 ? 4 : 3;  // sub-blocks in flight (registers)
  const int nb = G.nb, ib = G.ib;
  if constexpr (NARROW)
    if (t.kind != HQ_LARFB && t.kind != HQ_SSRFB) return;
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
        update_strips<false>(S, A, t_ptr(G, t.k, t.k), ib, A, nullptr, nb, t.j, t.j + 1, c0, c1);
      break;
    }
    case HQ_TSQRT_U: {
      double *A1 = tile_ptr(G, t.k, t.k), *A2 = tile_ptr(G, t.r, t.k);
      const int ii = t.j * ib, kb = min(ib, nb - ii), live = live_cols(G, t.k);
      const int c0 = max(t.slab * HQ_SLAB, ii + kb), c1 = min(live, (t.slab + 1) * HQ_SLAB);
      if (ii < live && c0 < c1)
        update_strips<true>(S, A2, t_ptr(G, t.r, t.k), ib, A2, A1, nb, t.j, t.j + 1, c0, c1);
      break;
    }
    case HQ_LARFB: {
      const double *V = tile_ptr(G, t.k, t.k), *T = t_ptr(G, t.k, t.k);
      double *X = xtile_ptr(G, t.k, t.j);
      const int uw = NARROW ? 16 * NS : hq_upd_slab(nb, 0), c0 = t.slab * uw, c1 = min(live_cols(G, t.j), c0 + uw);
      if (c0 >= c1) break;
      if constexpr (NARROW)
        update_narrow<false, (NARROW ? NRB : 1), PD, NS>(S, V, T, ib, X, nullptr, nb, c0, c1, Pipe{});
      else
        update_strips<false>(S, V, T, ib, X, nullptr, nb, 0, (nb + ib - 1) / ib, c0, c1);
      break;
    }
    case HQ_SSRFB: {
      const double *V = tile_ptr(G, t.r, t.k), *T = t_ptr(G, t.r, t.k);
      double *X1 = xtile_ptr(G, t.k, t.j), *X2 = xtile_ptr(G, t.r, t.j);
      const int uw = NARROW ? 16 * NS : hq_upd_slab(nb, 0), c0 = t.slab * uw, c1 = min(live_cols(G, t.j), c0 + uw);
      if (c0 >= c1) break;
      if constexpr (NARROW) {
        // dep[0] = -2 - predecessor: the previous SSRFB of the chain, waited
        // for sub-block by sub-block
        Pipe pp{};
        if (G.prog && self >= 0) {
          pp.prog = G.prog + self;
          pp.wait = t.ndep > 0 && t.dep[0] < 0 ? G.prog + (-2 - t.dep[0]) : nullptr;
          pp.epoch = (unsigned long long)(unsigned)epoch;
          pp.wait_limit = wait_limit;
          pp.err = err;
          pp.self = self;
        }
        update_narrow<true, (NARROW ? NRB : 1), PD, NS>(S, V, T, ib, X2, X1, nb, c0, c1, pp);
      } else
        update_strips<true>(S, V, T, ib, X2, X1, nb, 0, (nb + ib - 1) / ib, c0, c1);
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
template <int NRB, int NS = 1>
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
          for (int i = 0; i < HQ_MAXDEP; i++)  // (a negative dep is waited for inside the task)
            f[i] = i < nd && tk.dep[i] >= 0
                       ? __hip_atomic_load(&flags[tk.dep[i]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
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
    tk.ndep = __builtin_amdgcn_readfirstlane(S.next.ndep);
    tk.dep[0] = __builtin_amdgcn_readfirstlane(S.next.dep[0]);
    if (trace && threadIdx.x == 0) {
      trace[4 * t + 0] = t_claim;
      trace[4 * t + 1] = wall_clock64();
      trace[4 * t + 3] = blockIdx.x;
    }
    run_task<NRB, NS>(S, geoms ? geoms[tk.mat] : G0, tk, t, epoch, wait_limit, counter + 1);
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
  update_strips<false>(S, V, T, ib, C, nullptr, nb, 0, (nb + ib - 1) / ib, c0, c1);
}
__global__ __launch_bounds__(NT) void ssrfb_kernel(double *A1, double *A2, const double *V, const double *T, int nb,
                                                   int ib) {
  __shared__ Smem S;
  const int c0 = blockIdx.x * HQ_SLAB, c1 = min(nb, c0 + HQ_SLAB);
  update_strips<true>(S, V, T, ib, A2, A1, nb, 0, (nb + ib - 1) / ib, c0, c1);
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
  if (G.nb > hqg::NBMAX || G.ib > hqg::IBM || G.prec != 0 || (G.uw && G.uw != 16 && G.uw != 32) ||
      (G.uw == 32 && G.nb > 256))
    return hipErrorInvalidValue;
  // 16- / 32-column apply: row groups of NB / 8 rows, in 16-row blocks
  if (G.uw == 32)
    hipLaunchKernelGGL((hqg::sched_kernel<2, 2>), dim3(grid), dim3(hqg::NT), 0, st, G, geoms, tasks, ntasks, flags,
                       epoch, counter, trace, (const int *)epoch_dev, wait_limit);
  else if (G.uw && G.nb <= 256)
    hipLaunchKernelGGL(hqg::sched_kernel<2>, dim3(grid), dim3(hqg::NT), 0, st, G, geoms, tasks, ntasks, flags,
                       epoch, counter, trace, (const int *)epoch_dev, wait_limit);
  else if (G.uw)
    hipLaunchKernelGGL(hqg::sched_kernel<4>, dim3(grid), dim3(hqg::NT), 0, st, G, geoms, tasks, ntasks, flags,
                       epoch, counter, trace, (const int *)epoch_dev, wait_limit);
  else
    hipLaunchKernelGGL(hqg::sched_kernel<0>, dim3(grid), dim3(hqg::NT), 0, st, G, geoms, tasks, ntasks, flags,
                       epoch, counter, trace, (const int *)epoch_dev, wait_limit);
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
