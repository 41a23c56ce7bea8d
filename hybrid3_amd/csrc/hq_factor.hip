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

// Phase markers for timing the panel tasks (compile to nothing unless a
// build defines HQ_PHASE as a function-like macro).
#ifndef HQ_PHASE
#define HQ_PHASE(i)
#endif

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

// Wave-wide sums on the VALU with DPP lane moves (no LDS round trips):
// xor-1 and xor-2 quad permutes, half-row mirror and row mirror leave every
// lane of a 16-lane row holding the row sum (each step adds the same two
// operands in both partner lanes, so partners agree bit for bit); the four
// row sums are then read back with readlane and added in the same order by
// every lane.  N independent sums are interleaved to overlap latencies.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
  int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, true);
  int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_d(double x, int l) {
  int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
  int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
  return __hiloint2double(hi, lo);
}
// x + (x of the lane 16 apart: rows 0<->1, 2<->3), and + (lane 32 apart);
// v_permlane{16,32}_swap exchange rows of two registers, so swapping a value
// with itself and adding the pair gives the cross-row sum in every lane.
__device__ __forceinline__ double xrow_sum16(double x) {
  auto l = __builtin_amdgcn_permlane16_swap(__double2loint(x), __double2loint(x), false, false);
  auto h = __builtin_amdgcn_permlane16_swap(__double2hiint(x), __double2hiint(x), false, false);
  return __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
}
__device__ __forceinline__ double xrow_sum32(double x) {
  auto l = __builtin_amdgcn_permlane32_swap(__double2loint(x), __double2loint(x), false, false);
  auto h = __builtin_amdgcn_permlane32_swap(__double2hiint(x), __double2hiint(x), false, false);
  return __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
}
// One butterfly stage that halves M per-lane values: the two lanes of a pair
// (DPP pattern CTRL) keep opposite halves, so each sends the half it drops.
template <int M, int CTRL>
__device__ __forceinline__ void fold(double *u, bool hi_half) {
#pragma unroll
  for (int k = 0; k < M / 2; k++) {
    const double keep = hi_half ? u[k + M / 2] : u[k];
    const double send = hi_half ? u[k] : u[k + M / 2];
    u[k] = keep + dpp_move<CTRL>(send);
  }
}
// Wave-wide sums of N per-lane values (N = 1, 2, 4, 8), result v[t] uniform.
// Stages pair lanes i^1 (quad_perm [1,0,3,2]), i^2 (quad_perm [2,3,0,1]),
// i^7 (row_half_mirror), i^15 (row_mirror), then rows.  While more than one
// value is left, stage s keeps the half picked by f1 = b0^b2, f2 = b1^b2,
// f3 = b2^b3 of the lane bits: each is flipped by its own pairing and left
// unchanged by the later ones, so paired lanes always hold the same columns.
// After the stages lane L holds the total of value 4*f1+2*f2+f3 (N = 8),
// 2*f1+f2 (N = 4), f1 (N = 2); pairs of lanes add the same two numbers, so
// every copy of a total is bitwise identical.
template <int N>
__device__ __forceinline__ void wave_sums(double (&v)[N]) {
  static_assert(N == 1 || N == 2 || N == 4 || N == 8, "wave_sums: N in {1,2,4,8}");
  const int lane = threadIdx.x & 63;
  const bool f1 = (lane ^ (lane >> 2)) & 1;
  const bool f2 = ((lane >> 1) ^ (lane >> 2)) & 1;
  const bool f3 = ((lane >> 2) ^ (lane >> 3)) & 1;
  double u[N];
#pragma unroll
  for (int i = 0; i < N; i++) u[i] = v[i];
  if constexpr (N >= 2) fold<N, 0xB1>(u, f1); else u[0] += dpp_move<0xB1>(u[0]);
  if constexpr (N >= 4) fold<N / 2, 0x4E>(u, f2); else u[0] += dpp_move<0x4E>(u[0]);
  if constexpr (N >= 8) fold<N / 4, 0x141>(u, f3); else u[0] += dpp_move<0x141>(u[0]);
  double y = u[0] + dpp_move<0x140>(u[0]);
  y = xrow_sum32(xrow_sum16(y));
  if constexpr (N == 1) {
    v[0] = y;
  } else {
#pragma unroll
    for (int t = 0; t < N; t++) {
      // a lane (row 0, b3 = 0) holding value t
      int g1, g2, g3;
      if constexpr (N == 8) { g1 = t >> 2; g2 = (t >> 1) & 1; g3 = t & 1; }
      if constexpr (N == 4) { g1 = t >> 1; g2 = t & 1; g3 = 0; }
      if constexpr (N == 2) { g1 = t; g2 = 0; g3 = 0; }
      const int b2 = g3, b0 = g1 ^ b2, b1 = g2 ^ b2;
      v[t] = readlane_d(y, b0 + 2 * b1 + 4 * b2);
    }
  }
}
__device__ __forceinline__ double wave_sum1(double x) {
  double v[1] = {x};
  wave_sums<1>(v);
  return v[0];
}

template <int NB>
struct Smem {
  // column stride NB+2: each column starts 16-byte aligned (direct-to-LDS loads
  // write 64 lanes x 16 bytes contiguously) and consecutive columns fall on
  // different LDS banks (2*c mod 32 distinct for the 16 columns of an MFMA)
  static constexpr int LDV = NB + 2;
  static constexpr int LDT = HQ_IBMAX + 1;
  static constexpr int VSZ = HQ_IBMAX * LDV;
  static constexpr int TSZ = HQ_IBMAX * LDT;
  // buffer 0: V block / panel, column-major V[c*LDV + r]; T^T, TT[c*LDT + p] = T(p,c)
  double V[VSZ];
  double TT[TSZ];
  union {
    struct {  // buffer 1 of the LARFB / SSRFB pipeline
      double V1[VSZ];
      double TT1[TSZ];
    };
    struct {  // panel tasks only
      double A1[TSZ];  // TSQRT triangle block, A1[c*LDT + p] = R(p,c)
      double G[TSZ];   // Gram V^T V, G[p*LDT + c]
    };
  };
  double XA[16 * HQ_SLAB];  // staging of the pipelined split update (16 rows of a slab)
  double tau[HQ_IBMAX];
  int task;
  int nref;     // reflectors published in the current panel factorisation
  hq_task next; // descriptor of the task the workgroup runs next (scheduler)
};

// Global-address-space views of tile pointers.  Pointers read from the
// batch's geometry table are generic to the compiler, which then emits flat
// accesses; flat loads also count in the LDS counter, so every LDS operand
// wait would wait for them too.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}

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
      acc[rb] = *gmem(reinterpret_cast<const d4 *>(tile + (size_t)col * NB + 16 * rb + 4 * g));
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
    if (rb >= rb0) *gmem(reinterpret_cast<d4 *>(tile + (size_t)col * NB + 16 * rb + 4 * g)) = acc[rb];
}

// --------------------------------------------------------------------------
// Apply one block reflector (V, T from LDS) from the left, transposed:
//   W  = [A1 rows] + V^T X        (HAS_A1: SSRFB / TSQRT trailing; else LARFB)
//   W  = T^T W
//   A1 -= W ;  X -= V W
// X is the strip held in `acc` (rows >= 16*rb0 take part).  WB = IBP/16.
// --------------------------------------------------------------------------
// UNIT: Vs holds the raw tile (R above the diagonal): the operand is masked
// to the unit lower trapezoid with its diagonal at row i0 + c.  Columns
// c >= kb of Vs are treated as zero.
template <int WB>
__device__ __forceinline__ void load_a1(d4 (&A1v)[WB], const double *a1col, int kb, bool ok) {
  const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int wb = 0; wb < WB; wb++) {
    A1v[wb] = d4{0.0, 0.0, 0.0, 0.0};
    if (ok) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int row = 16 * wb + 4 * g + v;
        if (row < kb) A1v[wb][v] = a1col[row];
      }
    }
  }
}

template <int NB, int WB, bool HAS_A1, bool UNIT = false>
__device__ __forceinline__ void apply_block(d4 (&acc)[NB / 16], int rb0, const double *Vs,
                                            const double *TTs, const d4 (&A1v)[WB], double *a1col,
                                            int kb, bool ok, int i0 = 0) {
  constexpr int LDV = Smem<NB>::LDV, LDT = Smem<NB>::LDT;
  constexpr int NBR = NB / 16;
  const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
  const int pm = arow(m);
  // UNIT: Vs holds the raw tile; mask to the unit lower trapezoid whose
  // diagonal is at row i0 + c.  (Every other caller stages V with zero
  // padding, so no mask is needed there.)
  auto vmask = [&](double val, int r, int c) -> double {
    if (UNIT) {
      const int d = i0 + c;
      val = r > d ? val : (r == d ? 1.0 : 0.0);
      val = c < kb ? val : 0.0;
    }
    return val;
  };
  // ---- W = A1 + V^T X.  MFMA (rb, s): k index at lane group g = row 16rb+4g+s.
  // Operands of row block rb+1 are loaded while the MFMAs of rb issue; two
  // accumulator sets (even / odd row blocks) keep 2*WB independent chains.
  d4 W[WB], Wo[WB];
#pragma unroll
  for (int wb = 0; wb < WB; wb++) {
    W[wb] = HAS_A1 ? A1v[wb] : d4{0.0, 0.0, 0.0, 0.0};
    Wo[wb] = d4{0.0, 0.0, 0.0, 0.0};
  }
  double op1[2][4 * WB];
  auto ld1 = [&](double (&o)[4 * WB], int rb) {
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int wb = 0; wb < WB; wb++)
        o[s * WB + wb] = vmask(Vs[(16 * wb + pm) * LDV + 16 * rb + 4 * g + s],
                               16 * rb + 4 * g + s, 16 * wb + pm);
  };
  ld1(op1[0], 0);
#pragma unroll
  for (int rb = 0; rb < NBR; rb++) {
    if (rb + 1 < NBR) ld1(op1[(rb + 1) & 1], rb + 1);
    if (rb >= rb0) {
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int wb = 0; wb < WB; wb++) {
          const double a = op1[rb & 1][s * WB + wb];
          if (rb & 1)
            Wo[wb] = mfma(a, acc[rb][s], Wo[wb]);
          else
            W[wb] = mfma(a, acc[rb][s], W[wb]);
        }
    }
    __builtin_amdgcn_sched_group_barrier(0x100, 4 * WB, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 4 * WB, 0);
  }
#pragma unroll
  for (int wb = 0; wb < WB; wb++) W[wb] += Wo[wb];

  // ---- X -= V W operand chunks: chunk q = (wb, s, h) covers row blocks
  // h*CH .. h*CH+CH-1; k index of MFMA (wb, s) at lane group g = W row 16wb+4g+s.
  constexpr int CH = NBR < 8 ? NBR : 8;
  constexpr int NH = NBR / CH;
  constexpr int NCH = WB * 4 * NH;
  double op3[2][CH];
  auto ld3 = [&](double (&o)[CH], int q) {
    const int wb = q / (4 * NH), s = (q / NH) % 4, h = q % NH;
#pragma unroll
    for (int i = 0; i < CH; i++) {
      const int rb = h * CH + i;
      o[i] = vmask(Vs[(16 * wb + 4 * g + s) * LDV + 16 * rb + pm], 16 * rb + pm, 16 * wb + 4 * g + s);
    }
  };
  ld3(op3[0], 0);  // overlaps the T^T product below

  // ---- W2 = T^T W (T^T lower triangular)
  d4 W2[WB];
#pragma unroll
  for (int wb = 0; wb < WB; wb++) W2[wb] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int wi = 0; wi < WB; wi++)
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int wo = wi; wo < WB; wo++) {
        double a = TTs[(16 * wo + pm) * LDT + 16 * wi + 4 * g + s];
        W2[wo] = mfma(a, W[wi][s], W2[wo]);
      }
  if (HAS_A1 && ok) {  // A1 -= W2 (store only: no load in flight behind the prefetch)
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
  // ---- X -= V W, row blocks innermost (consecutive MFMAs are independent)
#pragma unroll
  for (int q = 0; q < NCH; q++) {
    if (q + 1 < NCH) ld3(op3[(q + 1) & 1], q + 1);
    const int wb = q / (4 * NH), s = (q / NH) % 4, h = q % NH;
#pragma unroll
    for (int i = 0; i < CH; i++) {
      const int rb = h * CH + i;
      if (rb >= rb0) acc[rb] = mfma(op3[q & 1][i], W2[wb][s], acc[rb]);
    }
    __builtin_amdgcn_sched_group_barrier(0x100, CH, 1);
    __builtin_amdgcn_sched_group_barrier(0x008, CH, 1);
  }
}

// --------------------------------------------------------------------------
// Panel factorisation (one ib-column panel of GEQRT or TSQRT).
//  GEQRT (TS=false): panel columns 0..sb-1 = tile columns ii.., rows ii..NB-1;
//    column c's reflector has its unit element at row d = ii+c.
//  TSQRT (TS=true): triangle block (sb x sb, S.A1) on top of the NB-row
//    panel; reflector c = [e_c ; v_c].
// Wave w owns the panel columns j = 4t + w (t < 8) and keeps them in
// registers for the whole panel (lane L holds rows L + 64q; for TSQRT lanes
// 0..31 also hold rows 0..31 of the triangle block).  Only the reflectors
// travel through LDS, each announced by an LDS counter instead of a barrier.
// --------------------------------------------------------------------------
template <int NB, bool TS>
__device__ __forceinline__ void gen_reflector(Smem<NB> &S, int c, int d, double (&x)[NB / 64],
                                              double &a1) {
  constexpr int LDV = Smem<NB>::LDV;
  const int lane = threadIdx.x & 63;
  double ss = 0.0;
#pragma unroll
  for (int q = 0; q < NB / 64; q++) {
    int r = lane + 64 * q;
    if (r > d) ss += x[q] * x[q];
  }
  ss = wave_sum1(ss);
  double alpha;
  if (TS) {
    alpha = readlane_d(a1, c);
  } else {
    // the diagonal element: register d / 64 of lane d % 64 (d is uniform)
    double xq = x[0];
#pragma unroll
    for (int q = 1; q < NB / 64; q++)
      if (q == (d >> 6)) xq = x[q];
    alpha = readlane_d(xq, d & 63);
  }
  // LAPACK dlarfg: beta = -sign(alpha) * ||(alpha, x)||.  ss is already an
  // unscaled sum of squares, so the norm is taken directly as
  // sqrt(alpha^2 + ss) (dlapy2's scaling would add no range here and costs a
  // division and a second square root on the per-column chain).
  double tau, beta, scal;
  if (ss == 0.0) {
    tau = 0.0;
    beta = alpha;
    scal = 1.0;
  } else {
    beta = -copysign(sqrt(fma(alpha, alpha, ss)), alpha);
    tau = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
  }
#pragma unroll
  for (int q = 0; q < NB / 64; q++) {
    int r = lane + 64 * q;
    // branch-free: every reader of this LDS column masks rows <= d itself,
    // and the panel's final write-back overwrites the whole column
    x[q] = r > d ? x[q] * scal : ((!TS && r == d) ? beta : x[q]);
    S.V[c * LDV + r] = x[q];
  }
  if (TS && lane == c) a1 = beta;
  if (lane == 0) S.tau[c] = tau;
}

// Device fault word of the library: nonzero after a bounded wait inside a
// task gave up (the task then finishes on whatever it read, so its results
// are wrong).  Read and cleared by the host after every synchronisation of a
// host-pointer entry point and by hq_sync.
__device__ int g_fault;

// Wait (bounded) until the LDS counter reaches val; acquire at workgroup
// scope.  Giving up records the fault word (1 + the awaited count).
__device__ __forceinline__ void lds_wait_geq(int *flag, int val) {
  for (unsigned spins = 0;; spins++) {
    int f = __hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (__builtin_amdgcn_readfirstlane(f) >= val) break;
    if (spins >= (1u << 22)) {
      if ((threadIdx.x & 63) == 0) __hip_atomic_store(&g_fault, 1 + val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
  }
}

// Apply reflector c (v, tau) to this wave's panel columns j = 4t + wave with
// j > c + 1, for the register slots t >= T0 (earlier slots are finished).
template <int NB, bool TS, int T0>
__device__ __forceinline__ void panel_apply(double (&x)[HQ_IBMAX / 4][NB / 64],
                                            double (&a1)[HQ_IBMAX / 4], const double (&v)[NB / 64],
                                            double tau, int c, int sb, int wave, int lane) {
  constexpr int NQ = NB / 64, JW = HQ_IBMAX / 4, N = JW - T0;
  double dot[N];
#pragma unroll
  for (int t = 0; t < N; t++) {
    // (TSQRT: the triangle's row c, held by lane c, joins the sum)
    dot[t] = (TS && lane == c) ? a1[T0 + t] : 0.0;
#pragma unroll
    for (int q = 0; q < NQ; q++) dot[t] += v[q] * x[T0 + t][q];
  }
  wave_sums<N>(dot);
#pragma unroll
  for (int t = T0; t < JW; t++) {
    const int j = 4 * t + wave;
    if (j > c + 1 && j < sb) {
      const double w = tau * dot[t - T0];
#pragma unroll
      for (int q = 0; q < NQ; q++) x[t][q] -= v[q] * w;
      if (TS && lane == c) a1[t] -= w;
    }
  }
}

template <int NB, bool TS>
__device__ __forceinline__ void panel_factor(Smem<NB> &S, int ii, int sb) {
  constexpr int LDV = Smem<NB>::LDV, LDT = Smem<NB>::LDT;
  constexpr int NQ = NB / 64, JW = HQ_IBMAX / 4;
  static_assert(JW == 8, "panel_factor: the live-slot dispatch assumes 8 register slots");
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double x[JW][NQ], a1[JW];
#pragma unroll
  for (int t = 0; t < JW; t++) {
    const int j = 4 * t + wave;
    const bool valid = j < sb;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int r = lane + 64 * q;
      x[t][q] = (valid && (TS || r >= ii)) ? S.V[j * LDV + r] : 0.0;
    }
    a1[t] = (TS && valid && lane < HQ_IBMAX) ? S.A1[j * LDT + lane] : 0.0;
  }
  // S.nref counts the published reflectors (0 on entry).  Reflector c+1 is
  // produced by the owner of column c+1 right after it applies reflector c
  // to that column; its other columns follow, so the next reflector's chain
  // overlaps the bulk of the column updates.
  if (wave == 0) {
    gen_reflector<NB, TS>(S, 0, TS ? -1 : ii, x[0], a1[0]);
    __hip_atomic_store(&S.nref, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  for (int c = 0; c < sb; c++) {
    lds_wait_geq(&S.nref, c + 1);
    const double tau = S.tau[c];
    const int d = TS ? -1 : ii + c;
    double v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      int r = lane + 64 * q;
      double val = S.V[c * LDV + r];
      v[q] = TS ? val : (r > d ? val : (r == d ? 1.0 : 0.0));
    }
    if (c + 1 < sb && wave == ((c + 1) & 3)) {
#pragma unroll
      for (int t = 0; t < JW; t++) {
        if (t == ((c + 1) >> 2)) {
          double dt = (TS && lane == c) ? a1[t] : 0.0;
#pragma unroll
          for (int q = 0; q < NQ; q++) dt += v[q] * x[t][q];
          dt = wave_sum1(dt);
          const double w = tau * dt;
#pragma unroll
          for (int q = 0; q < NQ; q++) x[t][q] -= v[q] * w;
          if (TS && lane == c) a1[t] -= w;
          gen_reflector<NB, TS>(S, c + 1, TS ? -1 : ii + c + 1, x[t], a1[t]);
        }
      }
      __hip_atomic_store(&S.nref, c + 2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    if (c + 2 < sb) {
      // this wave's columns j = 4t + wave still to update are those with
      // j > c + 1, i.e. t >= tmin; reduce only the slots from the last
      // power-of-two group covering them
      const int live = JW - ((c + 5 - wave) >> 2);
      if (live > 4) panel_apply<NB, TS, 0>(x, a1, v, tau, c, sb, wave, lane);
      else if (live > 2) panel_apply<NB, TS, JW - 4>(x, a1, v, tau, c, sb, wave, lane);
      else if (live == 2) panel_apply<NB, TS, JW - 2>(x, a1, v, tau, c, sb, wave, lane);
      else if (live == 1) panel_apply<NB, TS, JW - 1>(x, a1, v, tau, c, sb, wave, lane);
    }
  }
  __syncthreads();
  // the finished panel back to LDS (rows < ii of a GEQRT panel as zeros)
#pragma unroll
  for (int t = 0; t < JW; t++) {
    const int j = 4 * t + wave;
    if (j < sb) {
#pragma unroll
      for (int q = 0; q < NQ; q++) S.V[j * LDV + lane + 64 * q] = x[t][q];
      if (TS && lane < HQ_IBMAX) S.A1[j * LDT + lane] = a1[t];
    }
  }
  __syncthreads();
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
//   T(0:c, c) = -tau_c T(0:c, 0:c) G(0:c, c),  T(c,c) = tau_c,
// in right-looking order: lane p of wave 0 holds row p of T; as soon as
// T(p, q) is final its terms T(p,q) G(q,c) are added to every later column
// c, so finishing a column waits on one product, not on a dot product of
// growing length.  Row q+1 of G is read (broadcast LDS reads) while row q is
// applied.  Each step is its own branch and scheduling region, which keeps
// the register use of the unrolled recursion bounded.  Writes T^T to S.TT,
// T to global (ld = ib, block at column ii) and, if taug, tau to TAU.
template <int NB>
__device__ __forceinline__ void build_t(Smem<NB> &S, int sb, double *Tg, int ib, double *taug) {
  constexpr int LDT = Smem<NB>::LDT;
  const int p = threadIdx.x;  // wave 0 only
  double Tr[HQ_IBMAX];
  double gr[2][HQ_IBMAX];
#pragma unroll
  for (int c = 0; c < HQ_IBMAX; c++) {
    Tr[c] = 0.0;
    gr[1][c] = 0.0;
    gr[0][c] = c > 0 ? S.G[c] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < HQ_IBMAX; q++) {
    if (q < sb) {
      if (q + 1 < HQ_IBMAX) {
#pragma unroll
        for (int c = q + 2; c < HQ_IBMAX; c++) gr[(q + 1) & 1][c] = S.G[(q + 1) * LDT + c];
      }
      const double tq = S.tau[q];
      Tr[q] = (p < q) ? -tq * Tr[q] : (p == q ? tq : 0.0);
#pragma unroll
      for (int c = q + 1; c < HQ_IBMAX; c++) Tr[c] += Tr[q] * gr[q & 1][c];
    } else {
      Tr[q] = 0.0;
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  if (p < HQ_IBMAX) {
#pragma unroll
    for (int c = 0; c < HQ_IBMAX; c++) {
      S.TT[c * LDT + p] = Tr[c];
      if (c < sb && p < ib) Tg[(size_t)c * ib + p] = Tr[c];
    }
  }
  if (taug && p < sb) taug[p] = S.tau[p];
}

// --------------------------------------------------------------------------
// Tasks.
// --------------------------------------------------------------------------

typedef __attribute__((address_space(3))) char lds_char;

// Tile columns col0 .. col0+ncol-1 (all NB rows) into the LDS panel buffer
// Vs (column stride LDV); padding columns ncol .. IBP-1 set to zero.
// NB >= 128 uses direct-to-LDS 16-byte loads; the caller waits (s_waitcnt)
// and synchronises before reading.
template <int NB, int WB>
__device__ __forceinline__ void stage_columns(double *Vs, const double *tile, int col0, int ncol) {
  constexpr int LDV = Smem<NB>::LDV;
  constexpr int IBP = 16 * WB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if constexpr (NB >= 128) {
    constexpr int HALVES = NB / 128;
    for (int q = wave; q < ncol * HALVES; q += HQ_THREADS / 64) {
      const int c = q / HALVES, h = q - c * HALVES;
      const double *src = tile + (size_t)(col0 + c) * NB + 128 * h + 2 * lane;
      lds_char *dst = (lds_char *)(Vs + c * LDV + 128 * h);
      __builtin_amdgcn_global_load_lds((const void *)src, (void __attribute__((address_space(3))) *)dst,
                                       16, 0, 0);
    }
  } else {
    for (int e = tid; e < ncol * NB; e += HQ_THREADS) {
      const int c = e / NB, r = e - c * NB;
      Vs[c * LDV + r] = tile[(size_t)(col0 + c) * NB + r];
    }
  }
  for (int e = tid; e < (IBP - ncol) * NB; e += HQ_THREADS) {
    const int c = ncol + e / NB, r = e - (c - ncol) * NB;
    Vs[c * LDV + r] = 0.0;
  }
}

// GEQRT and TSQRT are sequences over the ib blocks of the tile of
//   panel b:     dgeqr2 / tsqr2 of the block's columns + dlarft into T, TAU
//   trailing b:  the block reflector applied to the columns right of it
// The whole-tile tasks run both in one workgroup; the DAG scheduler runs the
// panel as one task and the trailing update as one task per 64-column slab,
// so the flat-tree chain down a tile column (every TSQRT of a step updates
// R(k,k)) advances one panel at a time instead of one tile at a time.  Each
// column's arithmetic is the same either way.

// W = V T_s for each 16-column sub-block s of the panel's block (T_s its
// 16x16 diagonal block; ib % 16 == 0), from V in S.V and T^T in S.TT,
// stored to columns ii.. of the W tile Wt (ld = NB).  MFMA (rb, h): rows
// 16rb.. of sub-block h, contraction over its 16 reflectors.
template <int NB>
__device__ __forceinline__ void store_vt(Smem<NB> &S, double *Wt, int ii, int sb) {
  constexpr int LDV = Smem<NB>::LDV, LDT = Smem<NB>::LDT;
  const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int pm = arow(m);
  const int nq = (sb / 16) * (NB / 16);
  for (int q = wave; q < nq; q += HQ_THREADS / 64) {
    const int h = q / (NB / 16), rb = q - h * (NB / 16);
    d4 acc = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; s++)
      acc = mfma(S.V[(16 * h + 4 * g + s) * LDV + 16 * rb + pm], S.TT[(16 * h + m) * LDT + 16 * h + 4 * g + s],
                 acc);
    *reinterpret_cast<d4 *>(Wt + (size_t)(ii + 16 * h + m) * NB + 16 * rb + 4 * g) = acc;
  }
}

// GEQRT panel of columns ii .. ii+sb-1: R / V written back to A, T and TAU
// of the block written; leaves V (unit lower trapezoidal) in S.V and T^T in
// S.TT.
template <int NB, int WB>
__device__ __forceinline__ void geqrt_panel(Smem<NB> &S, double *A, double *T, double *TAU, int ib, int ii,
                                            double *Wt = nullptr, double *Vd = nullptr,
                                            unsigned short *VHt = nullptr) {
  constexpr int LDV = Smem<NB>::LDV;
  const int tid = threadIdx.x, wave = tid >> 6;
  const int sb = min(ib, NB - ii);
  stage_columns<NB, WB>(S.V, A, ii, sb);
  if (tid == 0) S.nref = 0;
  __builtin_amdgcn_s_waitcnt(0);
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
  if (VHt) store_vh<NB>(S, VHt, ib, ii, sb);
  if (Wt) {
    store_vt<NB>(S, Wt, ii, sb);
    for (int e = tid; e < sb * NB; e += HQ_THREADS) {
      const int c = e / NB, r = e - c * NB;
      Vd[(size_t)(ii + c) * NB + r] = S.V[c * LDV + r];
    }
  }
}

// TSQRT panel of columns ii .. ii+sb-1 of [R(A1); A2]: V2 written to A2, the
// R block to A1, T and TAU of the block written; leaves V2 in S.V and T^T in
// S.TT.
template <int NB, int WB>
__device__ __forceinline__ void tsqrt_panel(Smem<NB> &S, double *A1, double *A2, double *T, double *TAU, int ib,
                                            int ii, unsigned short *VHt = nullptr) {
  constexpr int LDV = Smem<NB>::LDV, LDT = Smem<NB>::LDT;
  constexpr int IBP = 16 * WB;
  const int tid = threadIdx.x, wave = tid >> 6;
  const int sb = min(ib, NB - ii);
  stage_columns<NB, WB>(S.V, A2, ii, sb);
  if (tid == 0) S.nref = 0;
  for (int e = tid; e < IBP * IBP; e += HQ_THREADS) {
    int c = e / IBP, p = e - c * IBP;
    S.A1[c * LDT + p] = (c < sb && p <= c) ? A1[(size_t)(ii + c) * NB + ii + p] : 0.0;
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  HQ_PHASE(0);
  panel_factor<NB, true>(S, ii, sb);
  HQ_PHASE(1);
  for (int e = tid; e < sb * NB; e += HQ_THREADS) {
    int c = e / NB, r = e - c * NB;
    A2[(size_t)(ii + c) * NB + r] = S.V[c * LDV + r];
  }
  for (int e = tid; e < sb * sb; e += HQ_THREADS) {
    int c = e / sb, p = e - c * sb;
    if (p <= c) A1[(size_t)(ii + c) * NB + ii + p] = S.A1[c * LDT + p];
  }
  __syncthreads();
  HQ_PHASE(2);
  gram<NB, WB>(S, 0);
  __syncthreads();
  HQ_PHASE(3);
  if (wave == 0) build_t<NB>(S, sb, T + (size_t)ii * ib, ib, TAU + ii);
  __syncthreads();
  HQ_PHASE(4);
  if (VHt) store_vh<NB>(S, VHt, ib, ii, sb);
}

// For a split trailing task: V of the block at column ii of tile Vt (GEQRT:
// made unit lower trapezoidal, as the panel leaves it) and T^T from the T
// array, into S.V / S.TT.
template <int NB, int WB, bool TS>
__device__ __forceinline__ void stage_panel(Smem<NB> &S, const double *Vt, const double *T, int ib, int ii) {
  constexpr int LDV = Smem<NB>::LDV, LDT = Smem<NB>::LDT;
  constexpr int IBP = 16 * WB;
  const int tid = threadIdx.x;
  const int sb = min(ib, NB - ii);
  stage_columns<NB, WB>(S.V, Vt, ii, sb);
  for (int e = tid; e < IBP * IBP; e += HQ_THREADS) {
    const int c = e / IBP, p = e - c * IBP;
    S.TT[c * LDT + p] = (c < sb && p <= c) ? T[(size_t)(ii + c) * ib + p] : 0.0;
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (!TS) {
    for (int e = tid; e < sb * NB; e += HQ_THREADS) {
      const int c = e / NB, r = e - c * NB;
      if (r <= ii + c) S.V[c * LDV + r] = (r == ii + c) ? 1.0 : 0.0;
    }
    __syncthreads();
  }
}

// Trailing update of the block at column ii (V in S.V, T^T in S.TT) on the
// 16-column groups c0 + 16*grp, grp = wave, wave + 4, ..., below c1; columns
// of the block and left of it are not touched.  GEQRT (!TS): rows >= ii of
// X = A.  TSQRT: [A1 rows ii..; X = A2].
template <int NB, int WB, bool TS>
__device__ __forceinline__ void panel_trailing(Smem<NB> &S, double *X, double *A1, int ib, int ii, int c0,
                                               int c1) {
  const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, wave = tid >> 6;
  const int sb = min(ib, NB - ii), j0 = ii + sb;
  const int rb0 = TS ? 0 : ii / 16;
  for (int grp = wave; c0 + 16 * grp < c1; grp += 4) {
    const int col = c0 + 16 * grp + m;
    const bool ok = col >= j0 && col < c1;
    d4 acc[NB / 16];
    load_strip<NB>(acc, X, col, ok, rb0);
    d4 A1v[WB];
    if (TS) {
      load_a1<WB>(A1v, A1 + (size_t)col * NB + ii, sb, ok);
      apply_block<NB, WB, true>(acc, 0, S.V, S.TT, A1v, A1 + (size_t)col * NB + ii, sb, ok);
      store_strip<NB>(acc, X, col, ok, 0);
    } else {
      apply_block<NB, WB, false>(acc, rb0, S.V, S.TT, A1v, nullptr, sb, ok);
      // rows >= ii only: rows above it in the first row block are R rows of
      // earlier blocks (unchanged here, and updated by other tasks meanwhile)
      const int g = lane >> 4;
      if (ok) {
#pragma unroll
        for (int rb = 0; rb < NB / 16; rb++) {
          if (rb > rb0 || (rb == rb0 && (ii & 15) == 0)) {
            *reinterpret_cast<d4 *>(X + (size_t)col * NB + 16 * rb + 4 * g) = acc[rb];
          } else if (rb == rb0) {
#pragma unroll
            for (int v = 0; v < 4; v++)
              if (16 * rb + 4 * g + v >= ii) X[(size_t)col * NB + 16 * rb + 4 * g + v] = acc[rb][v];
          }
        }
      }
    }
  }
}

// The outputs of panel blocks wholly inside a batch matrix's zero padding
// (columns ii.. of the tile, from column `live` on): the block's reflectors
// are identities, so TAU = 0 and T = 0 for its columns, and for a diagonal
// tile W = 0 and unit columns of V.
template <int NB>
__device__ __forceinline__ void pad_panel_blocks(double *T, double *TAU, int ib, int ii, double *Wt, double *Vd) {
  for (int e = threadIdx.x; e < (NB - ii) * ib; e += HQ_THREADS) T[(size_t)ii * ib + e] = 0.0;
  for (int c = ii + threadIdx.x; c < NB; c += HQ_THREADS) TAU[c] = 0.0;
  if (Wt)
    for (int e = threadIdx.x; e < (NB - ii) * NB; e += HQ_THREADS) {
      const int c = ii + e / NB, r = e % NB;
      Wt[(size_t)c * NB + r] = 0.0;
      Vd[(size_t)c * NB + r] = r == c ? 1.0 : 0.0;
    }
}

// GEQRT = CORE_dgeqrt(nb, nb, ib, A, nb, T, ib, TAU); columns from `live` on
// are zero padding (pad_panel_blocks)
template <int NB, int WB>
__device__ __forceinline__ void task_geqrt(Smem<NB> &S, double *A, double *T, double *TAU, int ib,
                                           double *Wt = nullptr, double *Vd = nullptr,
                                           unsigned short *VHt = nullptr, int live = NB) {
  const int c1 = min(NB, (live + 15) & ~15);
  int ii = 0;
  for (; ii < live; ii += ib) {
    geqrt_panel<NB, WB>(S, A, T, TAU, ib, ii, Wt, Vd, VHt);
    panel_trailing<NB, WB, false>(S, A, nullptr, ib, ii, ii + min(ib, NB - ii), c1);
    __syncthreads();
  }
  if (ii < NB) pad_panel_blocks<NB>(T, TAU, ib, ii, Wt, Vd);
}

// TSQRT = CORE_dtsqrt(nb, nb, ib, A1, nb, A2, nb, T, ib, TAU); columns from
// `live` on are zero padding
template <int NB, int WB>
__device__ __forceinline__ void task_tsqrt(Smem<NB> &S, double *A1, double *A2, double *T, double *TAU, int ib,
                                           unsigned short *VHt = nullptr, int live = NB) {
  const int c1 = min(NB, (live + 15) & ~15);
  int ii = 0;
  for (; ii < live; ii += ib) {
    tsqrt_panel<NB, WB>(S, A1, A2, T, TAU, ib, ii, VHt);
    panel_trailing<NB, WB, true>(S, A2, A1, ib, ii, ii + min(ib, NB - ii), c1);
    __syncthreads();
    HQ_PHASE(5);
  }
  if (ii < NB) pad_panel_blocks<NB>(T, TAU, ib, ii, nullptr, nullptr);
}

// Split tasks of the DAG scheduler: panel b, and the trailing update of
// block b on 64-column slab `slab` (V and T read back from memory).
template <int NB, int WB>
__device__ __forceinline__ void task_geqrt_update(Smem<NB> &S, double *A, const double *T, int ib, int b,
                                                  int slab) {
  const int ii = b * ib;
  stage_panel<NB, WB, false>(S, A, T, ib, ii);
  panel_trailing<NB, WB, false>(S, A, nullptr, ib, ii, slab * HQ_SLAB, min(NB, (slab + 1) * HQ_SLAB));
}
template <int NB, int WB>
__device__ __forceinline__ void task_tsqrt_update(Smem<NB> &S, double *A1, double *A2, const double *T, int ib,
                                                  int b, int slab) {
  const int ii = b * ib;
  stage_panel<NB, WB, true>(S, A2, T, ib, ii);
  panel_trailing<NB, WB, true>(S, A2, A1, ib, ii, slab * HQ_SLAB, min(NB, (slab + 1) * HQ_SLAB));
}

// Staging of one ib block of a reflector tile and its T^T for the LARFB /
// SSRFB pipeline.  V (raw tile columns i0..i0+kb) goes straight from global
// memory into LDS with direct-to-LDS 16-byte loads (one wave instruction per
// 128 doubles of a column; no registers involved); nb = 64 uses ordinary
// loads.  T^T (at most 32 x 32) goes through registers: each thread holds
// IBP*IBP/256 entries between the load and the LDS store.
template <int NB, int WB>
struct BlockRegs {
  static constexpr int IBP = 16 * WB;
  static constexpr int NPT = (IBP * IBP + HQ_THREADS - 1) / HQ_THREADS;
  static constexpr int NPV = NB >= 128 ? 1 : IBP * NB / HQ_THREADS;
  double t[NPT];
  double v[NPV];
};

template <int NB, int WB>
__device__ __forceinline__ void fetch_block(BlockRegs<NB, WB> &R, double *Vs_next, const double *V,
                                            const double *T, int ib, int i0, int kb) {
  constexpr int IBP = 16 * WB;
  constexpr int LDV = Smem<NB>::LDV;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if constexpr (NB >= 128) {
    constexpr int HALVES = NB / 128;
    const int nq = kb * HALVES;
    for (int q = wave; q < nq; q += HQ_THREADS / 64) {
      const int c = q / HALVES, h = q - c * HALVES;
      const double *src = V + (size_t)(i0 + c) * NB + 128 * h + 2 * lane;
      lds_char *dst = (lds_char *)(Vs_next + c * LDV + 128 * h);
      __builtin_amdgcn_global_load_lds((const void *)src, (void __attribute__((address_space(3))) *)dst,
                                       16, 0, 0);
    }
  } else {
#pragma unroll
    for (int i = 0; i < BlockRegs<NB, WB>::NPV; i++) {
      const int e = tid + HQ_THREADS * i;
      const int c = e / NB, r = e - c * NB;
      R.v[i] = c < kb ? V[(size_t)(i0 + c) * NB + r] : 0.0;
    }
  }
#pragma unroll
  for (int i = 0; i < BlockRegs<NB, WB>::NPT; i++) {
    const int e = tid + HQ_THREADS * i;
    const int c = e / IBP, p = e - c * IBP;
    R.t[i] = (e < IBP * IBP && c < kb && p <= c) ? T[(size_t)(i0 + c) * ib + p] : 0.0;
  }
}

template <int NB, int WB>
__device__ __forceinline__ void store_block(const BlockRegs<NB, WB> &R, double *Vs, double *TTs,
                                            int kb) {
  constexpr int IBP = 16 * WB;
  constexpr int LDV = Smem<NB>::LDV, LDT = Smem<NB>::LDT;
  const int tid = threadIdx.x;
  if constexpr (NB >= 128) {
    // a ragged last block: the columns the direct-to-LDS loads skipped are zero
    for (int e = tid; e < (IBP - kb) * NB; e += HQ_THREADS) {
      const int c = kb + e / NB, r = e - (c - kb) * NB;
      Vs[c * LDV + r] = 0.0;
    }
  }
  if constexpr (NB < 128) {
#pragma unroll
    for (int i = 0; i < BlockRegs<NB, WB>::NPV; i++) {
      const int e = tid + HQ_THREADS * i;
      const int c = e / NB, r = e - c * NB;
      Vs[c * LDV + r] = R.v[i];
    }
  }
#pragma unroll
  for (int i = 0; i < BlockRegs<NB, WB>::NPT; i++) {
    const int e = tid + HQ_THREADS * i;
    const int c = e / IBP, p = e - c * IBP;
    if (e < IBP * IBP) TTs[c * LDT + p] = R.t[i];
  }
  // the direct-to-LDS loads of this block complete on the vector memory
  // counter: wait for them before the barrier that publishes the buffer
  __builtin_amdgcn_s_waitcnt(0);
}

// LARFB (UNIT, no A1) = CORE_dormqr(Left, Trans, nb, nb, nb, ib, V, nb, T, ib, X, nb)
// SSRFB (full V, A1)  = CORE_dtsmqr(Left, Trans, nb, nb, nb, nb, nb, ib, A1, nb, X, nb, V, nb, T, ib)
// on one 64-column slab.  The strip of X stays in registers for all blocks;
// block b+1 of V/T is loaded while block b is applied (two LDS buffers, one
// barrier per block).
template <int NB, int WB, bool UNIT>
__device__ __forceinline__ void task_update(Smem<NB> &S, const double *V, const double *T,
                                            double *X, double *A1, int ib, int slab) {
  const int lane = threadIdx.x & 63, m = lane & 15, wave = threadIdx.x >> 6;
  const int col = slab * HQ_SLAB + 16 * wave + m;
  const bool ok = col < NB;
  const int nblk = (NB + ib - 1) / ib;
  double *a1c = UNIT ? nullptr : A1 + (size_t)col * NB;
  d4 acc[NB / 16];
  d4 A1v[WB], A1n[WB];
  BlockRegs<NB, WB> R;
  fetch_block<NB, WB>(R, S.V, V, T, ib, 0, min(ib, NB));
  load_strip<NB>(acc, X, col, ok, 0);
  if (!UNIT) load_a1<WB>(A1v, a1c, min(ib, NB), ok);
  store_block<NB, WB>(R, S.V, S.TT, min(ib, NB));
  __syncthreads();
  for (int b = 0; b < nblk; b++) {
    const int i0 = b * ib, kb = min(ib, NB - i0);
    const bool more = b + 1 < nblk;
    double *Vn = (b & 1) ? S.V : S.V1;
    double *TTn = (b & 1) ? S.TT : S.TT1;
    // everything for block b+1 is issued before the MFMAs of block b and
    // waited for only at the end of the block
    if (more) {
      fetch_block<NB, WB>(R, Vn, V, T, ib, i0 + ib, min(ib, NB - i0 - ib));
      if (!UNIT) load_a1<WB>(A1n, a1c + i0 + ib, min(ib, NB - i0 - ib), ok);
    }
    const double *Vs = (b & 1) ? S.V1 : S.V;
    const double *TTs = (b & 1) ? S.TT1 : S.TT;
    if (UNIT)
      apply_block<NB, WB, false, true>(acc, i0 / 16, Vs, TTs, A1v, nullptr, kb, ok, i0);
    else
      apply_block<NB, WB, true>(acc, 0, Vs, TTs, A1v, a1c + i0, kb, ok);
    if (more) {
      store_block<NB, WB>(R, Vn, TTn, min(ib, NB - i0 - ib));
#pragma unroll
      for (int wb = 0; wb < WB; wb++) A1v[wb] = A1n[wb];
    }
    __syncthreads();
  }
  store_strip<NB>(acc, X, col, ok, 0);
}

// --------------------------------------------------------------------------
// Fused update pipeline for LARFB / SSRFB on nb >= 128.
// Each ib block of reflectors is applied as sub-blocks of at most 16
// (Q_b = Q_b0 Q_b1 with the diagonal 16x16 blocks of the block's T: the
// compact WY form of a product of two blocks is exactly that).  Pass s of
// the pipeline interleaves, four row blocks at a time,
//   X -= V_s W2_s            (update with sub-block s)
//   W_{s+1} += V_{s+1}^T X   (the rows just updated)
// so the MFMA stream has no phase boundary inside a pass; after the pass
// W2_{s+1} = -T_{s+1}^T W_{s+1} (four MFMAs) and, for SSRFB, the A1 rows of
// sub-block s+1 are written.  V sub-blocks live in a ring of four LDS
// buffers of 16 columns: the direct-to-LDS loads of sub-block s+3 are issued
// one per phase in the first half of pass s and waited for at its end.  One
// barrier per sub-block.
// --------------------------------------------------------------------------
struct SubBlk {
  int i0, kb, p0;  // first reflector (tile column), width, row offset inside its T block
};
template <int NB>
__device__ __forceinline__ SubBlk sub_blk(int s, int ib) {
  const int nsb = (ib + 15) >> 4;
  const int b = s / nsb, h = s - b * nsb;
  SubBlk r;
  r.p0 = 16 * h;
  r.i0 = b * ib + r.p0;
  r.kb = min(16, min(ib, NB - b * ib) - r.p0);
  return r;
}
// The sub-block after x (x itself when x is the last one).
template <int NB>
__device__ __forceinline__ SubBlk next_sub(const SubBlk &x, int ib) {
  if (x.i0 + x.kb >= NB) return x;
  const int i0b = x.i0 - x.p0, kbb = min(ib, NB - i0b);
  SubBlk r;
  if (x.p0 + 16 < kbb) {
    r.i0 = x.i0 + 16;
    r.p0 = x.p0 + 16;
    r.kb = min(16, kbb - r.p0);
  } else {
    r.i0 = i0b + ib;
    r.p0 = 0;
    r.kb = min(16, min(ib, NB - r.i0));
  }
  return r;
}
template <int NB>
__device__ __forceinline__ int sub_count(int ib) {
  const int nsb = (ib + 15) >> 4, nblk = (NB + ib - 1) / ib;
  return (nblk - 1) * nsb + (NB - (nblk - 1) * ib + 15) / 16;
}
template <int NB>
__device__ __forceinline__ double *fvbuf(Smem<NB> &S, int i) {
  constexpr int LDV = Smem<NB>::LDV;
  return (i < 2 ? S.V : S.V1) + (i & 1) * 16 * LDV;
}
constexpr int LDT16 = 17;
// s_waitcnt immediate: vector memory count <= vm, LDS / scalar count 0, no
// export wait (gfx9 encoding: vm[3:0] at 3:0, vm[5:4] at 15:14, exp 6:4,
// lgkm 11:8)
constexpr int waitcnt_vm_lgkm0(int vm) { return (vm & 15) | ((vm >> 4) << 14) | (7 << 4); }
template <int NB>
__device__ __forceinline__ double *ftbuf(Smem<NB> &S, int i) {
  return (i < 2 ? S.TT : S.TT1) + (i & 1) * 16 * LDT16;
}
// direct-to-LDS loads per wave for one 16-column sub-block
template <int NB>
struct FDma {
  static constexpr int HALVES = NB / 128;
  static constexpr int NOPS = 16 * HALVES / (HQ_THREADS / 64);
};
// Load op q (< NOPS) of this wave for sub-block sb into ring slot Vb: 128
// rows of one column per op.  Ops past the block's width repeat its last
// column, so every sub-block costs exactly NOPS ops per wave.
template <int NB>
__device__ __forceinline__ void dma_op(int q, double *Vb, const double *V, const SubBlk &sb) {
  constexpr int LDV = Smem<NB>::LDV, HALVES = FDma<NB>::HALVES;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int Q = wave + (HQ_THREADS / 64) * q;
  const int c = min(Q / HALVES, sb.kb - 1), h = Q - (Q / HALVES) * HALVES;
  const double *src = V + (size_t)(sb.i0 + c) * NB + 128 * h + 2 * lane;
  lds_char *dst = (lds_char *)(Vb + c * LDV + 128 * h);
  __builtin_amdgcn_global_load_lds((const void *)src, (void __attribute__((address_space(3))) *)dst, 16, 0, 0);
}

// operands of update phase (group G = row blocks 2G, 2G+1): A(i, k) of MFMA
// (sp, rb) = V(16rb + arow(i), 4k + sp)
template <int NB, bool UNIT>
__device__ __forceinline__ void ld_u(double (&o)[8], const double *Vb, int G, int i0, int kb) {
  constexpr int LDV = Smem<NB>::LDV;
  const int lane = threadIdx.x & 63, g = lane >> 4, pm = arow(lane & 15);
#pragma unroll
  for (int sp = 0; sp < 4; sp++)
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const int r = 16 * (2 * G + i) + pm, c = 4 * g + sp;
      o[2 * sp + i] = Vb[c * LDV + r];
    }
}
// operands of accumulate phase: A(i, k) of MFMA (rb, sp) = V(16rb + 4k + sp, arow(i))
template <int NB, bool UNIT>
__device__ __forceinline__ void ld_a(double (&o)[8], const double *Vb, int G, int i0, int kb) {
  constexpr int LDV = Smem<NB>::LDV;
  const int lane = threadIdx.x & 63, g = lane >> 4, pm = arow(lane & 15);
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int sp = 0; sp < 4; sp++) {
      const int r = 16 * (2 * G + i) + 4 * g + sp, c = pm;
      o[4 * i + sp] = Vb[c * LDV + r];
    }
}

template <int NB, bool UNIT>
struct PassArgs {
  const double *Vu, *Vw, *TTw;
  int i0u, kbu, i0w, kbw;
  double *Vd;         // ring slot receiving sub-block sd
  const double *Vg;   // V tile in global memory
  SubBlk sd;
};

// One pass: phases U0 U1 A0 U2 A1 ... U(NG-1) A(NG-2) A(NG-1) (DO_U && DO_A),
// or only the U or only the A phases; a phase is 8 MFMAs on two row blocks.
// o[0] holds the operands of the first phase on entry; each phase loads the
// operands of the next one, the last the T^T operands (tto) when DO_A.
// Branch-free: LARFB runs the MFMAs of rows above its diagonal block too,
// on the zeros its staging put there.
template <int NB, bool UNIT, bool DO_U, bool DO_A, bool DO_DMA, bool DO_ST, class Hook>
__device__ __forceinline__ void fpass(d4 (&acc)[NB / 16], const d4 &W2, d4 (&Wn)[2],
                                      double (&o)[2][8], double (&tto)[4], const PassArgs<NB, UNIT> &P,
                                      double *xcol, Hook &&hook) {
  constexpr int NG = NB / 32;
  constexpr int NP = (DO_U ? NG : 0) + (DO_A ? NG : 0);
  static_assert(!DO_DMA || NP >= FDma<NB>::NOPS, "one direct-to-LDS op per phase");
  auto kind = [](int p) -> int {  // 1 = update, 0 = accumulate
    if (DO_U && !DO_A) return 1;
    if (!DO_U) return 0;
    if (p == 0) return 1;
    if (p == 2 * NG - 1) return 0;
    return (p & 1) ? 1 : 0;
  };
  auto grp = [](int p) -> int {
    if (!(DO_U && DO_A)) return p;
    if (p == 0) return 0;
    if (p == 2 * NG - 1) return NG - 1;
    return (p & 1) ? (p + 1) / 2 : p / 2 - 1;
  };
  const int lane = threadIdx.x & 63, g = lane >> 4, pm = arow(lane & 15);
#pragma unroll
  for (int p = 0; p < NP; p++) {
    if (DO_DMA && p < FDma<NB>::NOPS) dma_op<NB>(p, P.Vd, P.Vg, P.sd);
    if (p + 1 < NP) {
      if (kind(p + 1))
        ld_u<NB, UNIT>(o[(p + 1) & 1], P.Vu, grp(p + 1), P.i0u, P.kbu);
      else
        ld_a<NB, UNIT>(o[(p + 1) & 1], P.Vw, grp(p + 1), P.i0w, P.kbw);
    } else if (DO_A) {
#pragma unroll
      for (int sp = 0; sp < 4; sp++) tto[sp] = P.TTw[pm * LDT16 + 4 * g + sp];
    }
    const int G = grp(p);
    if (kind(p)) {
#pragma unroll
      for (int sp = 0; sp < 4; sp++)
#pragma unroll
        for (int i = 0; i < 2; i++) acc[2 * G + i] = mfma(o[p & 1][2 * sp + i], W2[sp], acc[2 * G + i]);
    } else {
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int sp = 0; sp < 4; sp++) Wn[sp & 1] = mfma(o[p & 1][4 * i + sp], acc[2 * G + i][sp], Wn[sp & 1]);
    }
    // next phase's LDS reads early in the phase, between its first MFMAs;
    // no instruction crosses a phase boundary, so the operands are always
    // one phase ahead and an accumulate phase never follows its update
    // directly
    // (the first MFMA goes before them: the wait for this phase's operands
    // then does not cover the new reads)
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 6, 0);
    __builtin_amdgcn_sched_barrier(0);
    // work of the pass that no MFMA of it waits for, behind its phases
    hook(p);
    // last pass: the rows of the previous phase are final; their stores go
    // out behind this phase's MFMAs instead of all at the end of the task
    if (DO_ST && p > 0) {
      const int Gp = grp(p - 1);
#pragma unroll
      for (int i = 0; i < 2; i++) *gmem(reinterpret_cast<d4 *>(xcol + 16 * (2 * Gp + i) + 4 * g)) = acc[2 * Gp + i];
    }
  }
  if (DO_ST) {
    const int Gp = grp(NP - 1);
#pragma unroll
    for (int i = 0; i < 2; i++) *gmem(reinterpret_cast<d4 *>(xcol + 16 * (2 * Gp + i) + 4 * g)) = acc[2 * Gp + i];
  }
}

// Columns kb..15 of ring slot Vb are zero (a ragged sub-block).
template <int NB>
__device__ __forceinline__ void zero_pad(double *Vb, const SubBlk &sb) {
  constexpr int LDV = Smem<NB>::LDV;
  for (int e = threadIdx.x; e < (16 - sb.kb) * NB; e += HQ_THREADS) {
    const int c = sb.kb + e / NB, r = e - (c - sb.kb) * NB;
    Vb[c * LDV + r] = 0.0;
  }
}
// This thread's entry of T^T for sub-block sb (TT slot layout c*17 + p),
// as a raw load clamped to an entry of the block, and the 0/1 factor that
// keeps only the upper triangle of the block.  The factor is applied where
// the value is stored, after the pass: a selected load would be turned
// into a load under an exec mask followed at once by a wait for every
// load in flight.
__device__ __forceinline__ double tt_raw(const double *T, int ib, const SubBlk &sb) {
  const int c = threadIdx.x >> 4, p = threadIdx.x & 15;
  const int cc = min(c, sb.kb - 1);
  return *gmem(T + (size_t)(sb.i0 + cc) * ib + sb.p0 + min(p, cc));
}
__device__ __forceinline__ double tt_mask(const SubBlk &sb) {
  const int c = threadIdx.x >> 4, p = threadIdx.x & 15;
  return (c < sb.kb && p <= c) ? 1.0 : 0.0;
}

// LARFB: the staged columns are raw GEQRT tile columns (R above the
// diagonal).  Make sub-block sb unit lower trapezoidal in LDS: rows
// r <= i0 + c of column c become (r == i0 + c).  Runs between two barriers.
template <int NB>
__device__ __forceinline__ void fix_unit(double *Vb, const SubBlk &sb) {
  constexpr int LDV = Smem<NB>::LDV;
  const int R = min(sb.i0 + 16, NB);
  for (int e = threadIdx.x; e < sb.kb * R; e += HQ_THREADS) {
    const int c = e / R, r = e - c * R;
    if (r <= sb.i0 + c) Vb[c * LDV + r] = (r == sb.i0 + c) ? 1.0 : 0.0;
  }
}

// Look-ahead of the scheduler during the last pass of a slab task: wave 0,
// in wave-uniform code, claims the next task (phase 0), reads its dependency
// list (a few phases later) and its dependency flags (later still), each
// behind MFMAs that do not wait for them.  The scheduler then takes the
// claimed task and skips its wait when every flag was already complete; a
// flag seen incomplete (including one of the task just finished) sends it
// to the normal wait.  Flags only move to the current epoch, so "complete"
// is never wrong.
// Field by field: an aggregate copy of a descriptor goes through scratch.
__device__ __forceinline__ void copy_task(hq_task &d, const hq_task &s) {
  d.kind = s.kind;
  d.k = s.k;
  d.r = s.r;
  d.j = s.j;
  d.slab = s.slab;
  d.mat = s.mat;
  d.ndep = s.ndep;
#pragma unroll
  for (int i = 0; i < HQ_MAXDEP; i++) d.dep[i] = s.dep[i];
}

struct Ahead {  // passed by value; counter == nullptr: no look-ahead
  int *counter;
  const hq_task *tasks;
  const int *flags;
  int ntasks, epoch;
};
// Stage p of the look-ahead (wave 0 only); the state lives in the caller's
// registers.
template <int NP>
__device__ __forceinline__ void ahead_stage(const Ahead &ah, int p, int &v, int &t, hq_task &tk,
                                            int (&f)[HQ_MAXDEP], int &stop) {
  constexpr int P_DEPS = NP / 2 - 1, P_FLAGS = NP - 2;
  static_assert(P_DEPS > 0 && P_FLAGS > P_DEPS, "stages in distinct phases after the claim");
  if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) != 0) return;
  const int lane = threadIdx.x & 63;
  // one lane claims; its result is read a few phases later, not waited for here
  if (p == 0 && lane == 0) v = __hip_atomic_fetch_add(ah.counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (p == P_DEPS) {
    t = __builtin_amdgcn_readlane(v, 0);
    tk.ndep = 0;
    if (t < ah.ntasks) copy_task(tk, ah.tasks[t]);
  }
  if (p == P_FLAGS) {
#pragma unroll
    for (int i = 0; i < HQ_MAXDEP; i++)
      f[i] = i < tk.ndep ? __hip_atomic_load(ah.flags + tk.dep[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                         : ah.epoch;
    // the scheduler's time-out word, read with the flags
    stop = __hip_atomic_load(ah.counter + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // acquire after the flag loads (in program order), behind the last phase:
  // if they show the task ready the scheduler starts it without a fence of
  // its own; the pass itself reads nothing more from global memory
  if (p == NP - 1) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Returns the task claimed by the look-ahead (x, -1: none; valid in wave 0,
// its descriptor left in S.next); y bit 0: all its dependency flags were
// seen complete, bit 1: the scheduler's time-out word was set.
template <int NB, bool UNIT>
__device__ __forceinline__ int2 task_update_f(Smem<NB> &S, const double *V, const double *T, double *X,
                                              double *A1, int ib, int slab, Ahead ah = Ahead{}) {
  static_assert(NB % HQ_SLAB == 0, "slabs cover the tile exactly");
  const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4, wave = tid >> 6;
  const int col = slab * HQ_SLAB + 16 * wave + m;
  const bool ok = true;
  double *a1col = UNIT ? nullptr : A1 + (size_t)col * NB;
  const int nsub = sub_count<NB>(ib);
  // sub-blocks s, s+1, s+2, s+3 of the current pass (clamped to the last)
  SubBlk w0 = sub_blk<NB>(0, ib), w1 = w0, w2 = next_sub<NB>(w0, ib), w3 = next_sub<NB>(w2, ib);
  d4 acc[NB / 16];
  d4 W2 = d4{0.0, 0.0, 0.0, 0.0}, Wn[2], a1v = W2, a1o = W2;
  double o[2][8], tto[4];
  double treg;
  // prologue: sub-blocks 0 and 1, T^T of both, the strip
  {
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const SubBlk sj = j ? w2 : w0;
#pragma unroll
      for (int q = 0; q < FDma<NB>::NOPS; q++) dma_op<NB>(q, fvbuf(S, j), V, sj);
      if (sj.kb < 16) zero_pad<NB>(fvbuf(S, j), sj);
      ftbuf(S, j)[(tid >> 4) * LDT16 + (tid & 15)] = tt_raw(T, ib, sj) * -tt_mask(sj);
    }
    // the strip's loads are issued last and stay in flight: the wait covers
    // the direct-to-LDS loads and the LDS stores above (vector memory loads
    // complete in issue order), and the first pass waits for each row group
    // of the strip where it reads it
    __builtin_amdgcn_sched_barrier(0);
    load_strip<NB>(acc, X, col, ok, 0);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(waitcnt_vm_lgkm0(2 * (NB / 16)));
    __syncthreads();
    if (UNIT) {
      fix_unit<NB>(fvbuf(S, 0), w0);
      __syncthreads();
    }
    ld_a<NB, UNIT>(o[0], fvbuf(S, 0), 0, 0, 0);
  }
  // pass s: (U) X -= V_s W2_s, (A) W_{s+1} += V_{s+1}^T X, loads of s+3
  auto pre = [&](int s, PassArgs<NB, UNIT> &P, SubBlk &sw) {
    const SubBlk su = w0;
    sw = w1;
    P.sd = w3;
    P.Vd = fvbuf(S, (s + 3) & 3);
    P.Vg = V;
    if (s + 3 < nsub && P.sd.kb < 16) zero_pad<NB>(P.Vd, P.sd);
    Wn[0] = d4{0.0, 0.0, 0.0, 0.0};
    Wn[1] = Wn[0];
    P.Vu = fvbuf(S, (s + 4) & 3);
    P.Vw = fvbuf(S, (s + 1) & 3);
    P.TTw = ftbuf(S, (s + 1) & 3);
    P.i0u = su.i0; P.kbu = su.kb; P.i0w = sw.i0; P.kbw = sw.kb;
  };
  // In the first phase of pass s, behind its MFMAs: (SSRFB) the A1 rows of
  // the update sub-block, a1 + W2 with W2 = -T^T W from the end of the
  // previous pass, then the A1 rows of the accumulate sub-block; T of s+2.
  int la_v = 0, la_t = -1, la_stop = 0, la_f[HQ_MAXDEP];
  hq_task la_tk;
  la_tk.ndep = 0;
#pragma unroll
  for (int i = 0; i < HQ_MAXDEP; i++) la_f[i] = 0;
  const bool look = ah.counter != nullptr;
  auto hook = [&](bool upd, bool acc_next, const PassArgs<NB, UNIT> &P, const SubBlk &sw) {
    return [&, upd, acc_next](int p) {
      if (look && !acc_next) ahead_stage<NB / 32>(ah, p, la_v, la_t, la_tk, la_f, la_stop);
      if (p != 0) return;
      if (!UNIT && upd) {
        a1o = a1v + W2;
        if (P.kbu == 16 && (P.i0u & 3) == 0) {
          *gmem(reinterpret_cast<d4 *>(a1col + P.i0u + 4 * g)) = a1o;
        } else {
#pragma unroll
          for (int v = 0; v < 4; v++)
            if (4 * g + v < P.kbu) *gmem(a1col + P.i0u + 4 * g + v) = a1o[v];
        }
      }
      if (!UNIT && acc_next) {
        // rows past a ragged sub-block load a clamped neighbour: their W rows
        // meet only zero rows and columns of T and are never stored
#pragma unroll
        for (int v = 0; v < 4; v++) a1v[v] = *gmem(a1col + min(sw.i0 + 4 * g + v, NB - 1));
      }
      if (acc_next) treg = tt_raw(T, ib, w2);
    };
  };
  // after pass s: W2_{s+1} = -T^T W (the T ring holds -T^T), T^T of s+2 into
  // the ring
  auto post = [&](int s, const PassArgs<NB, UNIT> &P, const SubBlk &sw) {
    const d4 W = UNIT ? Wn[0] + Wn[1] : (Wn[0] + Wn[1]) + a1v;
    d4 W2n = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int sp = 0; sp < 4; sp++) W2n = mfma(tto[sp], W[sp], W2n);
    W2 = W2n;
    ld_u<NB, UNIT>(o[0], P.Vw, 0, sw.i0, sw.kb);
    ftbuf(S, (s + 2) & 3)[(tid >> 4) * LDT16 + (tid & 15)] = treg * -tt_mask(w2);
    // Every load of the pass has landed, sub-block s+3 included (issued in
    // the first half of the pass); direct-to-LDS loads are not taken to
    // complete in order with the others, so no counted wait here.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (UNIT && s + 2 < nsub) {
      fix_unit<NB>(fvbuf(S, (s + 2) & 3), w2);
      __syncthreads();
    }
    // advance the sub-block window
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = next_sub<NB>(w3, ib);
  };
  {
    PassArgs<NB, UNIT> P;
    SubBlk sw;
    pre(-1, P, sw);
    fpass<NB, UNIT, false, true, true, false>(acc, W2, Wn, o, tto, P, nullptr, hook(false, true, P, sw));
    post(-1, P, sw);
  }
  for (int s = 0; s + 1 < nsub; s++) {
    PassArgs<NB, UNIT> P;
    SubBlk sw;
    pre(s, P, sw);
    fpass<NB, UNIT, true, true, true, false>(acc, W2, Wn, o, tto, P, nullptr, hook(true, true, P, sw));
    post(s, P, sw);
  }
  {
    PassArgs<NB, UNIT> P;
    SubBlk sw;
    pre(nsub - 1, P, sw);
    // the strip is stored row group by row group during the pass
    fpass<NB, UNIT, true, false, false, true>(acc, W2, Wn, o, tto, P, X + (size_t)col * NB,
                                              hook(true, false, P, sw));
  }
  // no direct-to-LDS load may still be in flight when the next task reuses LDS
  __builtin_amdgcn_s_waitcnt(0);
  bool ready = look;
#pragma unroll
  for (int i = 0; i < HQ_MAXDEP; i++) ready = ready && la_f[i] == ah.epoch;
  // the claimed task's descriptor for every thread of the next iteration
  if (look && __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0 && la_t < ah.ntasks) copy_task(S.next, la_tk);
  return make_int2(look ? la_t : -1, (ready ? 1 : 0) | (la_stop != 0 ? 2 : 0));
}

// --------------------------------------------------------------------------
// LARFB pipeline with W_s = V_s T_s precomputed for the diagonal tile (T_s
// the 16x16 diagonal block of T for sub-block s; ib a multiple of 16, so
// every sub-block is 16 wide, s = 0 .. NB/16-1).  Since T_s^T V_s^T X =
// W_s^T X, the accumulate phases of pass s produce T_{s+1}^T V_{s+1}^T X
// directly and the end of a pass is one add and a negation instead of a
// chain of four dependent MFMAs.  V is the unit lower trapezoidal copy the
// GEQRT panels write, so no fix-up pass in LDS either.
// Ring: V_s in slot s&1, W_s in slot 2+(s&1); pass t loads V_{t+1} and
// W_{t+2} (direct-to-LDS, one op per phase).
// (SSRFB keeps the T multiplication: there, streaming W beside V doubles the
// tile traffic of every slab task, which costs more than the chain.)
// --------------------------------------------------------------------------
template <int NB>
struct PassW {
  const double *Vu, *Vw;
  double *D1, *D2;        // ring slots receiving this pass's loads
  const double *G1, *G2;  // their source tiles
  SubBlk s1, s2;
};
__device__ __forceinline__ SubBlk sub16(int s) {
  SubBlk r;
  r.i0 = 16 * s;
  r.kb = 16;
  r.p0 = 0;
  return r;
}

template <int NB, bool DO_U, bool DO_A, int NDMA>
__device__ __forceinline__ void fpass_w(d4 (&acc)[NB / 16], const d4 &W2, d4 (&Wn)[2], double (&o)[2][8],
                                        const PassW<NB> &P) {
  constexpr int NG = NB / 32;
  constexpr int NP = (DO_U ? NG : 0) + (DO_A ? NG : 0);
  constexpr int NOPS = FDma<NB>::NOPS;
  static_assert(NP >= NDMA * NOPS, "one direct-to-LDS op per phase");
  auto kind = [](int p) -> int {  // 1 = update, 0 = accumulate
    if (DO_U && !DO_A) return 1;
    if (!DO_U) return 0;
    if (p == 0) return 1;
    if (p == 2 * NG - 1) return 0;
    return (p & 1) ? 1 : 0;
  };
  auto grp = [](int p) -> int {
    if (!(DO_U && DO_A)) return p;
    if (p == 0) return 0;
    if (p == 2 * NG - 1) return NG - 1;
    return (p & 1) ? (p + 1) / 2 : p / 2 - 1;
  };
#pragma unroll
  for (int p = 0; p < NP; p++) {
    if (NDMA >= 1 && p < NOPS)
      dma_op<NB>(p, P.D1, P.G1, P.s1);
    else if (NDMA >= 2 && p < 2 * NOPS)
      dma_op<NB>(p - NOPS, P.D2, P.G2, P.s2);
    if (p + 1 < NP) {
      if (kind(p + 1))
        ld_u<NB, true>(o[(p + 1) & 1], P.Vu, grp(p + 1), 0, 16);
      else
        ld_a<NB, true>(o[(p + 1) & 1], P.Vw, grp(p + 1), 0, 16);
    }
    const int G = grp(p);
    if (kind(p)) {
#pragma unroll
      for (int sp = 0; sp < 4; sp++)
#pragma unroll
        for (int i = 0; i < 2; i++) acc[2 * G + i] = mfma(o[p & 1][2 * sp + i], W2[sp], acc[2 * G + i]);
    } else {
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int sp = 0; sp < 4; sp++) Wn[sp & 1] = mfma(o[p & 1][4 * i + sp], acc[2 * G + i][sp], Wn[sp & 1]);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 6, 0);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// LARFB (X <- Q^T X) of one 64-column slab: V unit lower trapezoidal, W =
// V T16 (both NB x NB, ld = NB).
template <int NB>
__device__ __forceinline__ void task_larfb_w(Smem<NB> &S, const double *V, const double *Wt, double *X, int slab) {
  static_assert(NB % HQ_SLAB == 0 && NB % 32 == 0, "slabs and row-block groups cover the tile");
  constexpr int NSUB = NB / 16;
  const int lane = threadIdx.x & 63, m = lane & 15, wave = threadIdx.x >> 6;
  const int col = slab * HQ_SLAB + 16 * wave + m;
  d4 acc[NB / 16];
  const d4 zero = d4{0.0, 0.0, 0.0, 0.0};
  d4 W2 = zero, Wn[2];
  double o[2][8];
  // prologue: W_0, V_0, the strip
#pragma unroll
  for (int q = 0; q < FDma<NB>::NOPS; q++) dma_op<NB>(q, fvbuf(S, 2), Wt, sub16(0));
#pragma unroll
  for (int q = 0; q < FDma<NB>::NOPS; q++) dma_op<NB>(q, fvbuf(S, 0), V, sub16(0));
  load_strip<NB>(acc, X, col, true, 0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  Wn[0] = zero;
  Wn[1] = zero;
  ld_a<NB, true>(o[0], fvbuf(S, 2), 0, 0, 16);
  // after pass t-1: W2_t = -W_t^T X
  auto post = [&](int t) {
    W2 = -(Wn[0] + Wn[1]);
    // every load of the pass has landed (direct-to-LDS loads are not taken
    // to complete in order, so no counted wait); V_t came in during this
    // pass, so its first operands are read only after the barrier
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    ld_u<NB, true>(o[0], fvbuf(S, t & 1), 0, 0, 16);
    Wn[0] = zero;
    Wn[1] = zero;
  };
  {  // pass -1: W_0^T X; loads W_1
    PassW<NB> P;
    P.Vu = fvbuf(S, 0);
    P.Vw = fvbuf(S, 2);
    P.D1 = fvbuf(S, 3);
    P.G1 = Wt;
    P.s1 = sub16(1);
    P.D2 = P.D1;
    P.G2 = P.G1;
    P.s2 = P.s1;
    fpass_w<NB, false, true, 1>(acc, W2, Wn, o, P);
    post(0);
  }
  for (int t = 0; t + 1 < NSUB; t++) {
    PassW<NB> P;
    P.Vu = fvbuf(S, t & 1);
    P.Vw = fvbuf(S, 2 + ((t + 1) & 1));
    P.D1 = fvbuf(S, (t + 1) & 1);
    P.G1 = V;
    P.s1 = sub16(t + 1);
    P.D2 = fvbuf(S, 2 + (t & 1));
    P.G2 = Wt;
    P.s2 = sub16(min(t + 2, NSUB - 1));
    fpass_w<NB, true, true, 2>(acc, W2, Wn, o, P);
    post(t + 1);
  }
  {  // last pass: update only
    PassW<NB> P;
    P.Vu = fvbuf(S, (NSUB - 1) & 1);
    P.Vw = P.Vu;
    P.D1 = nullptr;
    P.G1 = nullptr;
    P.s1 = sub16(0);
    P.D2 = nullptr;
    P.G2 = nullptr;
    P.s2 = P.s1;
    fpass_w<NB, true, false, 0>(acc, W2, Wn, o, P);
  }
  // no direct-to-LDS load may still be in flight when the next task reuses LDS
  __builtin_amdgcn_s_waitcnt(0);
  store_strip<NB>(acc, X, col, true, 0);
}

// --------------------------------------------------------------------------
// FP32 update path (mixed precision): LARFB / SSRFB of the same slab tasks
// with FP32 operands and FP32 accumulation on v_mfma_f32_16x16x4f32; tiles
// stay FP64 in memory (converted on load and on store), the panels that
// produce V and T stay FP64.  Same fused pass structure as task_update_f.
// v_mfma_f32_16x16x4f32 lane layout (gfx950, measured): A(i,k) at lane
// i + 16k, B(k,j) at lane j + 16k, D(i,j) at lane j + 16*(i/4), register
// i%4 -- lane group g holds rows 4g..4g+3 of its column without any row
// permutation.  V and T^T are staged through registers (FP64 load, convert,
// FP32 LDS store at the end of the pass); LARFB's unit lower trapezoid is
// applied in that conversion.
// --------------------------------------------------------------------------
typedef float f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f4 mfma32(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
constexpr int LDV32 = 260;  // FP32 ring slot column stride (16-byte aligned columns)
template <int NB>
__device__ __forceinline__ float *f32vbuf(Smem<NB> &S, int i) {
  return reinterpret_cast<float *>(i < 2 ? S.V : S.V1) + (i & 1) * 16 * LDV32;
}
template <int NB>
__device__ __forceinline__ float *f32tbuf(Smem<NB> &S, int i) {
  return reinterpret_cast<float *>(S.TT) + i * 16 * LDT16;
}

template <int NB>
__device__ __forceinline__ void ld_u32(float (&o)[8], const float *Vb, int G) {
  const int lane = threadIdx.x & 63, g = lane >> 4, m = lane & 15;
#pragma unroll
  for (int sp = 0; sp < 4; sp++)
#pragma unroll
    for (int i = 0; i < 2; i++) o[2 * sp + i] = Vb[(4 * g + sp) * LDV32 + 16 * (2 * G + i) + m];
}
template <int NB>
__device__ __forceinline__ void ld_a32(float (&o)[8], const float *Vb, int G) {
  const int lane = threadIdx.x & 63, g = lane >> 4, m = lane & 15;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int sp = 0; sp < 4; sp++) o[4 * i + sp] = Vb[m * LDV32 + 16 * (2 * G + i) + 4 * g + sp];
}

// V columns of a sub-block held in registers between the global loads (issued
// one per phase, early in the pass) and the FP32 LDS store at its end.
template <int NB>
struct Stage32 {
  static constexpr int NLD = 16 * NB / 2 / HQ_THREADS;  // 16-byte loads per thread
  double2 v[NLD];
  double t;
};
// load q of this thread: column c = tid/16 of the sub-block, rows
// 2*(tid%16) + 32*q .. +1 (one 16-byte load; columns past kb load column kb-1)
template <int NB>
__device__ __forceinline__ void stage32_load(Stage32<NB> &R, int q, const double *V, const SubBlk &sb) {
  const int c = min(threadIdx.x >> 4, sb.kb - 1);
  const int r = 2 * (threadIdx.x & 15) + 32 * q;
  R.v[q] = *reinterpret_cast<const double2 *>(V + (size_t)(sb.i0 + c) * NB + r);
}
template <int NB, bool UNIT>
__device__ __forceinline__ void stage32_store(const Stage32<NB> &R, float *Vb, float *TTb, const SubBlk &sb) {
  const int c = threadIdx.x >> 4;
  const int d = sb.i0 + c;
#pragma unroll
  for (int q = 0; q < Stage32<NB>::NLD; q++) {
    const int r = 2 * (threadIdx.x & 15) + 32 * q;
    double a = R.v[q].x, b = R.v[q].y;
    if (UNIT) {
      a = r > d ? a : (r == d ? 1.0 : 0.0);
      b = r + 1 > d ? b : (r + 1 == d ? 1.0 : 0.0);
    }
    const bool live = c < sb.kb;
    float2 f = make_float2(live ? (float)a : 0.f, live ? (float)b : 0.f);
    *reinterpret_cast<float2 *>(Vb + c * LDV32 + r) = f;
  }
  TTb[(threadIdx.x >> 4) * LDT16 + (threadIdx.x & 15)] = (float)(R.t * tt_mask(sb));
}

template <int NB, int NS, bool DO_U, bool DO_A, bool DO_LD>
__device__ __forceinline__ void fpass32(f4 (&acc)[NS][NB / 16], const f4 (&W2)[NS], f4 (&Wn)[NS][2],
                                        float (&o)[2][8], float (&tto)[4], const float *Vu, const float *Vw,
                                        const float *TTw, Stage32<NB> &R, const double *Vg, const SubBlk &sd) {
  constexpr int NG = NB / 32;
  constexpr int NP = (DO_U ? NG : 0) + (DO_A ? NG : 0);
  static_assert(!DO_LD || NP >= Stage32<NB>::NLD, "one staging load per phase");
  auto kind = [](int p) -> int {
    if (DO_U && !DO_A) return 1;
    if (!DO_U) return 0;
    if (p == 0) return 1;
    if (p == 2 * NG - 1) return 0;
    return (p & 1) ? 1 : 0;
  };
  auto grp = [](int p) -> int {
    if (!(DO_U && DO_A)) return p;
    if (p == 0) return 0;
    if (p == 2 * NG - 1) return NG - 1;
    return (p & 1) ? (p + 1) / 2 : p / 2 - 1;
  };
  const int lane = threadIdx.x & 63, g = lane >> 4, m = lane & 15;
#pragma unroll
  for (int p = 0; p < NP; p++) {
    if (DO_LD && p < Stage32<NB>::NLD) stage32_load<NB>(R, p, Vg, sd);
    if (p + 1 < NP) {
      if (kind(p + 1))
        ld_u32<NB>(o[(p + 1) & 1], Vu, grp(p + 1));
      else
        ld_a32<NB>(o[(p + 1) & 1], Vw, grp(p + 1));
    } else if (DO_A) {
#pragma unroll
      for (int sp = 0; sp < 4; sp++) tto[sp] = TTw[m * LDT16 + 4 * g + sp];
    }
    const int G = grp(p);
    // the operands of a phase serve every strip of the wave: 2 NS
    // independent accumulator chains per phase
    if (kind(p)) {
#pragma unroll
      for (int sp = 0; sp < 4; sp++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int q = 0; q < NS; q++) acc[q][2 * G + i] = mfma32(o[p & 1][2 * sp + i], W2[q][sp], acc[q][2 * G + i]);
    } else {
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int sp = 0; sp < 4; sp++)
#pragma unroll
          for (int q = 0; q < NS; q++)
            Wn[q][sp & 1] = mfma32(o[p & 1][4 * i + sp], acc[q][2 * G + i][sp], Wn[q][sp & 1]);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 8 * NS - 2, 0);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// LARFB (UNIT) / SSRFB of one slab of 64 NS columns: wave w holds the
// 16-column strips w + 4q (q < NS) of the slab.
template <int NB, bool UNIT, int NS>
__device__ __forceinline__ void task_update_f32(Smem<NB> &S, const double *V, const double *T, double *X,
                                                double *A1, int ib, int slab) {
  constexpr int UW = 64 * NS;
  static_assert(NB % UW == 0, "slabs cover the tile exactly");
  const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4, wave = tid >> 6;
  // strip q: column col0 + 64 q
  const int col0 = slab * UW + 16 * wave + m;
  double *xcol = X + (size_t)col0 * NB;
  double *a1col = UNIT ? nullptr : A1 + (size_t)col0 * NB;
  constexpr size_t QS = (size_t)64 * NB;  // strip stride in doubles
  const int nsub = sub_count<NB>(ib);
  SubBlk w0 = sub_blk<NB>(0, ib), w1 = w0, w2 = next_sub<NB>(w0, ib), w3 = next_sub<NB>(w2, ib);
  f4 acc[NS][NB / 16];
  f4 W2[NS], Wn[NS][2], a1v[NS], a1o[NS];
#pragma unroll
  for (int q = 0; q < NS; q++) {
    W2[q] = f4{0.f, 0.f, 0.f, 0.f};
    a1v[q] = W2[q];
    a1o[q] = W2[q];
  }
  float o[2][8], tto[4];
  Stage32<NB> R;
  // prologue: sub-blocks 0 and 1 into ring slots 0, 1; the strips in FP32
  {
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const SubBlk sj = j ? w2 : w0;
#pragma unroll
      for (int q = 0; q < Stage32<NB>::NLD; q++) stage32_load<NB>(R, q, V, sj);
      R.t = tt_raw(T, ib, sj);
      stage32_store<NB, UNIT>(R, f32vbuf(S, j), f32tbuf(S, j), sj);
    }
#pragma unroll
    for (int q = 0; q < NS; q++)
#pragma unroll
      for (int rb = 0; rb < NB / 16; rb++) {
        const d4 x = *reinterpret_cast<const d4 *>(xcol + q * QS + 16 * rb + 4 * g);
        acc[q][rb] = f4{(float)x[0], (float)x[1], (float)x[2], (float)x[3]};
      }
    __syncthreads();
    ld_a32<NB>(o[0], f32vbuf(S, 0), 0);
  }
  // A1 rows of the accumulate sub-block, loaded before a pass as doubles and
  // converted after it: a conversion next to the load would make the wave
  // wait for each load in turn at the start of the pass
  d4 a1raw[NS];
  auto pre = [&](int s) {
    R.t = tt_raw(T, ib, w2);
    if (!UNIT) {
#pragma unroll
      for (int q = 0; q < NS; q++)
#pragma unroll
        for (int v = 0; v < 4; v++) a1raw[q][v] = a1col[q * QS + min(w1.i0 + 4 * g + v, NB - 1)];
    }
#pragma unroll
    for (int q = 0; q < NS; q++) {
      Wn[q][0] = f4{0.f, 0.f, 0.f, 0.f};
      Wn[q][1] = Wn[q][0];
    }
  };
  // after pass s: W2_{s+1}; sub-block s+2 into its ring slot; A1 rows of
  // s+1 stored after the barrier
  auto post = [&](int s) {
    f4 W[NS], W2n[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) {
      if (!UNIT) a1v[q] = f4{(float)a1raw[q][0], (float)a1raw[q][1], (float)a1raw[q][2], (float)a1raw[q][3]};
      W[q] = UNIT ? Wn[q][0] + Wn[q][1] : (Wn[q][0] + Wn[q][1]) + a1v[q];
      W2n[q] = f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int q = 0; q < NS; q++) W2n[q] = mfma32(tto[sp], W[q][sp], W2n[q]);
#pragma unroll
    for (int q = 0; q < NS; q++) {
      a1o[q] = a1v[q] - W2n[q];
      W2[q] = -W2n[q];
    }
    ld_u32<NB>(o[0], f32vbuf(S, (s + 1) & 3), 0);
    stage32_store<NB, UNIT>(R, f32vbuf(S, (s + 2) & 3), f32tbuf(S, (s + 2) & 3), w2);
    __syncthreads();
    if (!UNIT) {
#pragma unroll
      for (int q = 0; q < NS; q++)
#pragma unroll
        for (int v = 0; v < 4; v++)
          if (4 * g + v < w1.kb) a1col[q * QS + w1.i0 + 4 * g + v] = (double)a1o[q][v];
    }
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = next_sub<NB>(w3, ib);
  };
  pre(-1);
  fpass32<NB, NS, false, true, true>(acc, W2, Wn, o, tto, f32vbuf(S, 3), f32vbuf(S, 0), f32tbuf(S, 0), R, V, w2);
  post(-1);
  for (int s = 0; s + 1 < nsub; s++) {
    pre(s);
    fpass32<NB, NS, true, true, true>(acc, W2, Wn, o, tto, f32vbuf(S, s & 3), f32vbuf(S, (s + 1) & 3),
                                      f32tbuf(S, (s + 1) & 3), R, V, w2);
    post(s);
  }
  fpass32<NB, NS, true, false, false>(acc, W2, Wn, o, tto, f32vbuf(S, (nsub - 1) & 3), f32vbuf(S, 0),
                                      f32tbuf(S, 0), R, V, w2);
#pragma unroll
  for (int q = 0; q < NS; q++)
#pragma unroll
    for (int rb = 0; rb < NB / 16; rb++) {
      const f4 x = acc[q][rb];
      *reinterpret_cast<d4 *>(xcol + q * QS + 16 * rb + 4 * g) =
          d4{(double)x[0], (double)x[1], (double)x[2], (double)x[3]};
    }
}

// --------------------------------------------------------------------------
// Split-precision update path (prec 2): LARFB / SSRFB of the same slab tasks
// on FP16 MFMA with FP32 accumulation, every product taken as
// hi*hi + hi*lo + lo*hi of FP16 hi + lo splits (x = hi + lo, hi = fp16(x),
// lo = fp16(x - hi): 22 significant bits, the lo*lo term of 2^-22 relative
// size dropped).  The panels that produce V and T stay FP64.
//  * Scaling by powers of two (exact): V by 2^10 (Householder vectors have
//    |v_i| <= 1, so every entry of V and of its lo part stays inside FP16's
//    range); the strip by 2^e per column with e chosen from the column's
//    largest entry over [A1; A2] so that its 2-norm stays below 2^12 -- the
//    transform is orthogonal, so no intermediate column grows past it.  The lo
//    parts of small entries go subnormal at an absolute level of 2^-25, i.e.
//    2^-37 of the column norm.
//  * The strip lives in FP32 in MFMA accumulator layout; W = V^T X (K = 32
//    rows per v_mfma_f32_16x16x32_f16: the two row blocks a lane group holds
//    are its 8 K slots, so the strip's registers split straight into the B
//    operand), W2 = -T^T W on v_mfma_f32_16x16x4f32 (T in FP32), and
//    X += V (-W2) with K = 16 reflectors on v_mfma_f32_16x16x16f16 (W2's
//    accumulator layout is that instruction's B operand).
//  * v_mfma_f32_16x16x{16,32}_f16 lane layouts (gfx950, measured): A(i,k) at
//    lane i + 16*(k/E), element k%E, B(k,j) at lane j + 16*(k/E), element k%E
//    (E = K/4), D(i,j) at lane j + 16*(i/4), register i%4.
//  * LDS ring slot of a 16-reflector sub-block: V^T (reflector-major, rows
//    contiguous, the A phase's operands) and V (row-major, 16 reflectors
//    contiguous, the update phase's), hi and lo of each, as FP16.
// --------------------------------------------------------------------------
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f4 mfma_h16(h4 a, h4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 mfma_h32(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
constexpr float HV_SCALE = 1024.0f;          // V is staged as 2^10 V
constexpr float HV_UNSCALE = 1.0f / 1024.0f;
__device__ __forceinline__ void hsplit(float x, _Float16 &h, _Float16 &l) {
  h = (_Float16)x;
  l = (_Float16)(x - (float)h);
}
// A sub-block image (and LDS ring slot): 4 arrays of 16 x NB halves,
//   AH / AL: V^T hi / lo, reflector-major; element (i, row r) at
//            i*NB + 4*((r >> 2) ^ i) + (r & 3) (4-row chunks swizzled by
//            the reflector index, so the 16 reflectors of an accumulate
//            operand read fall on different LDS banks);
//   UH / UL: V hi / lo, row-major, element (r, i) at
//            r*16 + 4*((i >> 2) ^ ((r >> 2) & 3)) + (i & 3) (the 4-reflector
//            chunks of a row swizzled by the row, so that the 16 rows of an
//            update operand read fall on different LDS banks).
// 128 NB bytes, a whole number of 1 KiB direct-to-LDS wave copies.  (The
// update operands could also be read from AH / AL with the LDS transpose
// load, halving the image; but the compiler makes every such read wait for
// all direct-to-LDS copies in flight, which the pipelined update cannot
// afford.)
template <int NB>
struct HSlot {
  static constexpr int BYTES = 128 * NB;
  static constexpr int HALVES = 64 * NB;                                   // image size in halves
  static constexpr int AH = 0, AL = 16 * NB, UH = 32 * NB, UL = 48 * NB;  // offsets in halves
  static constexpr int NOPS = BYTES / 1024 / (HQ_THREADS / 64);            // copies per wave
};
__device__ __forceinline__ int hswz(int i, int r) { return 4 * ((r >> 2) ^ i) + (r & 3); }
__device__ __forceinline__ int uswz(int r, int q) { return r * 16 + 4 * (q ^ ((r >> 2) & 3)); }  // chunk q of row r
// Ring slot i (0..3): two slots in S.V, two in S.V1.
template <int NB>
__device__ __forceinline__ _Float16 *hvbuf(Smem<NB> &S, int i) {
  static_assert(2 * HSlot<NB>::BYTES <= (int)sizeof(S.V), "two split slots per V buffer");
  return reinterpret_cast<_Float16 *>(reinterpret_cast<unsigned char *>(i < 2 ? S.V : S.V1) +
                                      (i & 1) * HSlot<NB>::BYTES);
}

// Panel side: the images of the sub-blocks of the ib block at column ii
// (V in S.V, column-major, unit lower for GEQRT, V2 for TSQRT, zero past the
// block) into the tile's image area VHt (sub-block s = b*nsb + h).
template <int NB>
__device__ __forceinline__ void store_vh(Smem<NB> &S, unsigned short *VHt, int ib, int ii, int sb) {
  using L = HSlot<NB>;
  constexpr int LDV = Smem<NB>::LDV;
  const int nsb = (ib + 15) >> 4;
  const int s0 = (ii / ib) * nsb;
  const int nh = (sb + 15) >> 4;
  typedef _Float16 h4t __attribute__((ext_vector_type(4)));
  for (int h = 0; h < nh; h++) {
    const int kb = min(16, sb - 16 * h);
    _Float16 *img = reinterpret_cast<_Float16 *>(VHt) + (size_t)(s0 + h) * L::HALVES;
    // V^T arrays: chunk (i, c) = rows 4c..4c+3 of reflector i
    for (int e = threadIdx.x; e < 16 * (NB / 4); e += HQ_THREADS) {
      const int i = e / (NB / 4), c = e - i * (NB / 4);
      h4t hv, lv;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const float f = i < kb ? (float)S.V[(16 * h + i) * LDV + 4 * c + t] * HV_SCALE : 0.f;
        _Float16 a, b;
        hsplit(f, a, b);
        hv[t] = a;
        lv[t] = b;
      }
      *reinterpret_cast<h4t *>(img + L::AH + i * NB + hswz(i, 4 * c)) = hv;
      *reinterpret_cast<h4t *>(img + L::AL + i * NB + hswz(i, 4 * c)) = lv;
    }
    // V arrays: chunk (r, q) = reflectors 4q..4q+3 of row r
    for (int e = threadIdx.x; e < NB * 4; e += HQ_THREADS) {
      const int r = e >> 2, q = e & 3;
      h4t hv, lv;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int i = 4 * q + t;
        const float f = i < kb ? (float)S.V[(16 * h + i) * LDV + r] * HV_SCALE : 0.f;
        _Float16 a, b;
        hsplit(f, a, b);
        hv[t] = a;
        lv[t] = b;
      }
      *reinterpret_cast<h4t *>(img + L::UH + uswz(r, q)) = hv;
      *reinterpret_cast<h4t *>(img + L::UL + uswz(r, q)) = lv;
    }
  }
}

// Copy q (< NOPS) of this wave of a sub-block image into ring slot Vb.
template <int NB>
__device__ __forceinline__ void dmah_op(int q, _Float16 *Vb, const _Float16 *img) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int Q = wave + (HQ_THREADS / 64) * q;
  const unsigned char *src = reinterpret_cast<const unsigned char *>(img) + 1024 * Q + 16 * lane;
  lds_char *dst = (lds_char *)(reinterpret_cast<unsigned char *>(Vb) + 1024 * Q);
  __builtin_amdgcn_global_load_lds((const void *)src, (void __attribute__((address_space(3))) *)dst, 16, 0, 0);
}

// Operands of one update phase (row blocks 2G, 2G+1): V(16rb + i, 4g..4g+3)
// hi / lo for the lane (i, g).
struct HOpU {
  h4 h[2], l[2];
};
template <int NB>
__device__ __forceinline__ void ldh_u(HOpU &o, const _Float16 *Vb, int G) {
  using L = HSlot<NB>;
  const int lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int row = 16 * (2 * G + t) + i;
    o.h[t] = *reinterpret_cast<const h4 *>(Vb + L::UH + uswz(row, g));
    o.l[t] = *reinterpret_cast<const h4 *>(Vb + L::UL + uswz(row, g));
  }
}
// Operands of one accumulate phase: V^T(i, rows 32G + 4g.. and 32G + 16 + 4g..)
struct HOpA {
  h8 h, l;
};
template <int NB>
__device__ __forceinline__ void ldh_a(HOpA &o, const _Float16 *Vb, int G) {
  using L = HSlot<NB>;
  const int lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
  const int r0 = 32 * G + 4 * g;
  const h4 h0 = *reinterpret_cast<const h4 *>(Vb + L::AH + i * NB + hswz(i, r0));
  const h4 h1 = *reinterpret_cast<const h4 *>(Vb + L::AH + i * NB + hswz(i, r0 + 16));
  const h4 l0 = *reinterpret_cast<const h4 *>(Vb + L::AL + i * NB + hswz(i, r0));
  const h4 l1 = *reinterpret_cast<const h4 *>(Vb + L::AL + i * NB + hswz(i, r0 + 16));
  o.h = h8{h0[0], h0[1], h0[2], h0[3], h1[0], h1[1], h1[2], h1[3]};
  o.l = h8{l0[0], l0[1], l0[2], l0[3], l1[0], l1[1], l1[2], l1[3]};
}

// One pass of the split pipeline: update phases X += V_u (-W2) and
// accumulate phases W += V_w^T X interleaved as U0 U1 A0 U2 A1 ... A(NG-1)
// (or only U / only A); the copies of the next sub-block image, one per
// phase.
template <int NB, int NS, bool DO_U, bool DO_A, bool DO_DMA>
__device__ __forceinline__ void fpassh(f4 (&acc)[NS][NB / 16], const h4 (&W2h)[NS], const h4 (&W2l)[NS],
                                       f4 (&Wn)[NS][2], float (&tto)[4], const _Float16 *Vu, const _Float16 *Vw,
                                       const float *TTw, _Float16 *Vd, const _Float16 *img) {
  constexpr int NG = NB / 32;
  constexpr int NP = (DO_U ? NG : 0) + (DO_A ? NG : 0);
  static_assert(!DO_DMA || NP >= HSlot<NB>::NOPS, "one image copy per phase");
  auto kind = [](int p) -> int {
    if (DO_U && !DO_A) return 1;
    if (!DO_U) return 0;
    if (p == 0) return 1;
    if (p == 2 * NG - 1) return 0;
    return (p & 1) ? 1 : 0;
  };
  auto grp = [](int p) -> int {
    if (!(DO_U && DO_A)) return p;
    if (p == 0) return 0;
    if (p == 2 * NG - 1) return NG - 1;
    return (p & 1) ? (p + 1) / 2 : p / 2 - 1;
  };
  const int lane = threadIdx.x & 63, g = lane >> 4, m = lane & 15;
  // operands one phase ahead (the first phase's are read here); no
  // instruction crosses a phase boundary, which bounds the live registers
  HOpU ou[2];
  HOpA oa[2];
  if (kind(0))
    ldh_u<NB>(ou[0], Vu, grp(0));
  else
    ldh_a<NB>(oa[0], Vw, grp(0));
#pragma unroll
  for (int p = 0; p < NP; p++) {
    if (DO_DMA && p < HSlot<NB>::NOPS) dmah_op<NB>(p, Vd, img);
    if (p + 1 < NP) {
      if (kind(p + 1))
        ldh_u<NB>(ou[(p + 1) & 1], Vu, grp(p + 1));
      else
        ldh_a<NB>(oa[(p + 1) & 1], Vw, grp(p + 1));
    } else if (DO_A) {
#pragma unroll
      for (int sp = 0; sp < 4; sp++) tto[sp] = TTw[m * LDT16 + 4 * g + sp];
    }
    const int G = grp(p);
    // the operands of a phase serve every strip of the wave; each product
    // term runs over all 2 NS accumulators before the next (independent
    // chains back to back)
    if (kind(p)) {
      const HOpU &o = ou[p & 1];
#pragma unroll
      for (int t = 0; t < 2; t++)
#pragma unroll
        for (int q = 0; q < NS; q++) acc[q][2 * G + t] = mfma_h16(o.h[t], W2h[q], acc[q][2 * G + t]);
#pragma unroll
      for (int t = 0; t < 2; t++)
#pragma unroll
        for (int q = 0; q < NS; q++) acc[q][2 * G + t] = mfma_h16(o.h[t], W2l[q], acc[q][2 * G + t]);
#pragma unroll
      for (int t = 0; t < 2; t++)
#pragma unroll
        for (int q = 0; q < NS; q++) acc[q][2 * G + t] = mfma_h16(o.l[t], W2h[q], acc[q][2 * G + t]);
    } else {
      const HOpA &o = oa[p & 1];
      // strip by strip: the split of strip q+1 overlaps strip q's MFMAs
#pragma unroll
      for (int q = 0; q < NS; q++) {
        const f4 x0 = acc[q][2 * G], x1 = acc[q][2 * G + 1];
        h8 xh, xl;
#pragma unroll
        for (int e = 0; e < 4; e++) {
          _Float16 h, l;
          hsplit(x0[e], h, l);
          xh[e] = h;
          xl[e] = l;
          hsplit(x1[e], h, l);
          xh[4 + e] = h;
          xl[4 + e] = l;
        }
        Wn[q][0] = mfma_h32(o.h, xh, Wn[q][0]);
        Wn[q][1] = mfma_h32(o.h, xl, Wn[q][1]);
        Wn[q][0] = mfma_h32(o.l, xh, Wn[q][0]);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// LARFB (UNIT: the images of the diagonal tile's unit lower V) / SSRFB on
// one slab of 64 NS columns in split precision (wave w holds the 16-column
// strips w + 4q, q < NS).  VHt: the reflector tile's images.
template <int NB, bool UNIT, int NS>
__device__ __forceinline__ void task_update_h(Smem<NB> &S, const unsigned short *VHt, const double *T, double *X,
                                              double *A1, int ib, int slab, const int *CE) {
  constexpr int UW = 64 * NS;
  static_assert(NB % UW == 0, "slabs cover the tile exactly");
  const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4, wave = tid >> 6;
  // strip q: column col0 + 64 q
  const int col0 = slab * UW + 16 * wave + m;
  double *xcol = X + (size_t)col0 * NB;
  double *a1col = UNIT ? nullptr : A1 + (size_t)col0 * NB;
  constexpr size_t QS = (size_t)64 * NB;  // strip stride in doubles
  const int nsub = sub_count<NB>(ib);
  const _Float16 *img0 = reinterpret_cast<const _Float16 *>(VHt);
  auto sbk = [&](int s) __attribute__((always_inline)) { return sub_blk<NB>(min(s, nsub - 1), ib); };
  auto img = [&](int s) __attribute__((always_inline)) { return img0 + (size_t)min(s, nsub - 1) * HSlot<NB>::HALVES; };
  f4 acc[NS][NB / 16];
  f4 Wn[NS][2], a1v[NS];  // a1v: the A1 rows read before a pass, their new values after it
  h4 W2h[NS], W2l[NS];
#pragma unroll
  for (int q = 0; q < NS; q++) {
    a1v[q] = f4{0.f, 0.f, 0.f, 0.f};
    W2h[q] = h4{0, 0, 0, 0};
    W2l[q] = W2h[q];
  }
  float tto[4];
  double treg = 0.0;
  float xs[NS], xu[NS];  // each strip column's scale 2^e and 2^-e
  // prologue: the images of sub-blocks 0 and 1, T^T of both, the strips
  // (two sweeps over each column, the second from cache: its largest entry
  // over [A1; A2], then the scaled FP32 copy)
  {
#pragma unroll
    for (int j = 0; j < 2; j++) {
#pragma unroll
      for (int q = 0; q < HSlot<NB>::NOPS; q++) dmah_op<NB>(q, hvbuf(S, j), img(j));
      f32tbuf(S, j)[(tid >> 4) * LDT16 + (tid & 15)] = (float)(tt_raw(T, ib, sbk(j)) * tt_mask(sbk(j)));
    }
#pragma unroll
    for (int q = 0; q < NS; q++) {
      if (CE) {
        // the column's scale from its 2-norm: one sweep over the strip
        const int sh = CE[col0 + q * 64];
        const double dsc = ldexp(1.0, sh);
        xs[q] = (float)dsc;
        xu[q] = (float)ldexp(1.0, -sh);
#pragma unroll
        for (int rb = 0; rb < NB / 16; rb++) {
          const d4 x = *reinterpret_cast<const d4 *>(xcol + q * QS + 16 * rb + 4 * g);
          acc[q][rb] = f4{(float)(x[0] * dsc), (float)(x[1] * dsc), (float)(x[2] * dsc), (float)(x[3] * dsc)};
        }
        continue;
      }
      double mx = 0.0;
#pragma unroll 4
      for (int rb = 0; rb < NB / 16; rb++) {
        const d4 x = *reinterpret_cast<const d4 *>(xcol + q * QS + 16 * rb + 4 * g);
#pragma unroll
        for (int v = 0; v < 4; v++) mx = fmax(mx, fabs(x[v]));
      }
      if (!UNIT) {
#pragma unroll 4
        for (int rb = 0; rb < NB / 16; rb++) {
          const d4 a = *reinterpret_cast<const d4 *>(a1col + q * QS + 16 * rb + 4 * g);
#pragma unroll
          for (int v = 0; v < 4; v++) mx = fmax(mx, fabs(a[v]));
        }
      }
      mx = fmax(mx, __shfl_xor(mx, 16, 64));
      mx = fmax(mx, __shfl_xor(mx, 32, 64));
      // the column's largest entry is < 2^e, so its 2-norm over the 2 NB rows
      // of [A1; A2] (NB <= 256) is < 2^(e+5); the scaled norm is < 2^12
      int e = 0;
      if (mx > 0.0 && mx <= 1.7e308) (void)frexp(mx, &e);
      const int sh = mx > 0.0 ? 7 - e : 0;
      const double dsc = ldexp(1.0, sh);
      xs[q] = (float)dsc;
      xu[q] = (float)ldexp(1.0, -sh);
#pragma unroll
      for (int rb = 0; rb < NB / 16; rb++) {
        const d4 x = *reinterpret_cast<const d4 *>(xcol + q * QS + 16 * rb + 4 * g);
        acc[q][rb] = f4{(float)(x[0] * dsc), (float)(x[1] * dsc), (float)(x[2] * dsc), (float)(x[3] * dsc)};
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  // before pass s: A1 rows of the accumulate sub-block s+1 (as doubles,
  // converted after the pass: a conversion next to the load would make the
  // wave wait for each load in turn), T of s+2
  d4 a1raw[NS];
  auto pre = [&](int s) __attribute__((always_inline)) {
    if (!UNIT) {
      // rows past a ragged sub-block load a clamped neighbour: their W rows
      // meet only zero rows and columns of T and are never stored
      const SubBlk w1 = sbk(s + 1);
#pragma unroll
      for (int q = 0; q < NS; q++)
#pragma unroll
        for (int v = 0; v < 4; v++) a1raw[q][v] = a1col[q * QS + min(w1.i0 + 4 * g + v, NB - 1)];
    }
    treg = tt_raw(T, ib, sbk(s + 2));
#pragma unroll
    for (int q = 0; q < NS; q++) {
      Wn[q][0] = f4{0.f, 0.f, 0.f, 0.f};
      Wn[q][1] = Wn[q][0];
    }
  };
  // after pass s: W2 of sub-block s+1 (FP32 T chain), its FP16 split for the
  // next update phases, T^T of s+2 into the ring, A1 rows of s+1 after the
  // barrier (every copy of the pass has landed before it)
  auto post = [&](int s) __attribute__((always_inline)) {
    f4 W[NS], t[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) {
      W[q] = (Wn[q][0] + Wn[q][1]) * HV_UNSCALE;
      if (!UNIT) {
#pragma unroll
        for (int v = 0; v < 4; v++) a1v[q][v] = (float)a1raw[q][v] * xs[q];
        W[q] += a1v[q];
      }
      t[q] = f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int q = 0; q < NS; q++) t[q] = mfma32(tto[sp], W[q][sp], t[q]);
#pragma unroll
    for (int q = 0; q < NS; q++) {
      if (!UNIT) a1v[q] -= t[q];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        _Float16 h, l;
        hsplit(-t[q][v] * HV_UNSCALE, h, l);
        W2h[q][v] = h;
        W2l[q][v] = l;
      }
    }
    f32tbuf(S, (s + 2) & 3)[(tid >> 4) * LDT16 + (tid & 15)] = (float)(treg * tt_mask(sbk(s + 2)));
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (!UNIT) {
      const SubBlk w1 = sbk(s + 1);
#pragma unroll
      for (int q = 0; q < NS; q++)
#pragma unroll
        for (int v = 0; v < 4; v++)
          if (4 * g + v < w1.kb) a1col[q * QS + w1.i0 + 4 * g + v] = (double)a1v[q][v] * (double)xu[q];
    }
  };
  // pass -1: W of sub-block 0; copies sub-block 2 into slot 2
  pre(-1);
  fpassh<NB, NS, false, true, true>(acc, W2h, W2l, Wn, tto, hvbuf(S, 3), hvbuf(S, 0), f32tbuf(S, 0), hvbuf(S, 2),
                                    img(2));
  post(-1);
  // pass s: update with s (slot s&3), accumulate with s+1, copy s+3 into
  // slot (s+3)&3 (read last by pass s-1)
  for (int s = 0; s + 1 < nsub; s++) {
    pre(s);
    fpassh<NB, NS, true, true, true>(acc, W2h, W2l, Wn, tto, hvbuf(S, s & 3), hvbuf(S, (s + 1) & 3),
                                     f32tbuf(S, (s + 1) & 3), hvbuf(S, (s + 3) & 3), img(s + 3));
    post(s);
  }
  fpassh<NB, NS, true, false, false>(acc, W2h, W2l, Wn, tto, hvbuf(S, (nsub - 1) & 3), hvbuf(S, 0), f32tbuf(S, 0),
                                     hvbuf(S, 0), img(0));
#pragma unroll
  for (int q = 0; q < NS; q++)
#pragma unroll
    for (int rb = 0; rb < NB / 16; rb++) {
      const f4 x = acc[q][rb];
      const double u = (double)xu[q];
      *reinterpret_cast<d4 *>(xcol + q * QS + 16 * rb + 4 * g) =
          d4{(double)x[0] * u, (double)x[1] * u, (double)x[2] * u, (double)x[3] * u};
    }
  // no copy may still be in flight when the next task reuses LDS
  __builtin_amdgcn_s_waitcnt(0);
}

// Split-precision LARFB / SSRFB pipelined two passes deep (ib a multiple of
// 16, so every sub-block is 16 reflectors at row 16 s; precomputed column
// scales).  Pass p copies to LDS the image of sub-block p+3, its 16 x 16 T
// block and the A1 rows of sub-block p+2; the end of pass p waits only for
// the copies of pass p-1 (s_waitcnt vmcnt(N), N = this wave's copies of pass
// p: loads retire in issue order, and stores in between only make the wait
// longer) and synchronises with a barrier that leaves the copies of pass p in
// flight.  No register load is in flight during the passes: the compiler's
// own wait for one would wait for every copy.  Staging: the A1 rows of
// sub-block j in buffer j % 2 (S.TT1, S.XA; each wave copies and reads its
// own 16 columns), the T block of j in buffer j % 2 (S.TT past the FP32 T
// ring; wave w copies columns 8 (w / 2) .. +7, its own 4 and its partner's
// -- the same bytes -- and converts its own after its own wait).
// Workgroup barrier that leaves vector memory operations in flight: no
// memory-model fence (__syncthreads()'s fences, even LDS-only ones while a
// direct-to-LDS copy is in flight, make every wave wait for all of its
// vector memory operations), only the hardware barrier with this wave's LDS
// operations complete before it and compiler barriers that keep memory
// accesses on their side of it.  The caller waits explicitly, before it,
// for the copies whose LDS data is read after it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vmcnt, expcnt unconstrained
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

template <int NB>
struct H2Stage {
  static constexpr int A1B = 16 * HQ_SLAB;   // doubles per A1 buffer (16 rows of the slab)
  static constexpr int TB = 256;             // doubles per T buffer
  static constexpr int T0 = 4 * 16 * LDT16 / 2;  // doubles of S.TT before the T buffers (the FP32 ring)
  static_assert(A1B <= Smem<NB>::TSZ && A1B <= 16 * HQ_SLAB, "A1 buffers fit");
  static_assert(T0 + 2 * TB <= Smem<NB>::TSZ, "T buffers fit in S.TT");
};
template <int NB>
__device__ __forceinline__ void dmat_op(double *Ts, const double *T, int ib, int s) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = 8 * (wave >> 1) + (lane >> 3);      // column of the T block
  const int i0 = 16 * s, p0 = i0 % ib;
  const double *src = T + (size_t)(i0 + c) * ib + p0 + 2 * (lane & 7);
  lds_char *dst = (lds_char *)(Ts + 128 * (wave >> 1));
  __builtin_amdgcn_global_load_lds((const void *)src, (void __attribute__((address_space(3))) *)dst, 16, 0, 0);
}
// copy q (0, 1) of this wave's A1 rows 16 s .. 16 s + 15 of slab columns
// 16 w + 8 q .. +7 (column c of the slab at 16 c doubles of the buffer)
template <int NB>
__device__ __forceinline__ void dmaa_op(int q, double *Ab, const double *A1slab, int s) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = 16 * wave + 8 * q + (lane >> 3);
  const double *src = A1slab + (size_t)c * NB + 16 * s + 2 * (lane & 7);
  lds_char *dst = (lds_char *)(Ab + 16 * (16 * wave + 8 * q));
  __builtin_amdgcn_global_load_lds((const void *)src, (void __attribute__((address_space(3))) *)dst, 16, 0, 0);
}

// One pass of the two-deep split pipeline: fpassh's phases, the copies
// given by copy(k), k < NC, copy k in phase min(k, NP - 1).
template <int NB, bool DO_U, bool DO_A, int NC, class Copy>
__device__ __forceinline__ void fpassh2(f4 (&acc)[NB / 16], const h4 &W2h, const h4 &W2l, f4 (&Wn)[2],
                                        float (&tto)[4], const _Float16 *Vu, const _Float16 *Vw, const float *TTw,
                                        Copy &&copy) {
  constexpr int NG = NB / 32;
  constexpr int NP = (DO_U ? NG : 0) + (DO_A ? NG : 0);
  auto kind = [](int p) -> int {
    if (DO_U && !DO_A) return 1;
    if (!DO_U) return 0;
    if (p == 0) return 1;
    if (p == 2 * NG - 1) return 0;
    return (p & 1) ? 1 : 0;
  };
  auto grp = [](int p) -> int {
    if (!(DO_U && DO_A)) return p;
    if (p == 0) return 0;
    if (p == 2 * NG - 1) return NG - 1;
    return (p & 1) ? (p + 1) / 2 : p / 2 - 1;
  };
  const int lane = threadIdx.x & 63, g = lane >> 4, m = lane & 15;
  HOpU ou[2];
  HOpA oa[2];
  if (kind(0))
    ldh_u<NB>(ou[0], Vu, grp(0));
  else
    ldh_a<NB>(oa[0], Vw, grp(0));
#pragma unroll
  for (int p = 0; p < NP; p++) {
#pragma unroll
    for (int k = 0; k < NC; k++)
      if ((k < NP - 1 ? k : NP - 1) == p) copy(k);
    if (p + 1 < NP) {
      if (kind(p + 1))
        ldh_u<NB>(ou[(p + 1) & 1], Vu, grp(p + 1));
      else
        ldh_a<NB>(oa[(p + 1) & 1], Vw, grp(p + 1));
    } else if (DO_A) {
#pragma unroll
      for (int sp = 0; sp < 4; sp++) tto[sp] = TTw[m * LDT16 + 4 * g + sp];
    }
    const int G = grp(p);
    if (kind(p)) {
      const HOpU &o = ou[p & 1];
#pragma unroll
      for (int t = 0; t < 2; t++) acc[2 * G + t] = mfma_h16(o.h[t], W2h, acc[2 * G + t]);
#pragma unroll
      for (int t = 0; t < 2; t++) acc[2 * G + t] = mfma_h16(o.h[t], W2l, acc[2 * G + t]);
#pragma unroll
      for (int t = 0; t < 2; t++) acc[2 * G + t] = mfma_h16(o.l[t], W2h, acc[2 * G + t]);
    } else {
      const HOpA &o = oa[p & 1];
      const f4 x0 = acc[2 * G], x1 = acc[2 * G + 1];
      h8 xh, xl;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        _Float16 h, l;
        hsplit(x0[e], h, l);
        xh[e] = h;
        xl[e] = l;
        hsplit(x1[e], h, l);
        xh[4 + e] = h;
        xl[4 + e] = l;
      }
      Wn[0] = mfma_h32(o.h, xh, Wn[0]);
      Wn[1] = mfma_h32(o.h, xl, Wn[1]);
      Wn[0] = mfma_h32(o.l, xh, Wn[0]);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int NB, bool UNIT>
__device__ __forceinline__ void task_update_h2(Smem<NB> &S, const unsigned short *VHt, const double *T, double *X,
                                               double *A1, int ib, int slab, const int *CE) {
  static_assert(NB % HQ_SLAB == 0, "slabs cover the tile exactly");
  using ST = H2Stage<NB>;
  constexpr int NSUB = NB / 16;
  constexpr int NC = HSlot<NB>::NOPS + 1 + (UNIT ? 0 : 2);  // copies per wave and pass
  const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4, wave = tid >> 6;
  const int col = slab * HQ_SLAB + 16 * wave + m;
  double *xcol = X + (size_t)col * NB;
  const double *a1slab = UNIT ? nullptr : A1 + (size_t)slab * HQ_SLAB * NB;
  double *a1col = UNIT ? nullptr : A1 + (size_t)col * NB;
  const _Float16 *img0 = reinterpret_cast<const _Float16 *>(VHt);
  auto img = [&](int s) __attribute__((always_inline)) { return img0 + (size_t)min(s, NSUB - 1) * HSlot<NB>::HALVES; };
  auto abuf = [&](int s) __attribute__((always_inline)) { return (s & 1) ? S.XA : S.TT1; };
  auto tbuf = [&](int s) __attribute__((always_inline)) { return S.TT + ST::T0 + ST::TB * (s & 1); };
  // the copies of pass p: image p+3 into its ring slot, T of p+3, A1 rows of p+2
  auto copies = [&](int p) __attribute__((always_inline)) {
    const int si = min(p + 3, NSUB - 1), sa = min(p + 2, NSUB - 1);
    _Float16 *slot = hvbuf(S, (p + 3) & 3);
    const _Float16 *src = img(p + 3);
    double *tb = tbuf(p + 3), *ab = abuf(p + 2);
    return [=, &S](int k) __attribute__((always_inline)) {
      if (k < HSlot<NB>::NOPS)
        dmah_op<NB>(k, slot, src);
      else if (k == HSlot<NB>::NOPS)
        dmat_op<NB>(tb, T, ib, si);
      else if (!UNIT)
        dmaa_op<NB>(k - HSlot<NB>::NOPS - 1, ab, a1slab, sa);
    };
  };
  f4 acc[NB / 16];
  f4 Wn[2];
  h4 W2h = h4{0, 0, 0, 0}, W2l = W2h;
  float tto[4];
  const int sh = CE[col];
  const double dsc = ldexp(1.0, sh);
  const float xs = (float)dsc, xu = (float)ldexp(1.0, -sh);
  const int c = tid >> 4, pr = tid & 15;  // this thread's element of a 16 x 16 T block
  // prologue: images 0 and 1, the A1 rows of 0, T^T of 0 and 1 (register
  // loads), the strip; all waited for
  {
    if (!UNIT) {
      dmaa_op<NB>(0, abuf(0), a1slab, 0);
      dmaa_op<NB>(1, abuf(0), a1slab, 0);
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
#pragma unroll
      for (int q = 0; q < HSlot<NB>::NOPS; q++) dmah_op<NB>(q, hvbuf(S, j), img(j));
      const double tv = *gmem(T + (size_t)(16 * j + c) * ib + (16 * j) % ib + pr);
      f32tbuf(S, j)[c * LDT16 + pr] = pr <= c ? (float)tv : 0.f;
    }
#pragma unroll
    for (int rb = 0; rb < NB / 16; rb++) {
      const d4 x = *reinterpret_cast<const d4 *>(xcol + 16 * rb + 4 * g);
      acc[rb] = f4{(float)(x[0] * dsc), (float)(x[1] * dsc), (float)(x[2] * dsc), (float)(x[3] * dsc)};
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  // after pass s: this wave's copies of pass s-1 waited for; W2 of sub-block
  // s+1 and its FP16 split; T^T of s+2 into the ring; barrier; the new A1 rows
  // of s+1 stored
  auto post = [&](int s, bool copied) __attribute__((always_inline)) {
    if (copied)
      __builtin_amdgcn_s_waitcnt(waitcnt_vm_lgkm0(NC));
    else
      __builtin_amdgcn_s_waitcnt(0);
    f4 W = (Wn[0] + Wn[1]) * HV_UNSCALE;
    f4 a1 = f4{0.f, 0.f, 0.f, 0.f};
    if (!UNIT) {
      const double *ar = abuf(s + 1) + 16 * (16 * wave + m) + 4 * g;
#pragma unroll
      for (int v = 0; v < 4; v++) a1[v] = (float)(ar[v] * dsc);
      W += a1;
    }
    f4 t = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int sp = 0; sp < 4; sp++) t = mfma32(tto[sp], W[sp], t);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      _Float16 h, l;
      hsplit(-t[v] * HV_UNSCALE, h, l);
      W2h[v] = h;
      W2l[v] = l;
    }
    if (s >= 0 && s + 2 < NSUB)
      f32tbuf(S, (s + 2) & 3)[c * LDT16 + pr] = pr <= c ? (float)tbuf(s + 2)[16 * c + pr] : 0.f;
    lds_barrier();
    if (!UNIT) {
      const f4 a1o = a1 - t;
      *reinterpret_cast<d4 *>(a1col + 16 * (s + 1) + 4 * g) =
          d4{(double)a1o[0] * xu, (double)a1o[1] * xu, (double)a1o[2] * xu, (double)a1o[3] * xu};
    }
    Wn[0] = f4{0.f, 0.f, 0.f, 0.f};
    Wn[1] = Wn[0];
  };
  Wn[0] = f4{0.f, 0.f, 0.f, 0.f};
  Wn[1] = Wn[0];
  // pass -1: W of sub-block 0
  fpassh2<NB, false, true, NC>(acc, W2h, W2l, Wn, tto, hvbuf(S, 3), hvbuf(S, 0), f32tbuf(S, 0), copies(-1));
  post(-1, true);
  // pass s: update with s, accumulate with s+1 (copies clamped to the last
  // sub-block: every pass makes the same number)
  for (int s = 0; s + 1 < NSUB; s++) {
    fpassh2<NB, true, true, NC>(acc, W2h, W2l, Wn, tto, hvbuf(S, s & 3), hvbuf(S, (s + 1) & 3),
                                f32tbuf(S, (s + 1) & 3), copies(s));
    post(s, true);
  }
  fpassh2<NB, true, false, 0>(acc, W2h, W2l, Wn, tto, hvbuf(S, (NSUB - 1) & 3), hvbuf(S, 0), f32tbuf(S, 0),
                              [](int) {});
  const double u = (double)xu;
#pragma unroll
  for (int rb = 0; rb < NB / 16; rb++) {
    const f4 x = acc[rb];
    *reinterpret_cast<d4 *>(xcol + 16 * rb + 4 * g) = d4{(double)x[0] * u, (double)x[1] * u, (double)x[2] * u, (double)x[3] * u};
  }
  // no copy may still be in flight when the next task reuses LDS
  __builtin_amdgcn_s_waitcnt(0);
}

// PREC (hq_geom::prec) is a template parameter: each precision of the
// updates is its own kernel instantiation with its own register allocation.
template <int NB, int WB, int PREC = 0>
__device__ __forceinline__ int2 task_larfb(Smem<NB> &S, const double *V, const double *T, double *C,
                                           int ib, int slab, const double *Wt = nullptr,
                                           const double *Vd = nullptr, Ahead ah = Ahead{},
                                           const unsigned short *VHt = nullptr, const int *CE = nullptr) {
  if constexpr (NB >= 128 && PREC == 2) {
    if (CE && ib % 16 == 0)
      task_update_h2<NB, true>(S, VHt, T, C, nullptr, ib, slab, CE);
    else
      task_update_h<NB, true, hq_upd_slab(NB, 2) / 64>(S, VHt, T, C, nullptr, ib, slab, CE);
  } else if constexpr (NB >= 128 && PREC == 1) {
    task_update_f32<NB, true, hq_upd_slab(NB, 1) / 64>(S, V, T, C, nullptr, ib, slab);
  } else if constexpr (NB >= 128) {
    if (Wt)
      task_larfb_w<NB>(S, Vd, Wt, C, slab);
    else
      return task_update_f<NB, true>(S, V, T, C, nullptr, ib, slab, ah);
  } else {
    task_update<NB, WB, true>(S, V, T, C, nullptr, ib, slab);
  }
  return make_int2(-1, 0);
}

template <int NB, int WB, int PREC = 0>
__device__ __forceinline__ int2 task_ssrfb(Smem<NB> &S, double *A1, double *A2, const double *V,
                                           const double *T, int ib, int slab, Ahead ah = Ahead{},
                                           const unsigned short *VHt = nullptr, const int *CE = nullptr) {
  if constexpr (NB >= 128 && PREC == 2) {
    if (CE && ib % 16 == 0)
      task_update_h2<NB, false>(S, VHt, T, A2, A1, ib, slab, CE);
    else
      task_update_h<NB, false, hq_upd_slab(NB, 2) / 64>(S, VHt, T, A2, A1, ib, slab, CE);
  } else if constexpr (NB >= 128 && PREC == 1) {
    task_update_f32<NB, false, hq_upd_slab(NB, 1) / 64>(S, V, T, A2, A1, ib, slab);
  } else if constexpr (NB >= 128) {
    return task_update_f<NB, false>(S, V, T, A2, A1, ib, slab, ah);
  } else {
    task_update<NB, WB, false>(S, V, T, A2, A1, ib, slab);
  }
  return make_int2(-1, 0);
}

__device__ __forceinline__ double *tile_ptr(const hq_geom &G, int i, int j) {
  return G.A + ((size_t)i + (size_t)j * G.MT) * (size_t)G.nb * G.nb;
}
__device__ __forceinline__ double *xtile_ptr(const hq_geom &G, int i, int j) {
  return G.X + ((size_t)i + (size_t)j * G.MT) * (size_t)G.nb * G.nb;
}
__device__ __forceinline__ double *t_ptr(const hq_geom &G, int i, int j) {
  return G.T + ((size_t)i + (size_t)j * G.MT) * (size_t)G.ib * G.nb;
}
__device__ __forceinline__ double *wd_ptr(const hq_geom &G, int k) {
  return G.W ? G.W + (size_t)k * G.nb * G.nb : nullptr;
}
__device__ __forceinline__ double *vd_ptr(const hq_geom &G, int k) {
  return G.W ? G.VD + (size_t)k * G.nb * G.nb : nullptr;
}
// Images per tile: one per 16-reflector sub-block, ceil(nb/ib) ib blocks of
// ceil(ib/16) sub-blocks (nb/16 when ib is a multiple of 16).
__device__ __forceinline__ unsigned short *vh_ptr(const hq_geom &G, int i, int j) {
  const size_t per_tile = (size_t)((G.nb + G.ib - 1) / G.ib) * ((G.ib + 15) / 16) * 64 * G.nb;
  return G.VH ? G.VH + ((size_t)i + (size_t)j * G.MT) * per_tile : nullptr;
}
// column scales of tile column j (split factorisation), or nullptr
__device__ __forceinline__ const int *ce_ptr(const hq_geom &G, int j) {
  return G.CE ? G.CE + (size_t)j * G.nb : nullptr;
}
__device__ __forceinline__ double *tau_ptr(const hq_geom &G, int i, int j) {
  return G.TAU + ((size_t)i + (size_t)j * G.MT) * (size_t)G.nb;
}

// Columns of tile column j that hold matrix entries (hq_geom::ncol).
__device__ __forceinline__ int live_cols(const hq_geom &G, int j) {
  return G.ncol > 0 ? min(G.nb, max(0, G.ncol - j * G.nb)) : G.nb;
}

// Returns the look-ahead of a slab task (see task_update_f), else (-1, 0).
// Tasks wholly inside a batch matrix's zero padding write only the zeros
// their outputs hold (pad_panel_blocks) or nothing.
template <int NB, int WB, int PREC = 0>
__device__ __forceinline__ int2 run_task(Smem<NB> &S, const hq_geom &G, const hq_task &t, Ahead ah = Ahead{}) {
  switch (t.kind) {
    case HQ_GEQRT:  // j = ib block of the panel, or -1: the whole tile
      if (t.j < 0)
        task_geqrt<NB, WB>(S, tile_ptr(G, t.k, t.k), t_ptr(G, t.k, t.k), tau_ptr(G, t.k, t.k), G.ib,
                           wd_ptr(G, t.k), vd_ptr(G, t.k), PREC == 2 ? vh_ptr(G, t.k, t.k) : nullptr,
                           live_cols(G, t.k));
      else if (t.j * G.ib >= live_cols(G, t.k))
        pad_panel_blocks<NB>(t_ptr(G, t.k, t.k), tau_ptr(G, t.k, t.k), G.ib, t.j * G.ib, wd_ptr(G, t.k),
                             vd_ptr(G, t.k));
      else
        geqrt_panel<NB, WB>(S, tile_ptr(G, t.k, t.k), t_ptr(G, t.k, t.k), tau_ptr(G, t.k, t.k), G.ib,
                            t.j * G.ib, wd_ptr(G, t.k), vd_ptr(G, t.k), PREC == 2 ? vh_ptr(G, t.k, t.k) : nullptr);
      break;
    case HQ_TSQRT:
      if (t.j < 0)
        task_tsqrt<NB, WB>(S, tile_ptr(G, t.k, t.k), tile_ptr(G, t.r, t.k), t_ptr(G, t.r, t.k),
                           tau_ptr(G, t.r, t.k), G.ib, PREC == 2 ? vh_ptr(G, t.r, t.k) : nullptr, live_cols(G, t.k));
      else if (t.j * G.ib >= live_cols(G, t.k))
        pad_panel_blocks<NB>(t_ptr(G, t.r, t.k), tau_ptr(G, t.r, t.k), G.ib, t.j * G.ib, nullptr, nullptr);
      else
        tsqrt_panel<NB, WB>(S, tile_ptr(G, t.k, t.k), tile_ptr(G, t.r, t.k), t_ptr(G, t.r, t.k),
                            tau_ptr(G, t.r, t.k), G.ib, t.j * G.ib, PREC == 2 ? vh_ptr(G, t.r, t.k) : nullptr);
      break;
    case HQ_GEQRT_U:
      if (t.j * G.ib < live_cols(G, t.k) && t.slab * HQ_SLAB < live_cols(G, t.k))
        task_geqrt_update<NB, WB>(S, tile_ptr(G, t.k, t.k), t_ptr(G, t.k, t.k), G.ib, t.j, t.slab);
      break;
    case HQ_TSQRT_U:
      if (t.j * G.ib < live_cols(G, t.k) && t.slab * HQ_SLAB < live_cols(G, t.k))
        task_tsqrt_update<NB, WB>(S, tile_ptr(G, t.k, t.k), tile_ptr(G, t.r, t.k), t_ptr(G, t.r, t.k), G.ib, t.j,
                                  t.slab);
      break;
    case HQ_LARFB:
      if (t.slab * hq_upd_slab(NB, PREC) >= live_cols(G, t.j)) break;
      return task_larfb<NB, WB, PREC>(S, tile_ptr(G, t.k, t.k), t_ptr(G, t.k, t.k), xtile_ptr(G, t.k, t.j), G.ib,
                                      t.slab, wd_ptr(G, t.k), vd_ptr(G, t.k), ah, vh_ptr(G, t.k, t.k),
                                      ce_ptr(G, t.j));
    case HQ_SSRFB:
      if (t.slab * hq_upd_slab(NB, PREC) >= live_cols(G, t.j)) break;
      return task_ssrfb<NB, WB, PREC>(S, xtile_ptr(G, t.k, t.j), xtile_ptr(G, t.r, t.j), tile_ptr(G, t.r, t.k),
                                      t_ptr(G, t.r, t.k), G.ib, t.slab, ah, vh_ptr(G, t.r, t.k), ce_ptr(G, t.j));
    default:
      break;
  }
  return make_int2(-1, 0);
}

// Persistent DAG scheduler.  Workgroups claim task indices in order from one
// atomic counter; a claimed task waits (thread 0, with s_sleep back-off) until
// every dependency's flag equals this run's epoch, runs, then publishes its own
// flag with release semantics.  Dependencies have smaller indices and were
// claimed by running workgroups, so every wait ends.
// A batch passes one geometry per matrix (geoms[task.mat]); a single
// factorisation passes geoms = nullptr and its geometry by value.
template <int NB, int WB, int PREC>
__global__ __launch_bounds__(HQ_THREADS) void sched_kernel(hq_geom G, const hq_geom *__restrict__ geoms,
                                                          const hq_task *__restrict__ tasks,
                                                          int ntasks, int *flags, int epoch_arg,
                                                          int *counter, long long *trace,
                                                          const int *epoch_dev, unsigned wait_limit) {
  __shared__ Smem<NB> S;
  // this run's flag value: the argument, or the device epoch advanced by the
  // counter reset just before (graph replays)
  const int epoch = epoch_dev ? __hip_atomic_load(epoch_dev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : epoch_arg;
  int done = -1;  // task finished by this workgroup in the previous iteration
  long long t_claim = 0;
  const Ahead ah{counter, tasks, flags, ntasks, epoch};
  int2 la = make_int2(-1, 0);  // the next task, claimed during the last pass of a slab task
  for (;;) {
    // Thread 0 publishes the previous task and claims the next one in ONE
    // divergent region.  (Two separate thread-0 regions at the end and the
    // start of the loop body get merged across the back edge by the CFG
    // structurizer into an inner loop that the other lanes never leave.)
    if (threadIdx.x == 0) {
      if (done >= 0) {
        // every store of the task completed before the barrier that ended
        // it; the L2 write-back of the release must complete too before the
        // flag can be seen (the fence alone leaves no wait here)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_store(&flags[done], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (trace && done >= 0) trace[4 * done + 2] = wall_clock64();
      int t = la.x >= 0 ? la.x : atomicAdd(counter, 1);
      const bool seen_ready = la.x >= 0 && (la.y & 1) != 0;
      if (trace) t_claim = wall_clock64();
      // counter[1] != 0: a wait timed out somewhere; drain without work (a
      // look-ahead read it already)
      if (la.x >= 0 ? (la.y & 2) != 0
                    : __hip_atomic_load(&counter[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
        t = ntasks;
      if (t < ntasks && !seen_ready) {
        // (a look-ahead that saw the task ready left its descriptor in S.next
        // and read nothing else this needs)
        const hq_task &tk = tasks[t];
        if (la.x < 0) copy_task(S.next, tk);
        const int nd = tk.ndep;
        unsigned spins = 0;
        for (;;) {
          // every dependency flag in flight at once (one round trip per check)
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
          if (++spins > wait_limit) {  // (default 2^22: seconds) a dependency never completed
            __hip_atomic_store(&counter[1], 1 + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            t = ntasks;
            break;
          }
        }
      }
      if (t < ntasks && !seen_ready) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      S.task = t;
    }
    __syncthreads();
    // readfirstlane: the task index, the loop exit and every barrier after it
    // are wave-uniform for the compiler
    const int t = __builtin_amdgcn_readfirstlane(S.task);
    if (t >= ntasks) break;
    // the descriptor from LDS, each field made wave-uniform like the index
    hq_task tk;
    tk.kind = __builtin_amdgcn_readfirstlane(S.next.kind);
    tk.k = __builtin_amdgcn_readfirstlane(S.next.k);
    tk.r = __builtin_amdgcn_readfirstlane(S.next.r);
    tk.j = __builtin_amdgcn_readfirstlane(S.next.j);
    tk.slab = __builtin_amdgcn_readfirstlane(S.next.slab);
    tk.mat = __builtin_amdgcn_readfirstlane(S.next.mat);
    tk.ndep = 0;  // (dependencies are the scheduler's business only)
    if (trace && threadIdx.x == 0) {
      trace[4 * t + 0] = t_claim;
      trace[4 * t + 1] = wall_clock64();
      trace[4 * t + 3] = blockIdx.x;
    }
    if (geoms)
      la = run_task<NB, WB, PREC>(S, geoms[tk.mat], tk, ah);
    else
      la = run_task<NB, WB, PREC>(S, G, tk, ah);
    // each wave's stores have completed (a barrier waits for LDS only), so
    // none of them can land after the task's flag is published
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    done = t;
  }
  if (trace && threadIdx.x == 0 && done >= 0) trace[4 * done + 2] = wall_clock64();
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

// Rebuild the ib x nb T factor of every reflector tile (r, k), r >= k, from
// V and TAU (LAPACK dlarft('F','C') per ib block), for applying Q when only
// RY and TAU are given (the reference rebuilds T the same way in
// Hybrid_dorgqr.c:84,118 / Hybrid_make_Q.c:124-130,159-163).  Diagonal tiles
// use the unit lower part of the tile; their last reflector is dropped, as
// the reference does by zeroing Y(nb-1,nb-1) (Hybrid_dorgqr.c:83).
// grid.x = ib block, grid.y = flattened lower tile index.
template <int NB, int WB>
__global__ __launch_bounds__(HQ_THREADS) void tbuild_kernel(hq_geom G, const int2 *tiles) {
  __shared__ Smem<NB> S;
  constexpr int LDV = Smem<NB>::LDV;
  constexpr int IBP = 16 * WB;
  const int2 rk = tiles[blockIdx.y];
  const int r = rk.x, k = rk.y;
  const int ib = G.ib, i0 = blockIdx.x * ib;
  if (i0 >= NB) return;
  const int kb = min(ib, NB - i0);
  const bool diag = (r == k);
  const double *V = tile_ptr(G, r, k);
  const double *tau = tau_ptr(G, r, k) + i0;
  for (int e = threadIdx.x; e < IBP * NB; e += HQ_THREADS) {
    const int c = e / NB, rr = e - c * NB;
    double val = 0.0;
    if (c < kb) {
      const int d = i0 + c;
      val = V[(size_t)d * NB + rr];
      if (diag) val = rr > d ? val : (rr == d ? 1.0 : 0.0);
    }
    S.V[c * LDV + rr] = val;
  }
  if (threadIdx.x < HQ_IBMAX) {
    const int c = threadIdx.x;
    double tv = c < kb ? tau[c] : 0.0;
    if (diag && i0 + c == NB - 1) tv = 0.0;
    S.tau[c] = tv;
  }
  __syncthreads();
  gram<NB, WB>(S, diag ? i0 / 16 : 0);
  __syncthreads();
  if (threadIdx.x < 64) build_t<NB>(S, kb, t_ptr(G, r, k) + (size_t)i0 * ib, ib, nullptr);
}

// T of one block of kb <= ib reflectors wider than the LDS-staged builder
// takes (ib > 32; PLASMA's T layout with ld = ib for any ib <= nb): the
// block's columns i0 .. i0+kb-1 of reflector tile V (ld = NB; diag: unit lower
// with the diagonal at row i0 + c, else the V2 of a TSQRT tile under an
// identity), tau of the block, T block at Tb (column c at Tb + c*ib).
//   1. G(p, c) = v_p^T v_c for p < c, one wave per pair, into T's strict
//      upper part;
//   2. dlarft('F','C'): column by column, T(0:c, c) = -tau_c T(0:c, 0:c)
//      G(0:c, c), T(c, c) = tau_c; column c of G is read into LDS before the
//      column is overwritten.  Rows c < p < ib are written as zero.
template <int NB>
__device__ void tbuild_wide_body(const double *V, const double *tau, double *Tb, int ib, int i0, int kb,
                                 bool diag, bool drop_last) {
  __shared__ double gcol[NB];
  __shared__ double tv[NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int c = tid; c < kb; c += HQ_THREADS) {
    double t = tau[c];
    if (drop_last && diag && i0 + c == NB - 1) t = 0.0;
    tv[c] = t;
  }
  const int npair = kb * (kb - 1) / 2;
  for (int e = wave; e < npair; e += HQ_THREADS / 64) {
    // pair e -> (p, c), p < c: c = largest with c(c-1)/2 <= e
    int c = (int)((1.0 + sqrt(1.0 + 8.0 * (double)e)) * 0.5);
    while (c * (c - 1) / 2 > e) c--;
    while ((c + 1) * c / 2 <= e) c++;
    const int p = e - c * (c - 1) / 2;
    const double *vp = V + (size_t)(i0 + p) * NB, *vc = V + (size_t)(i0 + c) * NB;
    double s = 0.0;
    for (int row = lane; row < NB; row += 64) {
      double a = vp[row], b = vc[row];
      if (diag) {
        const int dp = i0 + p, dc = i0 + c;
        a = row > dp ? a : (row == dp ? 1.0 : 0.0);
        b = row > dc ? b : (row == dc ? 1.0 : 0.0);
      }
      s += a * b;
    }
    s = wave_sum1(s);
    if (lane == 0) Tb[(size_t)c * ib + p] = s;
  }
  __syncthreads();
  for (int c = 0; c < kb; c++) {
    for (int q = tid; q < c; q += HQ_THREADS) gcol[q] = Tb[(size_t)c * ib + q];
    __syncthreads();
    const double tc = tv[c];
    for (int p = tid; p < ib; p += HQ_THREADS) {
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
// the lower tile list), as tbuild_kernel for wide blocks.
template <int NB>
__global__ __launch_bounds__(HQ_THREADS) void tbuild_wide_kernel(hq_geom G, const int2 *tiles, int drop_last) {
  const int2 rk = tiles[blockIdx.y];
  const int r = rk.x, k = rk.y, ib = G.ib, i0 = blockIdx.x * ib;
  if (i0 >= NB) return;
  tbuild_wide_body<NB>(tile_ptr(G, r, k), tau_ptr(G, r, k) + i0, t_ptr(G, r, k) + (size_t)i0 * ib, ib, i0,
                       min(ib, NB - i0), r == k, drop_last != 0);
}
// One tile given by pointers (grid.x = ib block).
template <int NB>
__global__ __launch_bounds__(HQ_THREADS) void tbuild_wide_tile_kernel(const double *V, const double *tau,
                                                                     double *T, int ib, int diag) {
  const int i0 = blockIdx.x * ib;
  if (i0 >= NB) return;
  tbuild_wide_body<NB>(V, tau + i0, T + (size_t)i0 * ib, ib, i0, min(ib, NB - i0), diag != 0, false);
}

// Zero the entries of an MT x NT tile grid past row m / column n (the
// padding of edge tiles), touching only the last tile row and column.
// Exponent sh of each matrix column's split scale: one wave per column, the
// column's 2-norm N over the MT tiles by the scaled sum of squares (LAPACK
// dlassq: no overflow or underflow for any finite entries), then
// sh = 12 - e with N < 2^e, so 2^sh N < 2^12 (sh = 0 for a zero column).
__global__ __launch_bounds__(256) void colscale_kernel(const double *A, int MT, int NT, int nb, int *CE) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= NT * nb) return;
  const int tj = c / nb, cc = c - tj * nb;
  double sc = 0.0, sq = 1.0;
  for (int ti = 0; ti < MT; ti++) {
    const double *col = A + ((size_t)ti + (size_t)tj * MT) * nb * nb + (size_t)cc * nb;
    for (int r = lane; r < nb; r += 64) {
      const double x = fabs(col[r]);
      if (x > 0.0) {
        if (x > sc) {
          sq = 1.0 + sq * (sc / x) * (sc / x);
          sc = x;
        } else {
          sq += (x / sc) * (x / sc);
        }
      }
    }
  }
  for (int o = 32; o >= 1; o >>= 1) {
    const double s2 = __shfl_xor(sc, o, 64), q2 = __shfl_xor(sq, o, 64);
    const double m = fmax(sc, s2);
    if (m > 0.0) {
      const double a = sc / m, b = s2 / m;
      sq = (sc > 0.0 ? sq * a * a : 0.0) + (s2 > 0.0 ? q2 * b * b : 0.0);
    }
    sc = m;
  }
  if (lane == 0) {
    int e = 0;
    const double N = sc * sqrt(sq);
    if (N > 0.0 && N <= 1.7e308) (void)frexp(N, &e);
    CE[c] = N > 0.0 ? 12 - e : 0;
  }
}

__global__ void pad_zero_kernel(double *A, int MT, int NT, int nb, int m, int n) {
  const size_t ts = (size_t)nb * nb;
  const size_t total = ts * (size_t)(MT + NT);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t t = e / ts, w = e - t * ts;
    const size_t c = w / nb, rr = w - c * nb;
    size_t ti, tj;
    if (t < (size_t)NT) {  // last tile row, tile column t
      ti = MT - 1;
      tj = t;
    } else {               // last tile column, tile row t - NT
      ti = t - NT;
      tj = NT - 1;
    }
    const size_t row = ti * nb + rr, col = tj * nb + c;
    if (row >= (size_t)m || col >= (size_t)n) A[(ti + tj * (size_t)MT) * ts + w] = 0.0;
  }
}

// X <- identity (full = true: every entry written) or only ones on the
// diagonal of the diagonal tiles (Hybrid_make_Q.c:91-95 writes only those).
__global__ void identity_kernel(double *X, int MT, int XNT, int nb, int full) {
  const size_t ts = (size_t)nb * nb, total = ts * (size_t)MT * XNT;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    size_t t = e / ts, w = e - t * ts;
    size_t ti = t % MT, tj = t / MT, c = w / nb, rr = w - c * nb;
    bool on = (ti == tj) && (c == rr);
    if (full || on) X[e] = on ? 1.0 : 0.0;
  }
}

// Zero the reflectors of RY (below-diagonal part of the diagonal tiles and
// every tile below them), leaving R (Hybrid_make_Q.c:148-150,179-182).
__global__ void zero_reflectors_kernel(double *A, int MT, int NT, int nb) {
  const size_t ts = (size_t)nb * nb, total = ts * (size_t)MT * NT;
  const int KT = MT < NT ? MT : NT;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    size_t t = e / ts, w = e - t * ts;
    size_t ti = t % MT, tj = t / MT, c = w / nb, rr = w - c * nb;
    if (tj < (size_t)KT && (ti > tj || (ti == tj && rr > c))) A[e] = 0.0;
  }
}

// Claim counter reset before a scheduler launch; with epoch_dev (launches
// that may be replayed from a captured graph) the run's epoch is advanced on
// the device too, since a captured kernel argument would stay fixed.
__global__ void reset_counter_kernel(int *counter, int *epoch_dev) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *counter = 0;
    if (epoch_dev) *epoch_dev = *epoch_dev + 1;
  }
}

// p[0 .. n) <- value (clears counters and flags with a kernel, which a
// captured graph replays correctly, unlike a memset node).
__global__ void fill_int_kernel(int *p, size_t n, int value) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    p[e] = value;
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

// Layout operations on a batch, one matrix per blockIdx.y, grid-stride over
// its tiles: OP 0 fills uniform[-1,1) (value of (row, col) from
// (seed, col*m + row), as a column-major m x n fill), OP 1 converts
// column-major -> tile layout, OP 2 tile layout -> column-major.  Tile
// entries past row m or column n are written as zero by OP 0 and OP 1.
template <int OP>
__global__ __launch_bounds__(256) void batched_layout_kernel(const hq_bmat *__restrict__ mats, int nb) {
  const hq_bmat b = mats[blockIdx.y];
  const size_t ts = (size_t)nb * nb;
  const size_t total = ts * (size_t)b.MT * b.NT;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t t = e / ts, w = e - t * ts;
    const size_t ti = t % b.MT, tj = t / b.MT;
    const size_t c = w / nb, r = w - c * nb;
    const size_t row = ti * nb + r, col = tj * nb + c;
    const bool in = row < (size_t)b.m && col < (size_t)b.n;
    if (OP == 0) {
      b.A[e] = in ? uniform_pm1(b.seed, col * b.m + row) : 0.0;
    } else if (OP == 1) {
      b.A[e] = in ? b.C[row + col * (size_t)b.ld] : 0.0;
    } else if (OP == 2) {
      if (in) b.C[row + col * (size_t)b.ld] = b.A[e];
    } else if (OP == 3) {  // FP32 column-major -> FP64 tile layout
      b.A[e] = in ? (double)b.Cf[row + col * (size_t)b.ld] : 0.0;
    } else {               // FP64 tile layout -> FP32 column-major
      if (in) b.Cf[row + col * (size_t)b.ld] = (float)b.A[e];
    }
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

// Tile sizes without a specialised kernel set run the generic kernels
// (hq_generic.hip).
static bool hq_tuned_nb(int nb) { return nb == 64 || nb == 128 || nb == 256; }
hipError_t hqg_launch_sched(const hq_geom &G, const hq_geom *geoms, const hq_task *tasks, int ntasks, int *flags,
                            int epoch, int *counter, int grid, long long *trace, hipStream_t st, int *epoch_dev,
                            unsigned wait_limit);
hipError_t hqg_launch_task(const hq_geom &G, const hq_task &tk, hipStream_t st);
hipError_t hqg_launch_geqrt(int nb, int ib, double *A, double *T, double *TAU, hipStream_t st);
hipError_t hqg_launch_tsqrt(int nb, int ib, double *A1, double *A2, double *T, double *TAU, hipStream_t st);
hipError_t hqg_launch_larfb(int nb, int ib, const double *V, const double *T, double *C, hipStream_t st);
hipError_t hqg_launch_ssrfb(int nb, int ib, double *A1, double *A2, const double *V, const double *T,
                            hipStream_t st);
hipError_t hqg_launch_tbuild(const hq_geom &G, const int2 *tiles, int ntiles, int drop_last, hipStream_t st);
hipError_t hqg_launch_tbuild_tile(int nb, int ib, const double *V, const double *tau, double *T, int diag,
                                  hipStream_t st);

hipError_t hq_launch_sched(const hq_geom &G, const hq_geom *geoms, const hq_task *tasks, int ntasks,
                           int *flags, int epoch, int *counter, int grid, long long *trace, hipStream_t st,
                           int *epoch_dev, unsigned wait_limit) {
  hipLaunchKernelGGL(hq::reset_counter_kernel, dim3(1), dim3(64), 0, st, counter, epoch_dev);
  if (!hq_tuned_nb(G.nb) || G.uw)
    return hqg_launch_sched(G, geoms, tasks, ntasks, flags, epoch, counter, grid, trace, st, epoch_dev, wait_limit);
  return hq_dispatch(G.nb, G.ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    if constexpr (NB >= 128) {
      if (G.prec == 2) {
        // one instantiation (its panels take any ib <= 32) for the split
        // updates: the ib <= 16 one at nb = 256 crashes this compiler's
        // AGPR-copy rewrite pass
        hipLaunchKernelGGL((hq::sched_kernel<NB, 2, 2>), dim3(grid), dim3(HQ_THREADS), 0, st, G, geoms, tasks,
                           ntasks, flags, epoch, counter, trace, (const int *)epoch_dev, wait_limit);
        return;
      }
      if (G.prec == 1) {
        hipLaunchKernelGGL((hq::sched_kernel<NB, WB, 1>), dim3(grid), dim3(HQ_THREADS), 0, st, G, geoms, tasks,
                           ntasks, flags, epoch, counter, trace, (const int *)epoch_dev, wait_limit);
        return;
      }
    }
    hipLaunchKernelGGL((hq::sched_kernel<NB, WB, 0>), dim3(grid), dim3(HQ_THREADS), 0, st, G, geoms, tasks,
                       ntasks, flags, epoch, counter, trace, (const int *)epoch_dev, wait_limit);
  });
}

hipError_t hq_launch_batched_layout(int op, const hq_bmat *mats, int count, int nb, int max_tiles,
                                    hipStream_t st) {
  // enough blocks in x for the largest matrix at ~8 elements per thread
  const size_t elems = (size_t)max_tiles * nb * nb;
  const unsigned bx = (unsigned)std::min<size_t>(1024, (elems + 2047) / 2048);
  const dim3 grid(bx, (unsigned)count);
  if (op == 0)
    hipLaunchKernelGGL(hq::batched_layout_kernel<0>, grid, dim3(256), 0, st, mats, nb);
  else if (op == 1)
    hipLaunchKernelGGL(hq::batched_layout_kernel<1>, grid, dim3(256), 0, st, mats, nb);
  else if (op == 2)
    hipLaunchKernelGGL(hq::batched_layout_kernel<2>, grid, dim3(256), 0, st, mats, nb);
  else if (op == 3)
    hipLaunchKernelGGL(hq::batched_layout_kernel<3>, grid, dim3(256), 0, st, mats, nb);
  else
    hipLaunchKernelGGL(hq::batched_layout_kernel<4>, grid, dim3(256), 0, st, mats, nb);
  return hipGetLastError();
}

hipError_t hq_launch_task(const hq_geom &G, const hq_task &tk, hipStream_t st) {
  if (!hq_tuned_nb(G.nb)) return hqg_launch_task(G, tk, st);
  return hq_dispatch(G.nb, G.ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    hipLaunchKernelGGL((hq::task_kernel<NB, WB>), dim3(1), dim3(HQ_THREADS), 0, st, G, tk);
  });
}

hipError_t hq_launch_geqrt(int nb, int ib, double *A, double *T, double *TAU, hipStream_t st) {
  if (!hq_tuned_nb(nb)) return hqg_launch_geqrt(nb, ib, A, T, TAU, st);
  return hq_dispatch(nb, ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    hipLaunchKernelGGL((hq::geqrt_kernel<NB, WB>), dim3(1), dim3(HQ_THREADS), 0, st, A, T, TAU, ib);
  });
}
hipError_t hq_launch_tsqrt(int nb, int ib, double *A1, double *A2, double *T, double *TAU,
                           hipStream_t st) {
  if (!hq_tuned_nb(nb)) return hqg_launch_tsqrt(nb, ib, A1, A2, T, TAU, st);
  return hq_dispatch(nb, ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    hipLaunchKernelGGL((hq::tsqrt_kernel<NB, WB>), dim3(1), dim3(HQ_THREADS), 0, st, A1, A2, T, TAU,
                       ib);
  });
}
hipError_t hq_launch_larfb(int nb, int ib, const double *V, const double *T, double *C,
                           hipStream_t st) {
  if (!hq_tuned_nb(nb)) return hqg_launch_larfb(nb, ib, V, T, C, st);
  return hq_dispatch(nb, ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    hipLaunchKernelGGL((hq::larfb_kernel<NB, WB>), dim3((NB + HQ_SLAB - 1) / HQ_SLAB),
                       dim3(HQ_THREADS), 0, st, V, T, C, ib);
  });
}
hipError_t hq_launch_ssrfb(int nb, int ib, double *A1, double *A2, const double *V, const double *T,
                           hipStream_t st) {
  if (!hq_tuned_nb(nb)) return hqg_launch_ssrfb(nb, ib, A1, A2, V, T, st);
  return hq_dispatch(nb, ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    hipLaunchKernelGGL((hq::ssrfb_kernel<NB, WB>), dim3((NB + HQ_SLAB - 1) / HQ_SLAB),
                       dim3(HQ_THREADS), 0, st, A1, A2, V, T, ib);
  });
}

hipError_t hq_launch_tbuild(const hq_geom &G, const int2 *tiles, int ntiles, hipStream_t st) {
  if (!hq_tuned_nb(G.nb)) return hqg_launch_tbuild(G, tiles, ntiles, 1, st);
  return hq_dispatch(G.nb, G.ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    hipLaunchKernelGGL((hq::tbuild_kernel<NB, WB>), dim3((NB + G.ib - 1) / G.ib, ntiles),
                       dim3(HQ_THREADS), 0, st, G, tiles);
  });
}

static unsigned grid_for(size_t total) {
  size_t blocks = (total + 255) / 256;
  return (unsigned)(blocks > 65536 ? 65536 : (blocks ? blocks : 1));
}

hipError_t hq_launch_tbuild_wide(const hq_geom &G, const int2 *tiles, int ntiles, int drop_last, hipStream_t st) {
  if (!hq_tuned_nb(G.nb)) return hqg_launch_tbuild(G, tiles, ntiles, drop_last, st);
  return hq_dispatch(G.nb, G.ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    (void)WB;
    hipLaunchKernelGGL((hq::tbuild_wide_kernel<NB>), dim3((NB + G.ib - 1) / G.ib, ntiles), dim3(HQ_THREADS), 0,
                       st, G, tiles, drop_last);
  });
}

hipError_t hq_launch_tbuild_wide_tile(int nb, int ib, const double *V, const double *tau, double *T, int diag,
                                      hipStream_t st) {
  if (!hq_tuned_nb(nb)) return hqg_launch_tbuild_tile(nb, ib, V, tau, T, diag, st);
  return hq_dispatch(nb, ib, [&](auto nbc, auto wbc) {
    HQ_NBWB
    (void)WB;
    hipLaunchKernelGGL((hq::tbuild_wide_tile_kernel<NB>), dim3((NB + ib - 1) / ib), dim3(HQ_THREADS), 0, st, V,
                       tau, T, ib, diag);
  });
}

hipError_t hq_launch_fill_int(int *p, size_t n, int value, hipStream_t st) {
  hipLaunchKernelGGL(hq::fill_int_kernel, dim3(grid_for(n)), dim3(256), 0, st, p, n, value);
  return hipGetLastError();
}

hipError_t hq_launch_colscale(const double *A, int MT, int NT, int nb, int *CE, hipStream_t st) {
  const int cols = NT * nb;
  hipLaunchKernelGGL(hq::colscale_kernel, dim3((cols + 3) / 4), dim3(256), 0, st, A, MT, NT, nb, CE);
  return hipGetLastError();
}

hipError_t hq_launch_pad_zero(double *A, int MT, int NT, int nb, int m, int n, hipStream_t st) {
  hipLaunchKernelGGL(hq::pad_zero_kernel, dim3(grid_for((size_t)(MT + NT) * nb * nb)), dim3(256), 0, st, A, MT,
                     NT, nb, m, n);
  return hipGetLastError();
}

// Read the device fault word of the current device (after the work that may
// have raised it has completed) and clear it.
hipError_t hq_fault_take(int *value) {
  hipError_t e = hipMemcpyFromSymbol(value, HIP_SYMBOL(hq::g_fault), sizeof(int), 0, hipMemcpyDeviceToHost);
  if (e != hipSuccess || *value == 0) return e;
  const int zero = 0;
  return hipMemcpyToSymbol(HIP_SYMBOL(hq::g_fault), &zero, sizeof(int), 0, hipMemcpyHostToDevice);
}

hipError_t hq_launch_identity(double *X, int MT, int XNT, int nb, int full, hipStream_t st) {
  hipLaunchKernelGGL(hq::identity_kernel, dim3(grid_for((size_t)MT * XNT * nb * nb)), dim3(256), 0,
                     st, X, MT, XNT, nb, full);
  return hipGetLastError();
}

hipError_t hq_launch_zero_reflectors(double *A, int MT, int NT, int nb, hipStream_t st) {
  hipLaunchKernelGGL(hq::zero_reflectors_kernel, dim3(grid_for((size_t)MT * NT * nb * nb)),
                     dim3(256), 0, st, A, MT, NT, nb);
  return hipGetLastError();
}

hipError_t hq_launch_fill_uniform(double *A, int MT, int NT, int nb, uint64_t seed, hipStream_t st) {
  size_t total = (size_t)MT * NT * nb * nb;
  size_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(hq::fill_uniform_tile_kernel, dim3((unsigned)blocks), dim3(256), 0, st, A, MT,
                     NT, nb, seed);
  return hipGetLastError();
}
