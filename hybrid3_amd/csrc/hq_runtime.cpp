// hq_runtime.cpp -- host side of the tiled QR engine: the C-ABI entry points,
// the task-DAG builder that replaces the reference's dependency table
// (Hybrid_tile_QR.c:36-40,496-601) and TileSch FIFO (TileSch.c:10-61), and the
// per-device buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <queue>
#include <tuple>
#include <type_traits>
#include <vector>

#include <rocblas/rocblas.h>

#include "../../include/Hybrid_tile_QR.h"
#include "../../include/core_blas.h"
#include "../../include/hybrid3_dev.h"
#include "hq_internal.h"

// launch wrappers (hq_factor.hip)
hipError_t hq_launch_sched(const hq_geom &G, const hq_geom *geoms, const hq_task *tasks, int ntasks,
                           int *flags, int epoch, int *counter, int grid, long long *trace, hipStream_t st, int *epoch_dev = nullptr,
                           unsigned wait_limit = 1u << 22);
hipError_t hq_launch_task(const hq_geom &G, const hq_task &tk, hipStream_t st);
hipError_t hq_launch_geqrt(int nb, int ib, double *A, double *T, double *TAU, hipStream_t st);
hipError_t hq_launch_tsqrt(int nb, int ib, double *A1, double *A2, double *T, double *TAU,
                           hipStream_t st);
hipError_t hq_launch_larfb(int nb, int ib, const double *V, const double *T, double *C,
                           hipStream_t st);
hipError_t hq_launch_ssrfb(int nb, int ib, double *A1, double *A2, const double *V, const double *T,
                           hipStream_t st);
hipError_t hq_launch_fill_uniform(double *A, int MT, int NT, int nb, uint64_t seed, hipStream_t st);
hipError_t hq_launch_batched_layout(int op, const hq_bmat *mats, int count, int nb, int max_tiles,
                                    hipStream_t st);
hipError_t hq_launch_tbuild(const hq_geom &G, const int2 *tiles, int ntiles, hipStream_t st);
hipError_t hq_launch_identity(double *X, int MT, int XNT, int nb, int full, hipStream_t st);
hipError_t hq_launch_zero_reflectors(double *A, int MT, int NT, int nb, hipStream_t st);
hipError_t hq_launch_tbuild_wide(const hq_geom &G, const int2 *tiles, int ntiles, int drop_last, hipStream_t st);
hipError_t hq_launch_tbuild_wide_tile(int nb, int ib, const double *V, const double *tau, double *T, int diag,
                                      hipStream_t st);
hipError_t hq_launch_pad_zero(double *A, int MT, int NT, int nb, int m, int n, hipStream_t st);
hipError_t hq_launch_colscale(const double *A, int MT, int NT, int nb, int *CE, hipStream_t st);
hipError_t hq_fault_take(int *value);
hipError_t hq_launch_fill_int(int *p, size_t n, int value, hipStream_t st);

// homegrown tile_kernel.h API and verification helpers (hq_homegrown.hip)
hipError_t hqh_launch_dgeqrt(int row, int col, int block, double *A, int lda, double *TAU,
                             double *Y, double *T, int ldt, double *work, hipStream_t st);
hipError_t hqh_launch_dtsqrt(int row, int col, int block, double *A1, int lda1, double *A2,
                             int lda2, double *TAU, double *T, int ldt, double *work,
                             hipStream_t st);
hipError_t hqh_launch_dlarfb(int row, int col, double *A, int lda, const double *Y,
                             const double *T, int ldt, double *work, hipStream_t st);
hipError_t hqh_launch_dlarft(int row, int col, int block, const double *Y, int lda,
                             const double *TAU, double *T, int ldt, double *work, hipStream_t st);
hipError_t hqh_launch_tile_gemm_tn(const double *A, const double *B, double *Cm, int MT, int NT,
                                   int nb, hipStream_t st);

namespace {

// ---------------------------------------------------------------- errors
thread_local int g_err = HQ_OK;
thread_local char g_msg[512] = "";

int set_error(int code, const char *fmt, ...) {
  g_err = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof g_msg, fmt, ap);
  va_end(ap);
  return code;
}
int ok() {
  g_err = HQ_OK;
  g_msg[0] = 0;
  return HQ_OK;
}
#define HQ_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return set_error(HQ_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_));     \
  } while (0)

// Any tile size up to 512 runs; 64, 128 and 256 have the specialised kernel
// sets (register-resident updates, MFMA), the others the generic ones, and
// only 128 and 256 have the reduced-precision updates.
bool nb_ok(int nb) { return nb >= 1 && nb <= 512; }
bool tuned_nb(int nb) { return nb == 64 || nb == 128 || nb == 256; }
bool rp_nb(int nb) { return nb == 128 || nb == 256; }

bool geometry_ok(int MT, int NT, int nb, int ib) {
  if (MT < 1 || NT < 1) return set_error(HQ_ERR_ARG, "tile grid %d x %d is empty", MT, NT), false;
  if (!nb_ok(nb)) return set_error(HQ_ERR_ARG, "nb=%d unsupported (1..512)", nb), false;
  if (ib < 1 || ib > nb) return set_error(HQ_ERR_ARG, "ib=%d unsupported (1..nb=%d)", ib, nb), false;
  return true;
}

// Inner block width the kernels run with.  The panels factor at most
// HQ_IBMAX columns at a time; R, V and TAU do not depend on the inner
// blocking (every ib applies the same reflectors, only grouped differently,
// so results agree to rounding), and T for a wider ib is rebuilt from V and
// TAU (dlarft per ib block) where a caller asks for it.
int ibi(int ib) { return ib < HQ_IBMAX ? ib : HQ_IBMAX; }

// ---------------------------------------------------------------- DAG
// Canonical task enumeration (step-major, a topological order):
//   for k: GEQRT(k); LARFB(k,j,s) for j>k; for r>k: TSQRT(k,r); SSRFB(k,j,r,s) for j>k
// Dependencies (the reference's MEM / LIN / YTM bits, Hybrid_tile_QR.c:36-40):
//   GEQRT(k)      MEM: SSRFB(k-1,k,k,*)
//   TSQRT(k,r)    LIN: TSQRT(k,r-1) | GEQRT(k);  MEM: SSRFB(k-1,k,r,*)
//   LARFB(k,j,s)  YTM: GEQRT(k);                 MEM: SSRFB(k-1,j,k,s)
//   SSRFB(k,j,r,s) YTM: TSQRT(k,r); LIN: SSRFB(k,j,r-1,s) | LARFB(k,j,s); MEM: SSRFB(k-1,j,r,s)
struct Schedule {
  std::vector<hq_task> tasks;  // in claim order
  int grid = 0;
};

// Apply mode (Q^T X, the sweep of Hybrid_dorgqr.c:92-138 / Hybrid_dormqr.c:
// 88-127 / Hybrid_make_Q.c:139-178): no panel tasks; every tile column j of X
// (0 <= j < XNT) gets
//   LARFB(k,j,s)   MEM: SSRFB(k-1,j,k,s)
//   SSRFB(k,j,r,s) LIN: SSRFB(k,j,r-1,s) | LARFB(k,j,s); MEM: SSRFB(k-1,j,r,s)
// for each slab s of UW columns (uw_ overrides the factorisation's slab
// width) that holds one of X's first ncol_ columns (0: all); the slabs are
// independent chains, so a narrow X gets as many chains as it has slabs.
struct DagBuilder {
  int MT, NT, nb, ib, S, KT;
  bool apply;
  int XNT;
  int NBLK;  // ib blocks per tile
  int prec;  // precision of the LARFB / SSRFB updates (their slab width and cost)
  int UW, SU;  // LARFB / SSRFB slab width (hq_upd_slab) and slabs per tile
  int NCX;     // apply mode: X's columns that hold entries (0: all)
  bool NARROWX;  // apply mode on the narrow (16- / 32-column) tasks, whose SSRFB chains are pipelined
  // pipe[i]: the task that task i follows sub-block by sub-block (its dep[0],
  // the previous SSRFB of a 16-column chain; -1 otherwise)
  std::vector<int> pipe;
  std::vector<hq_task> t;
  std::vector<int> idl, ids;
  // panel tasks of the current step k: (row r, block b); their trailing
  // tasks: (row r, block b, slab s).  Row k is the GEQRT of the step.
  std::vector<int> idp, idu;
  bool split;  // GEQRT / TSQRT as panel + trailing tasks (else one task per tile)
  DagBuilder(int MT_, int NT_, int nb_, int ib_, bool apply_ = false, int XNT_ = 0, bool split_ = true,
             int prec_ = 0, int uw_ = 0, int ncol_ = 0)
      : MT(MT_), NT(NT_), nb(nb_), ib(ib_), apply(apply_), XNT(apply_ ? XNT_ : NT_), prec(apply_ ? 0 : prec_),
        split(split_) {
    S = (nb + HQ_SLAB - 1) / HQ_SLAB;
    UW = apply && uw_ ? uw_ : hq_upd_slab(nb, prec);
    NARROWX = apply && uw_ > 0;
    NCX = apply ? ncol_ : 0;
    SU = (nb + UW - 1) / UW;
    KT = std::min(MT, NT);
    NBLK = (nb + ib - 1) / ib;
    idl.assign((size_t)KT * XNT * SU, -1);
    ids.assign((size_t)KT * XNT * MT * SU, -1);
    idp.assign((size_t)MT * NBLK, -1);
    idu.assign((size_t)MT * NBLK * S, -1);
  }
  // L / Sr: update tasks by update slab u (UW columns); Su: the update task
  // covering the 64-column slab s of the panel tasks
  int &L(int k, int j, int u) { return idl[((size_t)k * XNT + j) * SU + u]; }
  int &Sr(int k, int j, int r, int u) { return ids[(((size_t)k * XNT + j) * MT + r) * SU + u]; }
  int Su(int k, int j, int r, int s) { return Sr(k, j, r, s * HQ_SLAB / UW); }
  int &P(int r, int b) { return idp[(size_t)r * NBLK + b]; }
  int &U(int r, int b, int s) { return idu[((size_t)r * NBLK + b) * S + s]; }
  int add(int kind, int k, int r, int j, int s) {
    hq_task x{};
    x.kind = kind;
    x.k = k;
    x.r = r;
    x.j = j;
    x.slab = s;
    x.ndep = 0;
    t.push_back(x);
    pipe.push_back(-1);
    return (int)t.size() - 1;
  }
  void dep(int id, int d) {
    hq_task &x = t[id];
    if (d < 0) return;
    if (x.ndep >= HQ_MAXDEP) {
      fprintf(stderr, "hybrid3: internal error: too many dependencies\n");
      return;
    }
    x.dep[x.ndep++] = d;
  }
  bool live_slab(int j, int s) const { return NCX <= 0 || j * nb + s * UW < NCX; }
  void build_apply() {
    for (int k = 0; k < KT; k++) {
      for (int j = 0; j < XNT; j++)
        for (int s = 0; s < SU; s++) {
          if (!live_slab(j, s)) continue;
          int l = L(k, j, s) = add(HQ_LARFB, k, k, j, s);
          if (k > 0) dep(l, Sr(k - 1, j, k, s));
        }
      for (int r = k + 1; r < MT; r++)
        for (int j = 0; j < XNT; j++)
          for (int s = 0; s < SU; s++) {
            if (!live_slab(j, s)) continue;
            int x = Sr(k, j, r, s) = add(HQ_SSRFB, k, r, j, s);
            dep(x, r == k + 1 ? L(k, j, s) : Sr(k, j, r - 1, s));
            if (k > 0) dep(x, Sr(k - 1, j, r, s));
            if (NARROWX && r > k + 1) pipe[x] = Sr(k, j, r - 1, s);
          }
    }
  }
  // The panel QR of tile (r, k) (GEQRT for r == k, TSQRT below it) as one
  // task per ib block b plus one trailing task per (b, slab) for the slabs
  // right of the block.  Footprints: panel b writes the block's columns of
  // tile (r,k) (rows >= b*ib for GEQRT) and, for TSQRT, the diagonal block
  // (b,b) of R(k,k); trailing (b, s) reads them and writes tile (r,k) and,
  // for TSQRT, R(k,k) rows of block b, in slab s right of the block.  So
  //   panel(r,b)    <- panel(r-1,b) [R(b,b)], trailing(r,b-1,s) for the
  //                    slabs s of block b (or the step k-1 update of them)
  //   trailing(r,b,s) <- panel(r,b), trailing(r,b-1,s) (or step k-1),
  //                    trailing(r-1,b,s) [R rows of block b in slab s]
  // and the last panel of a tile closes its whole panel QR: every trailing
  // task of the tile precedes it through the chains above.
  void panel_tasks(int k, int r) {
    const int kind = r == k ? HQ_GEQRT : HQ_TSQRT;
    const int ukind = r == k ? HQ_GEQRT_U : HQ_TSQRT_U;
    if (!split) {  // the whole panel QR of the tile in one task (j = -1)
      const int p = add(kind, k, r, -1, 0);
      if (r > k) dep(p, P(r - 1, 0));
      if (k > 0)
        for (int u = 0; u < SU; u++) dep(p, Sr(k - 1, k, r, u));
      for (int b = 0; b < NBLK; b++) P(r, b) = p;
      return;
    }
    for (int b = 0; b < NBLK; b++) {
      const int ii = b * ib, j0 = std::min(nb, ii + ib);
      const int p = P(r, b) = add(kind, k, r, b, 0);
      if (r > k) dep(p, P(r - 1, b));
      for (int s = ii / HQ_SLAB; s <= (j0 - 1) / HQ_SLAB; s++)
        if (b > 0 || s * HQ_SLAB % UW == 0 || s == ii / HQ_SLAB)  // one dependency per update task
          dep(p, b == 0 ? (k > 0 ? Su(k - 1, k, r, s) : -1) : U(r, b - 1, s));
      if (j0 >= nb) continue;
      for (int s = j0 / HQ_SLAB; s < S; s++) {
        const int u = U(r, b, s) = add(ukind, k, r, b, s);
        dep(u, p);
        dep(u, b == 0 ? (k > 0 ? Su(k - 1, k, r, s) : -1) : U(r, b - 1, s));
        if (r > k) dep(u, U(r - 1, b, s));
      }
    }
  }
  void build() {
    if (apply) {
      build_apply();
      return;
    }
    for (int k = 0; k < KT; k++) {
      std::fill(idp.begin(), idp.end(), -1);
      std::fill(idu.begin(), idu.end(), -1);
      panel_tasks(k, k);
      const int g = P(k, NBLK - 1);
      for (int j = k + 1; j < NT; j++)
        for (int s = 0; s < SU; s++) {
          int l = L(k, j, s) = add(HQ_LARFB, k, k, j, s);
          dep(l, g);
          if (k > 0) dep(l, Sr(k - 1, j, k, s));
        }
      for (int r = k + 1; r < MT; r++) {
        panel_tasks(k, r);
        const int q = P(r, NBLK - 1);
        for (int j = k + 1; j < NT; j++)
          for (int s = 0; s < SU; s++) {
            int x = Sr(k, j, r, s) = add(HQ_SSRFB, k, r, j, s);
            dep(x, q);
            dep(x, r == k + 1 ? L(k, j, s) : Sr(k, j, r - 1, s));
            if (k > 0) dep(x, Sr(k - 1, j, r, s));
          }
      }
    }
  }
  // Estimated durations (us) of each task kind for the priority computation:
  // the per-task means measured by the trace build at nb = 256, ib = 32
  // (GEQRT panel 58, TSQRT panel 43, trailing tasks 22 per 64-column slab of
  // the full tile height, LARFB 80 and SSRFB 85 per slab), scaled by the
  // work of the kind; a whole-tile panel QR (batches) ≈ 540.
  // Reduced-precision updates: the per-column cost of an update task
  // relative to FP64 (measured means per 64 columns at nb = 256: FP32 PF1,
  // split PF2, against 85 us in FP64).
  double cost(const hq_task &x) const {
    static constexpr double PF1 = 0.7, PF2 = 1.0;
    const double f = nb / 256.0;
    const double w = (double)std::min(HQ_SLAB, nb) / 64.0;  // trailing slab width / 64
    const double wu = (double)UW / 64.0 * (prec == 1 ? PF1 : prec == 2 ? PF2 : 1.0);
    if ((x.kind == HQ_GEQRT || x.kind == HQ_TSQRT) && x.j < 0) return 540.0 * f * f * f;
    const double c = std::min(ib, nb - x.j * ib) / 32.0;     // panel block width / 32
    switch (x.kind) {
      case HQ_GEQRT: return 58.0 * f * c;
      case HQ_TSQRT: return 43.0 * f * c;
      case HQ_GEQRT_U: return 40.0 * f * c * w * (nb - x.j * ib) / nb;
      case HQ_TSQRT_U: return 22.0 * f * c * w;
      case HQ_LARFB: return 80.0 * f * f * wu;
      default: return 85.0 * f * f * wu;
    }
  }
  // List-scheduling simulation on P workers with bottom-level priorities;
  // the dispatch order is a topological order in which tasks are claimed
  // roughly when they become ready.
  // A pipelined link (pipe) lets the successor start when its predecessor
  // starts and end no earlier than one sub-block after the predecessor ends.
  std::vector<int> order(int P) {
    const int n = (int)t.size();
    const bool piped = (int)pipe.size() == n;
    const double nsub = std::max(1, nb / 16);
    std::vector<std::vector<int>> succ(n), psucc(n);
    std::vector<int> indeg(n, 0);
    for (int i = 0; i < n; i++)
      for (int d = 0; d < t[i].ndep; d++) {
        if (piped && pipe[i] == t[i].dep[d])
          psucc[t[i].dep[d]].push_back(i);
        else
          succ[t[i].dep[d]].push_back(i);
        indeg[i]++;
      }
    std::vector<double> bl(n, 0.0);
    for (int i = n - 1; i >= 0; i--) {
      double mx = 0.0, mp = 0.0;
      for (int s : succ[i]) mx = std::max(mx, bl[s]);
      for (int s : psucc[i]) mp = std::max(mp, cost(t[s]) / nsub + bl[s]);
      bl[i] = std::max(cost(t[i]) + mx, mp);
    }
    std::vector<double> fin(n, 0.0);
    typedef std::pair<double, int> PI;
    std::priority_queue<PI> ready;  // max bottom level first
    std::priority_queue<PI, std::vector<PI>, std::greater<PI>> running;  // (finish, id)
    for (int i = 0; i < n; i++)
      if (indeg[i] == 0) ready.push(PI(bl[i], -i));
    std::vector<int> out;
    out.reserve(n);
    double now = 0.0;
    int busy = 0;
    while ((int)out.size() < n) {
      while (busy < P && !ready.empty()) {
        int id = -ready.top().second;
        ready.pop();
        out.push_back(id);
        fin[id] = now + cost(t[id]);
        if (piped && pipe[id] >= 0) fin[id] = std::max(fin[id], fin[pipe[id]] + cost(t[id]) / nsub);
        running.push(PI(fin[id], id));
        busy++;
        for (int s : psucc[id])
          if (--indeg[s] == 0) ready.push(PI(bl[s], -s));
      }
      if (running.empty()) break;  // cannot happen for a DAG
      PI f = running.top();
      running.pop();
      now = f.first;
      busy--;
      for (int s : succ[f.second])
        if (--indeg[s] == 0) ready.push(PI(bl[s], -s));
    }
    return out;
  }
};

// ---------------------------------------------------------------- device context
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int reserve(size_t n) {
    if (n <= bytes) return HQ_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, n) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(HQ_ERR_NOMEM, "hipMalloc(%zu) failed", n);
    }
    bytes = n;
    return HQ_OK;
  }
};

struct CachedSched {
  DevBuf tasks;
  int ntasks = 0;
  int grid = 0;
};

// A stream and a scratch buffer for one host-pointer tile-kernel call at a
// time; calls from different host threads take different lanes, so
// tile_kernel.h calls are reentrant as the reference's are (it calls them
// from inside `omp for` loops, Hybrid_dorgqr.c:92-108,124-137).
struct Lane {
  hipStream_t stream = nullptr;
  DevBuf buf;
  rocblas_handle blas = nullptr;  // made on first use, bound to `stream`
};

struct Context {
  int device = -1;
  int num_cu = 0;
  hipStream_t stream = nullptr;
  DevBuf flags, counter, T, A, TAU, X, T1;
  DevBuf prog;  // per-task progress words of the 16-column apply (hq_geom::prog)
  DevBuf W, VD;  // W = V T16 and unit lower V of the diagonal tiles (FP64 factorisation)
  DevBuf VH;     // FP16 hi / lo images of the reflector sub-blocks (split factorisation)
  DevBuf CE;     // per-column scale exponents (split factorisation)
  int epoch = 0;
  // Device-pointer calls share this context's buffers (claim counter,
  // flags, T / W / VD scratch, batch tables).  Each call waits for `busy`
  // (recorded after the previous call's launches) on its own stream before
  // enqueueing, so calls on different streams run one after another on the
  // device instead of overwriting each other's scratch.
  hipEvent_t busy = nullptr;
  unsigned wait_limit = 1u << 22;  // spins of a scheduler dependency wait before it gives up
  // host-pointer entry points (Hybrid_*) use A / TAU / X and `stream`: one
  // such call per device at a time
  std::mutex host_mu;
  std::mutex lane_mu;
  std::vector<std::unique_ptr<Lane>> lanes;  // idle lanes
  std::map<std::tuple<int, int, int, int, int, int, int, int>, std::unique_ptr<CachedSched>> scheds;
  std::map<std::pair<int, int>, std::unique_ptr<DevBuf>> tilelists;  // (MT, NT) -> int2 list
  // batches: key = (nb, ib, MT_0, NT_0, MT_1, NT_1, ...)
  std::map<std::vector<int>, std::unique_ptr<CachedSched>> batch_scheds;
  DevBuf geoms, bmats, Tbatch, Wbatch, VDbatch, VHbatch;
  std::map<int, std::unique_ptr<struct BatchPlan>> plans;  // prepared batches
  int next_plan = 1;
};

std::mutex g_mu;
std::map<int, std::unique_ptr<Context>> g_ctx;

int get_ctx(Context **out) {
  int dev = 0;
  HQ_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  auto &c = g_ctx[dev];
  if (!c) {
    c.reset(new Context());
    c->device = dev;
    HQ_HIP(hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, dev));
    HQ_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HQ_HIP(hipEventCreateWithFlags(&c->busy, hipEventDisableTiming));
    if (c->counter.reserve(256) != HQ_OK) return g_err;
    HQ_HIP(hipMemset(c->counter.p, 0, 256));
  }
  *out = c.get();
  return HQ_OK;
}

// Order this call's launches on `st` after the previous call's (see busy).
int order_in(Context *c, hipStream_t st) {
  HQ_HIP(hipStreamWaitEvent(st, c->busy, 0));
  return HQ_OK;
}
int order_out(Context *c, hipStream_t st) {
  HQ_HIP(hipEventRecord(c->busy, st));
  return HQ_OK;
}

// After the work of a call has completed: a dependency wait of the scheduler
// that timed out (counter[1], the task index + 1) or a bounded wait inside a
// task that gave up (the device fault word) is reported and cleared, so the
// next call starts clean.
int take_device_faults(int *counter) {
  int st[2] = {0, 0};
  HQ_HIP(hipMemcpy(st, counter, sizeof st, hipMemcpyDeviceToHost));
  int fault = 0;
  HQ_HIP(hq_fault_take(&fault));
  if (st[1] != 0) {
    HQ_HIP(hipMemset(counter + 1, 0, sizeof(int)));
    return set_error(HQ_ERR_DEVICE, "scheduler dependency wait timed out at task %d", st[1] - 1);
  }
  if (fault != 0)
    return set_error(HQ_ERR_DEVICE, "a panel's reflector wait timed out (awaited count %d)", fault - 1);
  return HQ_OK;
}

// NULL is HIP's default (null) stream, as for every HIP API; the host-pointer
// entry points pass the library's own stream explicitly.
hipStream_t pick_stream(Context *, void *s) { return (hipStream_t)s; }

// flags: one int per task, and the progress words (hq_geom::prog), one
// 64-bit word per task, zero-initialised together; epochs start at 1 (both
// hold epochs, so both are cleared whenever the epoch restarts)
int reserve_flags(Context *c, int n) {
  size_t fbytes = sizeof(int) * (size_t)n, pbytes = sizeof(unsigned long long) * (size_t)n;
  if (c->flags.bytes < fbytes || c->prog.bytes < pbytes) {
    if (c->flags.reserve(fbytes) != HQ_OK || c->prog.reserve(pbytes) != HQ_OK) return g_err;
    HQ_HIP(hipMemset(c->flags.p, 0, c->flags.bytes));
    HQ_HIP(hipMemset(c->prog.p, 0, c->prog.bytes));
    c->epoch = 0;
  }
  return HQ_OK;
}

// Claim order of tasks tl (canonical order, dependencies as indices into
// tl) from the list-scheduling simulation; dependencies remapped.
std::vector<hq_task> claim_order(std::vector<hq_task> &&tl, int nb, int ib, int P, int prec = 0) {
  DagBuilder b(1, 1, nb, ib, false, 0, true, prec);
  b.t = std::move(tl);
  std::vector<int> ord = b.order(P);
  const int n = (int)b.t.size();
  std::vector<int> pos(n);
  for (int i = 0; i < n; i++) pos[ord[i]] = i;
  std::vector<hq_task> out(n);
  for (int i = 0; i < n; i++) {
    hq_task x = b.t[ord[i]];
    for (int d = 0; d < x.ndep; d++) x.dep[d] = pos[x.dep[d]];
    out[i] = x;
  }
  return out;
}

// The task list the device scheduler runs, in claim order (list scheduling
// on P workers), dependencies as claim indices; in a 16- / 32-column apply an
// SSRFB below the first waits for the previous SSRFB of its chain (its
// dep[0], build_apply) sub-block by sub-block, inside the task
// (hq_geom::prog): that dependency is marked -2 - index.
std::vector<hq_task> claim_list(int MT, int NT, int nb, int ib, int XNT, int prec, int uw, int ncol, int P) {
  DagBuilder b(MT, NT, nb, ib, XNT >= 1, XNT, true, prec, uw, ncol);
  b.build();
  std::vector<int> ord = b.order(P);
  const int n = (int)b.t.size();
  std::vector<int> pos(n);
  for (int i = 0; i < n; i++) pos[ord[i]] = i;
  std::vector<hq_task> tl(n);
  for (int i = 0; i < n; i++) {
    hq_task x = b.t[ord[i]];
    for (int d = 0; d < x.ndep; d++) x.dep[d] = pos[x.dep[d]];
    if (XNT >= 1 && (uw == 16 || uw == 32) && x.kind == HQ_SSRFB && x.r > x.k + 1 && x.ndep > 0)
      x.dep[0] = -2 - x.dep[0];
    tl[i] = x;
  }
  return tl;
}

int get_schedule(Context *c, int MT, int NT, int nb, int ib, CachedSched **out, int XNT = -1, int prec = 0,
                 int uw = 0, int ncol = 0) {
  // XNT < 0: factorisation DAG (prec: precision of its updates); XNT >= 1:
  // Q^T application to MT x XNT tiles (FP64) in slabs of uw columns (0: the
  // factorisation's) over X's first ncol columns (0: all)
  if (XNT >= 1) prec = 0;
  else uw = ncol = 0;
  auto key = std::make_tuple(MT, NT, nb, ib, XNT, prec, uw, ncol);
  auto it = c->scheds.find(key);
  if (it != c->scheds.end()) {
    *out = it->second.get();
    return HQ_OK;
  }
  const int grid = c->num_cu > 0 ? c->num_cu : 256;
  std::vector<hq_task> tl = claim_list(MT, NT, nb, ib, XNT, prec, uw, ncol, grid);
  const int n = (int)tl.size();
  std::unique_ptr<CachedSched> cs(new CachedSched());
  if (cs->tasks.reserve(sizeof(hq_task) * (size_t)n) != HQ_OK) return g_err;
  HQ_HIP(hipMemcpy(cs->tasks.p, tl.data(), sizeof(hq_task) * (size_t)n, hipMemcpyHostToDevice));
  cs->ntasks = n;
  cs->grid = std::min(grid, n);
  if (reserve_flags(c, n) != HQ_OK) return g_err;
  *out = cs.get();
  c->scheds[key] = std::move(cs);
  return HQ_OK;
}

int host_schedule(int MT, int NT, int nb, int ib, int P, std::vector<hq_task> &tl) {
  DagBuilder b(MT, NT, nb, ib);
  b.build();
  std::vector<int> ord = b.order(P);
  tl.resize(ord.size());
  for (size_t i = 0; i < ord.size(); i++) tl[i] = b.t[ord[i]];
  return HQ_OK;
}

int ensure_T(Context *c, int MT, int NT, int nb, int ib, double **T) {
  if (*T) return HQ_OK;
  size_t bytes = sizeof(double) * (size_t)MT * NT * ib * nb;
  if (c->T.reserve(bytes) != HQ_OK) return g_err;
  *T = (double *)c->T.p;
  return HQ_OK;
}

int tile_list(Context *c, int MT, int NT, const int2 **out, int *n);

// Library scratch of a factorisation (G.ib: the inner block width it runs
// with), reserved and entered into G.
int reserve_scratch(Context *c, hq_geom &G) {
  const int MT = G.MT, NT = G.NT, nb = G.nb, ik = G.ib;
  if (!G.prec && nb >= 128 && tuned_nb(nb) && ik % 16 == 0) {
    // the GEQRT panels also write W = V T16 and the unit lower V of the
    // diagonal tiles; LARFB then needs no T multiplication between passes
    const size_t tiles = sizeof(double) * (size_t)nb * nb * std::min(MT, NT);
    if (c->W.reserve(tiles) != HQ_OK || c->VD.reserve(tiles) != HQ_OK) return g_err;
    G.W = (double *)c->W.p;
    G.VD = (double *)c->VD.p;
  }
  if (G.prec == 2) {
    if (c->VH.reserve(2 * hq_vh_tile(nb, ik) * MT * NT) != HQ_OK) return g_err;
    G.VH = (unsigned short *)c->VH.p;
    if (c->CE.reserve(sizeof(int) * (size_t)NT * nb) != HQ_OK) return g_err;
    G.CE = (const int *)c->CE.p;
  }
  return HQ_OK;
}

int tile_qr_dev_impl(int MT, int NT, int nb, int ib, double *A, double *TAU, double *T,
                     void *stream, bool serial, long long *trace = nullptr, int prec = 0, int ncol = 0) {
  if (!geometry_ok(MT, NT, nb, ib)) return g_err;
  if (prec && !rp_nb(nb)) return set_error(HQ_ERR_ARG, "reduced-precision updates need nb = 128 or 256 (nb=%d)", nb);
  if (!A || !TAU) return set_error(HQ_ERR_ARG, "A and TAU must be device pointers");
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  hipStream_t st = pick_stream(c, stream);
  // the run itself uses the inner block width; a caller's T for a wider ib
  // is rebuilt from V and TAU after it
  const int ik = ibi(ib);
  double *Tout = ik != ib ? T : nullptr;
  if (ik != ib) T = nullptr;
  if (ensure_T(c, MT, NT, nb, ik, &T) != HQ_OK) return g_err;
  if (order_in(c, st) != HQ_OK) return g_err;
  hq_geom G{A, TAU, T, MT, NT, nb, ik, A, NT, prec};
  G.ncol = ncol;
  if (reserve_scratch(c, G) != HQ_OK) return g_err;
  if (G.CE) HQ_HIP(hq_launch_colscale(A, MT, NT, nb, (int *)G.CE, st));
  if (serial) {
    std::vector<hq_task> tl;
    host_schedule(MT, NT, nb, ik, c->num_cu, tl);
    for (const hq_task &tk : tl) HQ_HIP(hq_launch_task(G, tk, st));
  } else {
    CachedSched *cs;
    if (get_schedule(c, MT, NT, nb, ik, &cs, -1, prec) != HQ_OK) return g_err;
    c->epoch++;
    HQ_HIP(hq_launch_sched(G, nullptr, (const hq_task *)cs->tasks.p, cs->ntasks, (int *)c->flags.p,
                           c->epoch, (int *)c->counter.p, cs->grid, trace, st, nullptr, c->wait_limit));
  }
  if (Tout) {
    const int2 *tiles;
    int n;
    if (tile_list(c, MT, NT, &tiles, &n) != HQ_OK) return g_err;
    hq_geom Go{A, TAU, Tout, MT, NT, nb, ib, nullptr, 0};
    HQ_HIP(hq_launch_tbuild_wide(Go, tiles, n, 0, st));
  }
  if (order_out(c, st) != HQ_OK) return g_err;
  return ok();
}

int tile_list(Context *c, int MT, int NT, const int2 **out, int *n) {
  auto key = std::make_pair(MT, NT);
  auto &slot = c->tilelists[key];
  std::vector<int2> v;
  int KT = std::min(MT, NT);
  for (int k = 0; k < KT; k++)
    for (int r = k; r < MT; r++) v.push_back(make_int2(r, k));
  if (!slot) {
    slot.reset(new DevBuf());
    if (slot->reserve(sizeof(int2) * v.size()) != HQ_OK) return g_err;
    HQ_HIP(hipMemcpy(slot->p, v.data(), sizeof(int2) * v.size(), hipMemcpyHostToDevice));
  }
  *out = (const int2 *)slot->p;
  *n = (int)v.size();
  return HQ_OK;
}

int build_t_impl(Context *c, int MT, int NT, int nb, int ib, const double *RY, const double *TAU,
                 double *T, hipStream_t st) {
  const int2 *tiles;
  int n;
  if (tile_list(c, MT, NT, &tiles, &n) != HQ_OK) return g_err;
  hq_geom G{(double *)RY, (double *)TAU, T, MT, NT, nb, ib, nullptr, 0};
  if (ib > HQ_IBMAX)
    HQ_HIP(hq_launch_tbuild_wide(G, tiles, n, 1, st));
  else
    HQ_HIP(hq_launch_tbuild(G, tiles, n, st));
  return HQ_OK;
}

// Task width of a Q^T application to ncol columns on the narrow tasks: one
// 16-column strip per chain up to 32 columns (the most chains, the chain's
// pace is the run), two strips sharing each V and T operand above that (the
// workgroups are then the limit: 64 columns 4.7 against 5.3 ms, 256 columns
// 13.6 against 18.3 ms at 65536 rows after a 65536 x 4096 factorisation);
// two strips need nb <= 256.
int narrow_width(int ncol, int nb) { return ncol > 32 && nb <= 256 ? 32 : 16; }

// X (MT x XNT tiles) <- Q^T X, Q from the reflectors of RY (MT x NT tiles).
// A caller's T is applied with its own ib (the update kernels take the 16 x 16
// diagonal blocks of T for any ib); without T (or with nb = 64, whose update
// kernel stages whole ib blocks of at most HQ_IBMAX), T is rebuilt from RY
// and TAU with the inner block width.
// ncol > 0: only X's first ncol columns hold entries (the rest hold zeros
// and still do afterwards); width 16 runs the updates as 16-column tasks on
// the generic kernels (hq_geom::uw), 0 the factorisation's slabs.
int apply_qt_impl(int MT, int NT, int nb, int ib, const double *RY, const double *TAU,
                  const double *T, double *X, int XNT, void *stream, int ncol = 0, int width = 0,
                  long long *trace = nullptr) {
  if (!geometry_ok(MT, NT, nb, ib)) return g_err;
  if (XNT < 1) return set_error(HQ_ERR_ARG, "X needs at least one tile column");
  if (!RY || !TAU || !X) return set_error(HQ_ERR_ARG, "RY, TAU and X must be device pointers");
  // a whole X of at most HQ_NARROW_COLS - 1 columns (dormqr on one tile
  // column, dorgqr of a single tile column): the narrow tasks too
  if (width == 0 && ncol == 0 && XNT * nb < HQ_NARROW_COLS && nb % 16 == 0) width = narrow_width(XNT * nb, nb);
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  hipStream_t st = pick_stream(c, stream);
  if (order_in(c, st) != HQ_OK) return g_err;
  double *Tw = (double *)T;
  int iq = ib;
  const bool narrow = width == 16 || width == 32;
  if (Tw && (nb < 128 || !tuned_nb(nb) || narrow) && ib > HQ_IBMAX) Tw = nullptr;
  if (!Tw) {
    iq = ibi(ib);
    if (ensure_T(c, MT, NT, nb, iq, &Tw) != HQ_OK) return g_err;
    if (build_t_impl(c, MT, NT, nb, iq, RY, TAU, Tw, st) != HQ_OK) return g_err;
  }
  CachedSched *cs;
  if (get_schedule(c, MT, NT, nb, iq, &cs, XNT, 0, width, ncol) != HQ_OK) return g_err;
  hq_geom G{(double *)RY, (double *)TAU, Tw, MT, NT, nb, iq, X, XNT};
  G.ncol = ncol;
  G.uw = width;
  G.prog = narrow ? (unsigned long long *)c->prog.p : nullptr;
  c->epoch++;
  HQ_HIP(hq_launch_sched(G, nullptr, (const hq_task *)cs->tasks.p, cs->ntasks, (int *)c->flags.p,
                         c->epoch, (int *)c->counter.p, cs->grid, trace, st, nullptr, c->wait_limit));
  if (order_out(c, st) != HQ_OK) return g_err;
  return ok();
}

// Batch of independent factorisations, one persistent launch: every
// matrix's DAG (the single-matrix DAG, canonical order) is concatenated with
// its matrix index and the whole batch gets one claim order, so large
// matrices' critical paths start first and small ones fill the gaps.
// Panel mapping: split_panels 0 runs each GEQRT / TSQRT as one task per tile
// (one workgroup owns the tile's whole panel QR), 1 as one task per ib block
// plus one trailing task per (block, 64-column slab), as for one matrix.
int get_batch_schedule(Context *c, int count, int nb, int ib, const int *MT, const int *NT,
                       CachedSched **out, int prec = 0, int split_panels = 0) {
  std::vector<int> key;
  key.reserve(4 + 2 * (size_t)count);
  key.push_back(nb);
  key.push_back(ib);
  key.push_back(prec);
  key.push_back(split_panels);
  for (int i = 0; i < count; i++) {
    key.push_back(MT[i]);
    key.push_back(NT[i]);
  }
  auto it = c->batch_scheds.find(key);
  if (it != c->batch_scheds.end()) {
    *out = it->second.get();
    return HQ_OK;
  }
  std::map<std::pair<int, int>, std::vector<hq_task>> per_geom;  // canonical DAG per (MT, NT)
  std::vector<hq_task> all;
  for (int i = 0; i < count; i++) {
    auto &dag = per_geom[std::make_pair(MT[i], NT[i])];
    if (dag.empty()) {
      DagBuilder b(MT[i], NT[i], nb, ib, false, 0, split_panels != 0, prec);
      b.build();
      dag = b.t;
    }
    const int base = (int)all.size();
    for (hq_task x : dag) {
      x.mat = i;
      for (int d = 0; d < x.ndep; d++) x.dep[d] += base;
      all.push_back(x);
    }
  }
  const int grid = c->num_cu > 0 ? c->num_cu : 256;
  std::vector<hq_task> tl = claim_order(std::move(all), nb, ib, grid, prec);
  const int n = (int)tl.size();
  std::unique_ptr<CachedSched> cs(new CachedSched());
  if (cs->tasks.reserve(sizeof(hq_task) * (size_t)n) != HQ_OK) return g_err;
  HQ_HIP(hipMemcpy(cs->tasks.p, tl.data(), sizeof(hq_task) * (size_t)n, hipMemcpyHostToDevice));
  cs->ntasks = n;
  cs->grid = std::min(grid, n);
  if (reserve_flags(c, n) != HQ_OK) return g_err;
  *out = cs.get();
  c->batch_scheds[key] = std::move(cs);
  return HQ_OK;
}

int batch_args_ok(int count, int nb, int ib, const int *MT, const int *NT, double *const *A,
                  double *const *TAU) {
  if (count < 1 || count > 65535) return set_error(HQ_ERR_ARG, "batch of %d matrices (1..65535)", count);
  if (!MT || !NT || !A || !TAU) return set_error(HQ_ERR_ARG, "MT, NT, A and TAU arrays are required");
  if (ib > HQ_IBMAX) return set_error(HQ_ERR_ARG, "batches take ib <= %d (ib=%d)", HQ_IBMAX, ib);
  for (int i = 0; i < count; i++) {
    if (!geometry_ok(MT[i], NT[i], nb, ib)) return g_err;
    if (!A[i] || !TAU[i]) return set_error(HQ_ERR_ARG, "matrix %d: A and TAU must be device pointers", i);
  }
  return HQ_OK;
}

// Per-matrix geometries of a batch.  T scratch for the matrices that bring
// none, and the diagonal tiles' W = V T16 / unit lower V (as for one matrix),
// are carved from Tb / Wb / VDb, reserved here.
int batch_geoms(int count, int nb, int ib, const int *MT, const int *NT, double *const *A,
                double *const *TAU, double *const *T, DevBuf &Tb, DevBuf &Wb, DevBuf &VDb,
                std::vector<hq_geom> &g, int prec = 0, DevBuf *VHb = nullptr, const int *ncol = nullptr) {
  if (ncol)
    for (int i = 0; i < count; i++)
      if (ncol[i] <= (NT[i] - 1) * nb || ncol[i] > NT[i] * nb)
        return set_error(HQ_ERR_ARG, "matrix %d: %d columns do not end in its last tile column (NT=%d, nb=%d)", i,
                         ncol[i], NT[i], nb);
  size_t tdoubles = 0;
  for (int i = 0; i < count; i++)
    if (!T || !T[i]) tdoubles += (size_t)MT[i] * NT[i] * ib * nb;
  if (tdoubles && Tb.reserve(sizeof(double) * tdoubles) != HQ_OK) return g_err;
  const bool wpath = !prec && nb >= 128 && tuned_nb(nb) && ib % 16 == 0;
  size_t vhalves = 0;
  if (prec == 2) {
    for (int i = 0; i < count; i++) vhalves += hq_vh_tile(nb, ib) * MT[i] * NT[i];
    if (VHb->reserve(2 * vhalves) != HQ_OK) return g_err;
  }
  size_t wdoubles = 0;
  if (wpath) {
    for (int i = 0; i < count; i++) wdoubles += (size_t)std::min(MT[i], NT[i]) * nb * nb;
    if (Wb.reserve(sizeof(double) * wdoubles) != HQ_OK || VDb.reserve(sizeof(double) * wdoubles) != HQ_OK)
      return g_err;
  }
  g.resize(count);
  size_t toff = 0, woff = 0, vhoff = 0;
  for (int i = 0; i < count; i++) {
    double *Ti = T ? T[i] : nullptr;
    if (!Ti) {
      Ti = (double *)Tb.p + toff;
      toff += (size_t)MT[i] * NT[i] * ib * nb;
    }
    g[i] = hq_geom{A[i], TAU[i], Ti, MT[i], NT[i], nb, ib, A[i], NT[i], prec};
    g[i].ncol = ncol ? ncol[i] : 0;
    if (prec == 2) {
      g[i].VH = (unsigned short *)VHb->p + vhoff;
      vhoff += hq_vh_tile(nb, ib) * MT[i] * NT[i];
    }
    if (wpath) {
      g[i].W = (double *)Wb.p + woff;
      g[i].VD = (double *)VDb.p + woff;
      woff += (size_t)std::min(MT[i], NT[i]) * nb * nb;
    }
  }
  return HQ_OK;
}

int batched_qr_impl(int count, int nb, int ib, const int *MT, const int *NT, double *const *A,
                    double *const *TAU, double *const *T, void *stream, int prec = 0, int split_panels = 0,
                    const int *ncol = nullptr) {
  if (batch_args_ok(count, nb, ib, MT, NT, A, TAU) != HQ_OK) return g_err;
  if (prec < 0 || prec > 2) return set_error(HQ_ERR_ARG, "prec=%d (0 FP64, 1 FP32, 2 split)", prec);
  if (split_panels < 0 || split_panels > 1) return set_error(HQ_ERR_ARG, "panel mapping %d (0 or 1)", split_panels);
  if (prec && !rp_nb(nb)) return set_error(HQ_ERR_ARG, "reduced-precision updates need nb = 128 or 256 (nb=%d)", nb);
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  hipStream_t st = pick_stream(c, stream);
  if (order_in(c, st) != HQ_OK) return g_err;
  std::vector<hq_geom> g;
  if (batch_geoms(count, nb, ib, MT, NT, A, TAU, T, c->Tbatch, c->Wbatch, c->VDbatch, g, prec, &c->VHbatch,
                  ncol) != HQ_OK)
    return g_err;
  CachedSched *cs;
  if (get_batch_schedule(c, count, nb, ib, MT, NT, &cs, prec, split_panels) != HQ_OK) return g_err;
  if (c->geoms.reserve(sizeof(hq_geom) * (size_t)count) != HQ_OK) return g_err;
  HQ_HIP(hipMemcpyAsync(c->geoms.p, g.data(), sizeof(hq_geom) * (size_t)count, hipMemcpyHostToDevice, st));
  c->epoch++;
  HQ_HIP(hq_launch_sched(g[0], (const hq_geom *)c->geoms.p, (const hq_task *)cs->tasks.p, cs->ntasks,
                         (int *)c->flags.p, c->epoch, (int *)c->counter.p, cs->grid, nullptr, st, nullptr,
                         c->wait_limit));
  if (order_out(c, st) != HQ_OK) return g_err;
  return ok();
}

// A batch prepared once and launched many times (hq_batched_plan_dev /
// hq_batched_run_dev): everything a run reads lives in device memory owned by
// the plan, and a run enqueues two kernels and nothing else, so it can be
// captured in a graph.  Counter block: [0] claim counter, [1] time-out
// report, [2] device epoch (advanced by every run; the plan's own flags
// start at 0).
struct BatchPlan {
  CachedSched *cs = nullptr;  // task list, owned by the context's schedule cache
  hq_geom g0{};
  DevBuf geoms, flags, counter, T, W, VD;
};

int batched_plan_impl(int count, int nb, int ib, const int *MT, const int *NT, double *const *A,
                      double *const *TAU, double *const *T, int *plan, int split_panels = 0,
                      const int *ncol = nullptr) {
  if (!plan) return set_error(HQ_ERR_ARG, "plan must not be NULL");
  if (split_panels < 0 || split_panels > 1) return set_error(HQ_ERR_ARG, "panel mapping %d (0 or 1)", split_panels);
  if (batch_args_ok(count, nb, ib, MT, NT, A, TAU) != HQ_OK) return g_err;
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  std::unique_ptr<BatchPlan> p(new BatchPlan());
  std::vector<hq_geom> g;
  if (batch_geoms(count, nb, ib, MT, NT, A, TAU, T, p->T, p->W, p->VD, g, 0, nullptr, ncol) != HQ_OK) return g_err;
  if (get_batch_schedule(c, count, nb, ib, MT, NT, &p->cs, 0, split_panels) != HQ_OK) return g_err;
  if (p->geoms.reserve(sizeof(hq_geom) * (size_t)count) != HQ_OK ||
      p->flags.reserve(sizeof(int) * (size_t)p->cs->ntasks) != HQ_OK || p->counter.reserve(64) != HQ_OK)
    return g_err;
  HQ_HIP(hipMemcpy(p->geoms.p, g.data(), sizeof(hq_geom) * (size_t)count, hipMemcpyHostToDevice));
  HQ_HIP(hipMemset(p->flags.p, 0, sizeof(int) * (size_t)p->cs->ntasks));
  HQ_HIP(hipMemset(p->counter.p, 0, 64));
  p->g0 = g[0];
  std::lock_guard<std::mutex> lk(g_mu);
  const int id = c->next_plan++;
  c->plans[id] = std::move(p);
  *plan = id;
  return ok();
}

BatchPlan *find_plan(Context *c, int plan) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = c->plans.find(plan);
  return it == c->plans.end() ? nullptr : it->second.get();
}

// Layout operations on a batch (op 0: fill uniform, 1: column-major -> tile,
// 2: tile -> column-major); m x n entries inside ceil(m/nb) x ceil(n/nb)
// tiles, padding zero.
int batched_layout_impl(int op, int count, int nb, const int *m, const int *n, const uint64_t *seeds,
                        const double *const *C, const int *ldc, double *const *A, void *stream) {
  if (count < 1 || count > 65535) return set_error(HQ_ERR_ARG, "batch of %d matrices (1..65535)", count);
  if (nb < 1) return set_error(HQ_ERR_ARG, "nb=%d", nb);
  if (!m || !n || !A) return set_error(HQ_ERR_ARG, "m, n and A arrays are required");
  std::vector<hq_bmat> v(count);
  int max_tiles = 0;
  for (int i = 0; i < count; i++) {
    if (m[i] < 1 || n[i] < 1) return set_error(HQ_ERR_ARG, "matrix %d is %d x %d", i, m[i], n[i]);
    if (!A[i]) return set_error(HQ_ERR_ARG, "matrix %d: A must be a device pointer", i);
    hq_bmat b{};
    b.A = A[i];
    b.m = m[i];
    b.n = n[i];
    b.MT = (m[i] + nb - 1) / nb;
    b.NT = (n[i] + nb - 1) / nb;
    if (op == 0) {
      b.seed = seeds ? seeds[i] : 0;
    } else {
      if (!C || !C[i] || !ldc || ldc[i] < m[i])
        return set_error(HQ_ERR_ARG, "matrix %d: column-major pointer and ld >= m required", i);
      if (op >= 3)
        b.Cf = (float *)C[i];
      else
        b.C = (double *)C[i];
      b.ld = ldc[i];
    }
    max_tiles = std::max(max_tiles, b.MT * b.NT);
    v[i] = b;
  }
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  hipStream_t st = pick_stream(c, stream);
  if (order_in(c, st) != HQ_OK) return g_err;
  if (c->bmats.reserve(sizeof(hq_bmat) * (size_t)count) != HQ_OK) return g_err;
  HQ_HIP(hipMemcpyAsync(c->bmats.p, v.data(), sizeof(hq_bmat) * (size_t)count, hipMemcpyHostToDevice, st));
  HQ_HIP(hq_launch_batched_layout(op, (const hq_bmat *)c->bmats.p, count, nb, max_tiles, st));
  if (order_out(c, st) != HQ_OK) return g_err;
  return ok();
}

}  // namespace

// =====================================================================
// C-ABI
// =====================================================================
extern "C" {

int hq_last_error(void) { return g_err; }
const char *hq_last_error_message(void) { return g_msg; }
const char *hq_version(void) { return "hybrid3-mi355x 0.1 (gfx950)"; }

void mat_init(mat_stat_t *status, long int mat_m, long int mat_n, int tile_m, int tile_n,
              int my_rank, int proc_num) {
  long int mt = mat_m / tile_m;
  long int nt = mat_n / tile_n;
  long int mine = nt / proc_num + ((nt % proc_num) > my_rank ? 1 : 0);
  status->mat_M = mat_m;
  status->mat_N = mat_n;
  status->my_mat_M = mat_m;
  status->my_tile_M = (unsigned)mt;
  status->glob_tile_M = (unsigned)mt;
  status->glob_tile_N = (unsigned)nt;
  status->my_mat_N = (unsigned long)tile_n * mine;
  status->my_tile_N = (unsigned)mine;
  status->tile_M = tile_m;
  status->tile_N = tile_n;
}

int hq_tile_qr_dev(int MT, int NT, int nb, int ib, double *A, double *TAU, double *T,
                   void *stream) {
  return tile_qr_dev_impl(MT, NT, nb, ib, A, TAU, T, stream, false);
}

int hq_tile_qr_prepare_dev(int MT, int NT, int nb, int ib, int prec) {
  if (!geometry_ok(MT, NT, nb, ib)) return g_err;
  if (prec < 0 || prec > 2) return set_error(HQ_ERR_ARG, "prec=%d (0 FP64, 1 FP32, 2 split)", prec);
  if (prec && !rp_nb(nb)) return set_error(HQ_ERR_ARG, "reduced-precision updates need nb = 128 or 256 (nb=%d)", nb);
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  const int ik = ibi(ib);
  CachedSched *cs;
  if (get_schedule(c, MT, NT, nb, ik, &cs, -1, prec) != HQ_OK) return g_err;
  hq_geom G{nullptr, nullptr, nullptr, MT, NT, nb, ik, nullptr, NT, prec};
  if (reserve_scratch(c, G) != HQ_OK) return g_err;
  return ok();
}

// ---------------------------------------------------------------------------
// Allocation-free tile QR (hq_tile_qr_noalloc_dev): the scheduler's claim
// counter, device epoch, task flags and task list live in `scratch`, a
// file-scope device array of HQ_NOALLOC_SCRATCH bytes that each GPU holds
// from library load to process end.  Host code never uses the array's name
// as an address: the device address comes from hipGetSymbolAddress, taken
// once per device with that device current, and is kept per device.
// Layout: [0, 256) counter block (int 0: claim counter, 1: time-out word,
// 2: device epoch), then the flags (one int per task, 256-byte aligned), then
// the task list.  Every call shares the array of its GPU, which keeps the
// last call's contents: the epoch only grows, so flags left by earlier calls
// of the same geometry never read as complete; when the geometry changes,
// the task list is copied again and the flag range is cleared (it may cover
// the previous task list).  Calls that overlap on one GPU race.
static __device__ __align__(256) unsigned char scratch[HQ_NOALLOC_SCRATCH];

namespace {

constexpr int kMaxDevices = 64;
struct NoAllocDev {
  bool ready = false;
  unsigned char *ptr = nullptr;  // device address of `scratch` on this device
  int num_cu = 0;
  int MT = 0, NT = 0, nb = 0, ib = 0, ntasks = 0, grid = 0;  // resident task list
};
NoAllocDev g_noalloc[kMaxDevices];
std::mutex g_noalloc_mu;

size_t noalloc_flags_off() { return 256; }
size_t noalloc_tasks_off(int ntasks) { return 256 + (((size_t)ntasks * sizeof(int) + 255) & ~(size_t)255); }

int noalloc_impl(int MT, int NT, int nb, int ib, double *A, double *TAU, double *T, void *stream) {
  if (!geometry_ok(MT, NT, nb, ib)) return g_err;
  if (ib > HQ_IBMAX)
    return set_error(HQ_ERR_ARG, "the allocation-free tile QR takes ib <= %d (ib=%d)", HQ_IBMAX, ib);
  if (!A || !TAU || !T) return set_error(HQ_ERR_ARG, "A, TAU and T must be device pointers");
  int dev = 0;
  HQ_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(HQ_ERR_ARG, "device %d beyond %d", dev, kMaxDevices);
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  NoAllocDev &D = g_noalloc[dev];
  if (!D.ready) {  // set-up on this device, with it current
    void *p = nullptr;
    HQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    D.ptr = (unsigned char *)p;
    HQ_HIP(hipDeviceGetAttribute(&D.num_cu, hipDeviceAttributeMultiprocessorCount, dev));
    // counter block and every flag start at zero
    HQ_HIP(hq_launch_fill_int((int *)D.ptr, HQ_NOALLOC_SCRATCH / sizeof(int), 0, st));
    D.ready = true;
  }
  int *counter = (int *)D.ptr;
  int *flags = (int *)(D.ptr + noalloc_flags_off());
  if (D.MT != MT || D.NT != NT || D.nb != nb || D.ib != ib) {
    // the task list of this geometry in the claim order of hq_tile_qr_dev
    // (host memory for the list, this call only), copied into the array
    DagBuilder b(MT, NT, nb, ib);
    b.build();
    const int n = (int)b.t.size();
    const size_t need = noalloc_tasks_off(n) + sizeof(hq_task) * (size_t)n;
    if (need > HQ_NOALLOC_SCRATCH)
      return set_error(HQ_ERR_ARG,
                       "the allocation-free tile QR holds at most %d MiB of schedule; %d x %d tiles of nb=%d, "
                       "ib=%d need %zu bytes (%d tasks)",
                       HQ_NOALLOC_SCRATCH >> 20, MT, NT, nb, ib, need, n);
    const int grid = D.num_cu > 0 ? D.num_cu : 256;
    std::vector<int> ord = b.order(grid);
    std::vector<int> pos(ord.size());
    for (size_t i = 0; i < ord.size(); i++) pos[ord[i]] = (int)i;
    std::vector<hq_task> tl(ord.size());
    for (size_t i = 0; i < ord.size(); i++) {
      hq_task x = b.t[ord[i]];
      for (int d = 0; d < x.ndep; d++) x.dep[d] = pos[x.dep[d]];
      tl[i] = x;
    }
    // the new flag range may cover bytes of the previous geometry's task
    // list, which could equal a later epoch: clear it (epochs start at 1)
    HQ_HIP(hq_launch_fill_int(flags, (size_t)n, 0, st));
    HQ_HIP(hipMemcpyAsync(D.ptr + noalloc_tasks_off(n), tl.data(), sizeof(hq_task) * tl.size(),
                          hipMemcpyHostToDevice, st));
    HQ_HIP(hipStreamSynchronize(st));  // the host list is released on return
    D.MT = MT;
    D.NT = NT;
    D.nb = nb;
    D.ib = ib;
    D.ntasks = n;
    D.grid = std::min(grid, n);
  }
  // LARFB multiplies by T itself (no W = V T16 scratch)
  hq_geom G{A, TAU, T, MT, NT, nb, ib, A, NT, 0};
  HQ_HIP(hq_launch_sched(G, nullptr, (const hq_task *)(D.ptr + noalloc_tasks_off(D.ntasks)), D.ntasks, flags, 0,
                         counter, D.grid, nullptr, st, counter + 2));
  return ok();
}

}  // namespace

extern "C" int hq_tile_qr_noalloc_dev(int MT, int NT, int nb, int ib, double *A, double *TAU, double *T,
                                      void *stream) {
  return noalloc_impl(MT, NT, nb, ib, A, TAU, T, stream);
}

extern "C" int hq_tile_qr_noalloc_check(void *stream) {
  int dev = 0;
  HQ_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices || !g_noalloc[dev].ready)
    return set_error(HQ_ERR_ARG, "no allocation-free call was made on device %d", dev);
  HQ_HIP(hipStreamSynchronize((hipStream_t)stream));
  if (take_device_faults((int *)g_noalloc[dev].ptr) != HQ_OK) return g_err;
  return ok();
}

int hq_tile_qr_mn_dev(int m, int n, int nb, int ib, double *A, double *TAU, double *T, void *stream) {
  if (m < 1 || n < 1 || !nb_ok(nb)) return set_error(HQ_ERR_ARG, "bad m=%d n=%d nb=%d", m, n, nb);
  if (!A || !TAU) return set_error(HQ_ERR_ARG, "A and TAU must be device pointers");
  const int MT = (m + nb - 1) / nb, NT = (n + nb - 1) / nb;
  if (!geometry_ok(MT, NT, nb, ib)) return g_err;
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  hipStream_t st = pick_stream(c, stream);
  // padding entries are zero, so every reflector is zero in the padded rows
  // and the padded columns get tau = 0 (their reflector is the identity)
  if (m % nb || n % nb) HQ_HIP(hq_launch_pad_zero(A, MT, NT, nb, m, n, st));
  // tasks wholly inside the padding columns only write zeros (hq_geom::ncol)
  return tile_qr_dev_impl(MT, NT, nb, ib, A, TAU, T, stream, false, nullptr, 0, n);
}

int hq_tile_qr_mixed_dev(int MT, int NT, int nb, int ib, double *A, double *TAU, double *T,
                         void *stream) {
  return tile_qr_dev_impl(MT, NT, nb, ib, A, TAU, T, stream, false, nullptr, 1);
}

int hq_tile_qr_split_dev(int MT, int NT, int nb, int ib, double *A, double *TAU, double *T,
                         void *stream) {
  return tile_qr_dev_impl(MT, NT, nb, ib, A, TAU, T, stream, false, nullptr, 2);
}

int hq_tile_qr_dev_serial(int MT, int NT, int nb, int ib, double *A, double *TAU, double *T,
                          void *stream) {
  return tile_qr_dev_impl(MT, NT, nb, ib, A, TAU, T, stream, true);
}

int hq_tile_qr_dev_traced(int MT, int NT, int nb, int ib, double *A, double *TAU, double *T,
                          long long *trace, void *stream) {
  if (!trace) return set_error(HQ_ERR_ARG, "trace must be a device buffer of 4*ntasks int64");
  return tile_qr_dev_impl(MT, NT, nb, ib, A, TAU, T, stream, false, trace);
}

int hq_tile_qr_traced_prec_dev(int MT, int NT, int nb, int ib, int prec, double *A, double *TAU, double *T,
                               long long *trace, void *stream) {
  if (!trace) return set_error(HQ_ERR_ARG, "trace must be a device buffer of 4*ntasks int64");
  if (prec < 0 || prec > 2) return set_error(HQ_ERR_ARG, "prec=%d (0 FP64, 1 FP32, 2 split)", prec);
  return tile_qr_dev_impl(MT, NT, nb, ib, A, TAU, T, stream, false, trace, prec);
}

static int prec_ok(int prec, int nb) {
  if (prec < 0 || prec > 2) return set_error(HQ_ERR_ARG, "prec=%d (0 FP64, 1 FP32, 2 split)", prec);
  if (prec && !rp_nb(nb)) return set_error(HQ_ERR_ARG, "reduced-precision updates need nb = 128 or 256 (nb=%d)", nb);
  return HQ_OK;
}

int hq_schedule_tasks_prec(int MT, int NT, int nb, int ib, int prec, int *out, int capacity) {
  if (!geometry_ok(MT, NT, nb, ib) || prec_ok(prec, nb) != HQ_OK) return g_err;
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  CachedSched *cs;
  if (get_schedule(c, MT, NT, nb, ibi(ib), &cs, -1, prec) != HQ_OK) return g_err;
  if (capacity < cs->ntasks * 5) return set_error(HQ_ERR_ARG, "need capacity >= 5*ntasks");
  std::vector<hq_task> tl(cs->ntasks);
  HQ_HIP(hipMemcpy(tl.data(), cs->tasks.p, sizeof(hq_task) * tl.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < cs->ntasks; i++) {
    out[5 * i + 0] = tl[i].kind;
    out[5 * i + 1] = tl[i].k;
    out[5 * i + 2] = tl[i].r;
    out[5 * i + 3] = tl[i].j;
    out[5 * i + 4] = tl[i].slab;
  }
  return ok();
}

int hq_schedule_tasks(int MT, int NT, int nb, int ib, int *out, int capacity) {
  return hq_schedule_tasks_prec(MT, NT, nb, ib, 0, out, capacity);
}

int hq_dag_host_prec(int MT, int NT, int nb, int ib, int prec, int XNT, int P, int *tasks, int *deps,
                     int capacity, int *ntasks) {
  if (!geometry_ok(MT, NT, nb, ib) || prec_ok(prec, nb) != HQ_OK) return g_err;
  if (P < 1) return set_error(HQ_ERR_ARG, "P must be >= 1");
  DagBuilder b(MT, NT, nb, ibi(ib), XNT >= 1, XNT, true, prec);
  b.build();
  const int n = (int)b.t.size();
  if (ntasks) *ntasks = n;
  if (!tasks || !deps) return ok();
  if (capacity < n) return set_error(HQ_ERR_ARG, "capacity %d < %d tasks", capacity, n);
  std::vector<int> ord = b.order(P), pos(n);
  for (int i = 0; i < n; i++) pos[ord[i]] = i;
  for (int i = 0; i < n; i++) {
    const hq_task &x = b.t[ord[i]];
    tasks[5 * i + 0] = x.kind;
    tasks[5 * i + 1] = x.k;
    tasks[5 * i + 2] = x.r;
    tasks[5 * i + 3] = x.j;
    tasks[5 * i + 4] = x.slab;
    for (int d = 0; d < HQ_MAXDEP; d++) deps[HQ_MAXDEP * i + d] = d < x.ndep ? pos[x.dep[d]] : -1;
  }
  return ok();
}

int hq_dag_host(int MT, int NT, int nb, int ib, int XNT, int P, int *tasks, int *deps,
                int capacity, int *ntasks) {
  return hq_dag_host_prec(MT, NT, nb, ib, 0, XNT, P, tasks, deps, capacity, ntasks);
}

int hq_dag_host_apply(int MT, int NT, int nb, int ib, int XNT, int ncol, int width, int P, int *tasks, int *deps,
                      int capacity, int *ntasks) {
  if (!geometry_ok(MT, NT, nb, ib)) return g_err;
  if (XNT < 1) return set_error(HQ_ERR_ARG, "X needs at least one tile column");
  if (ncol < 0 || ncol > XNT * nb) return set_error(HQ_ERR_ARG, "ncol=%d outside X (%d columns)", ncol, XNT * nb);
  if (width != 0 && width != 16 && width != 32) return set_error(HQ_ERR_ARG, "width %d (0, 16 or 32)", width);
  if (P < 1) return set_error(HQ_ERR_ARG, "P must be >= 1");
  const std::vector<hq_task> tl = claim_list(MT, NT, nb, ibi(ib), XNT, 0, width, ncol, P);
  const int n = (int)tl.size();
  if (ntasks) *ntasks = n;
  if (!tasks || !deps) return ok();
  if (capacity < n) return set_error(HQ_ERR_ARG, "capacity %d < %d tasks", capacity, n);
  for (int i = 0; i < n; i++) {
    const hq_task &x = tl[i];
    tasks[5 * i + 0] = x.kind;
    tasks[5 * i + 1] = x.k;
    tasks[5 * i + 2] = x.r;
    tasks[5 * i + 3] = x.j;
    tasks[5 * i + 4] = x.slab;
    for (int d = 0; d < HQ_MAXDEP; d++) deps[HQ_MAXDEP * i + d] = d < x.ndep ? x.dep[d] : -1;
  }
  return ok();
}

int hq_wall_clock_khz(int *khz) {
  int dev = 0;
  HQ_HIP(hipGetDevice(&dev));
  HQ_HIP(hipDeviceGetAttribute(khz, hipDeviceAttributeWallClockRate, dev));
  return ok();
}

int hq_schedule_info_prec(int MT, int NT, int nb, int ib, int prec, int *ntasks, int *grid) {
  if (!geometry_ok(MT, NT, nb, ib) || prec_ok(prec, nb) != HQ_OK) return g_err;
  ib = ibi(ib);
  const int KT = std::min(MT, NT), S = (nb + HQ_SLAB - 1) / HQ_SLAB;
  const int UW = hq_upd_slab(nb, prec), SU = (nb + UW - 1) / UW;
  // tasks of one panel QR (GEQRT or TSQRT): a panel per ib block and a
  // trailing task per slab right of the block
  long per_tile = 0;
  for (int ii = 0; ii < nb; ii += ib) {
    const int j0 = std::min(nb, ii + ib);
    per_tile += 1 + (j0 < nb ? S - j0 / HQ_SLAB : 0);
  }
  long n = 0;
  for (int k = 0; k < KT; k++)
    n += (MT - k) * per_tile + (long)(NT - k - 1) * SU + (long)(NT - k - 1) * (MT - k - 1) * SU;
  if (ntasks) *ntasks = (int)n;
  if (grid) {
    int dev = 0, cu = 256;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
      *grid = (int)std::min<long>(cu, n);
    else
      *grid = (int)std::min<long>(256, n);
  }
  return ok();
}

int hq_schedule_info(int MT, int NT, int nb, int ib, int *ntasks, int *grid) {
  return hq_schedule_info_prec(MT, NT, nb, ib, 0, ntasks, grid);
}

// Single-tile kernels: any 1 <= ib <= nb.  The panels (GEQRT / TSQRT) run
// with the inner block width and T for a wider ib is rebuilt from V and TAU;
// the updates take T with its own ib, except at nb = 64, whose update stages
// whole ib blocks (ib <= HQ_IBMAX there).
static int tile_args_ok(int nb, int ib, bool consumer) {
  if (!nb_ok(nb) || ib < 1 || ib > nb) return set_error(HQ_ERR_ARG, "unsupported nb=%d / ib=%d", nb, ib);
  if (consumer && (nb < 128 || !tuned_nb(nb)) && ib > HQ_IBMAX)
    return set_error(HQ_ERR_ARG, "nb=%d takes ib <= %d in LARFB / SSRFB (ib=%d)", nb, HQ_IBMAX, ib);
  return HQ_OK;
}

// GEQRT (ts = false: A1 = the tile) or TSQRT (A1 on A2) of one tile.
static int panel_tile(int nb, int ib, double *A1, double *A2, double *T, double *TAU, void *stream, bool ts) {
  if (tile_args_ok(nb, ib, false) != HQ_OK) return g_err;
  if (!A1 || !T || !TAU || (ts && !A2)) return set_error(HQ_ERR_ARG, "tile, T and TAU must be device pointers");
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  hipStream_t st = pick_stream(c, stream);
  const int ik = ibi(ib);
  double *Tk = T;
  if (ik != ib) {  // inner T in the context's scratch, then T for ib
    if (order_in(c, st) != HQ_OK) return g_err;
    if (c->T1.reserve(sizeof(double) * (size_t)nb * ik) != HQ_OK) return g_err;
    Tk = (double *)c->T1.p;
  }
  if (ts)
    HQ_HIP(hq_launch_tsqrt(nb, ik, A1, A2, Tk, TAU, st));
  else
    HQ_HIP(hq_launch_geqrt(nb, ik, A1, Tk, TAU, st));
  if (ik != ib) {
    HQ_HIP(hq_launch_tbuild_wide_tile(nb, ib, ts ? A2 : A1, TAU, T, ts ? 0 : 1, st));
    if (order_out(c, st) != HQ_OK) return g_err;
  }
  return ok();
}

int hq_dgeqrt_dev(int nb, int ib, double *A, double *T, double *TAU, void *stream) {
  return panel_tile(nb, ib, A, nullptr, T, TAU, stream, false);
}
int hq_dtsqrt_dev(int nb, int ib, double *A1, double *A2, double *T, double *TAU, void *stream) {
  return panel_tile(nb, ib, A1, A2, T, TAU, stream, true);
}
int hq_dormqr_dev(int nb, int ib, const double *V, const double *T, double *C, void *stream) {
  if (tile_args_ok(nb, ib, true) != HQ_OK) return g_err;
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  HQ_HIP(hq_launch_larfb(nb, ib, V, T, C, pick_stream(c, stream)));
  return ok();
}
int hq_dtsmqr_dev(int nb, int ib, double *A1, double *A2, const double *V, const double *T,
                  void *stream) {
  if (tile_args_ok(nb, ib, true) != HQ_OK) return g_err;
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  HQ_HIP(hq_launch_ssrfb(nb, ib, A1, A2, V, T, pick_stream(c, stream)));
  return ok();
}

int hq_tile_build_t_dev(int MT, int NT, int nb, int ib, const double *RY, const double *TAU,
                        double *T, void *stream) {
  if (!geometry_ok(MT, NT, nb, ib)) return g_err;
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  if (build_t_impl(c, MT, NT, nb, ib, RY, TAU, T, pick_stream(c, stream)) != HQ_OK) return g_err;
  return ok();
}

int hq_tile_ormqr_dev(int MT, int NT, int nb, int ib, const double *RY, const double *TAU,
                      const double *T, double *X, int XNT, void *stream) {
  return apply_qt_impl(MT, NT, nb, ib, RY, TAU, T, X, XNT, stream);
}

// Narrow X: width 0 picks 16- or 32-column tasks (narrow_width) when X
// holds fewer than HQ_NARROW_COLS columns (a short block of right-hand
// sides, whose few column chains along the flat tree are the whole critical
// path), else the factorisation's slabs.
static int ormqr_cols(int MT, int NT, int nb, int ib, const double *RY, const double *TAU, const double *T,
                      double *X, int XNT, int ncol, int width, long long *trace, void *stream) {
  if (XNT < 1) return set_error(HQ_ERR_ARG, "X needs at least one tile column");
  if (ncol <= (XNT - 1) * nb || ncol > XNT * nb)
    return set_error(HQ_ERR_ARG, "ncol=%d outside tile column %d of X (nb=%d)", ncol, XNT, nb);
  if (width != 0 && width != 16 && width != 32 && width != 64)
    return set_error(HQ_ERR_ARG, "update width %d (0, 16, 32 or 64)", width);
  if (width == 32 && nb > 256) return set_error(HQ_ERR_ARG, "32-column updates need nb <= 256 (nb=%d)", nb);
  if (width == 0) width = ncol < HQ_NARROW_COLS ? narrow_width(ncol, nb) : 64;
  return apply_qt_impl(MT, NT, nb, ib, RY, TAU, T, X, XNT, stream, ncol, width == 64 ? 0 : width, trace);
}

int hq_tile_ormqr_cols_dev(int MT, int NT, int nb, int ib, const double *RY, const double *TAU,
                           const double *T, double *X, int XNT, int ncol, int width, void *stream) {
  return ormqr_cols(MT, NT, nb, ib, RY, TAU, T, X, XNT, ncol, width, nullptr, stream);
}

int hq_tile_ormqr_cols_traced_dev(int MT, int NT, int nb, int ib, const double *RY, const double *TAU,
                                  const double *T, double *X, int XNT, int ncol, int width, long long *trace,
                                  void *stream) {
  if (!trace) return set_error(HQ_ERR_ARG, "trace must be a device buffer of 4*ntasks int64");
  return ormqr_cols(MT, NT, nb, ib, RY, TAU, T, X, XNT, ncol, width, trace, stream);
}

int hq_tile_orgqr_dev(int MT, int NT, int nb, int ib, const double *RY, const double *TAU,
                      const double *T, double *Q, void *stream) {
  if (!geometry_ok(MT, NT, nb, ib)) return g_err;
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  HQ_HIP(hq_launch_identity(Q, MT, NT, nb, 1, pick_stream(c, stream)));
  return apply_qt_impl(MT, NT, nb, ib, RY, TAU, T, Q, NT, stream);
}

int hq_tile_make_q_dev(int MT, int NT, int nb, int ib, double *RY, const double *TAU, double *Q,
                       void *stream) {
  if (!geometry_ok(MT, NT, nb, ib)) return g_err;
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  hipStream_t st = pick_stream(c, stream);
  HQ_HIP(hq_launch_identity(Q, MT, NT, nb, 0, st));
  if (apply_qt_impl(MT, NT, nb, ib, RY, TAU, nullptr, Q, NT, stream) != HQ_OK) return g_err;
  HQ_HIP(hq_launch_zero_reflectors(RY, MT, NT, nb, st));
  return ok();
}

int hq_batched_tile_qr_dev(int count, int nb, int ib, const int *MT, const int *NT, double *const *A,
                           double *const *TAU, double *const *T, void *stream) {
  return batched_qr_impl(count, nb, ib, MT, NT, A, TAU, T, stream);
}

int hq_batched_prepare_dev(int prec, int count, int nb, int ib, const int *MT, const int *NT) {
  return hq_batched_prepare_ex_dev(prec, 0, count, nb, ib, MT, NT);
}

int hq_batched_prepare_ex_dev(int prec, int split_panels, int count, int nb, int ib, const int *MT,
                              const int *NT) {
  if (split_panels < 0 || split_panels > 1) return set_error(HQ_ERR_ARG, "panel mapping %d (0 or 1)", split_panels);
  if (count < 1 || count > 65535) return set_error(HQ_ERR_ARG, "batch of %d matrices (1..65535)", count);
  if (!MT || !NT) return set_error(HQ_ERR_ARG, "MT and NT arrays are required");
  if (ib > HQ_IBMAX) return set_error(HQ_ERR_ARG, "batches take ib <= %d (ib=%d)", HQ_IBMAX, ib);
  if (prec < 0 || prec > 2) return set_error(HQ_ERR_ARG, "prec=%d (0 FP64, 1 FP32, 2 split)", prec);
  if (prec && !rp_nb(nb)) return set_error(HQ_ERR_ARG, "reduced-precision updates need nb = 128 or 256 (nb=%d)", nb);
  for (int i = 0; i < count; i++)
    if (!geometry_ok(MT[i], NT[i], nb, ib)) return g_err;
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  CachedSched *cs;
  if (get_batch_schedule(c, count, nb, ib, MT, NT, &cs, prec, split_panels) != HQ_OK) return g_err;
  return ok();
}

int hq_batched_tile_qr_prec_dev(int prec, int count, int nb, int ib, const int *MT, const int *NT,
                                double *const *A, double *const *TAU, double *const *T, void *stream) {
  return batched_qr_impl(count, nb, ib, MT, NT, A, TAU, T, stream, prec);
}

int hq_batched_tile_qr_ex_dev(int prec, int split_panels, int count, int nb, int ib, const int *MT,
                              const int *NT, const int *ncol, double *const *A, double *const *TAU,
                              double *const *T, void *stream) {
  return batched_qr_impl(count, nb, ib, MT, NT, A, TAU, T, stream, prec, split_panels, ncol);
}

int hq_batched_plan_dev(int count, int nb, int ib, const int *MT, const int *NT, double *const *A,
                        double *const *TAU, double *const *T, int *plan) {
  return batched_plan_impl(count, nb, ib, MT, NT, A, TAU, T, plan);
}

int hq_batched_plan_ex_dev(int split_panels, int count, int nb, int ib, const int *MT, const int *NT,
                           const int *ncol, double *const *A, double *const *TAU, double *const *T, int *plan) {
  return batched_plan_impl(count, nb, ib, MT, NT, A, TAU, T, plan, split_panels, ncol);
}

int hq_batched_run_dev(int plan, void *stream) {
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  BatchPlan *p = find_plan(c, plan);
  if (!p) return set_error(HQ_ERR_ARG, "no batch plan %d", plan);
  int *cnt = (int *)p->counter.p;
  HQ_HIP(hq_launch_sched(p->g0, (const hq_geom *)p->geoms.p, (const hq_task *)p->cs->tasks.p, p->cs->ntasks,
                         (int *)p->flags.p, 0, cnt, p->cs->grid, nullptr, pick_stream(c, stream), cnt + 2));
  return ok();
}

int hq_batched_plan_check(int plan, void *stream) {
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  BatchPlan *p = find_plan(c, plan);
  if (!p) return set_error(HQ_ERR_ARG, "no batch plan %d", plan);
  HQ_HIP(hipStreamSynchronize(pick_stream(c, stream)));
  if (take_device_faults((int *)p->counter.p) != HQ_OK) return g_err;
  return ok();
}

int hq_batched_plan_free(int plan) {
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  HQ_HIP(hipDeviceSynchronize());
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c->plans.erase(plan)) return set_error(HQ_ERR_ARG, "no batch plan %d", plan);
  return ok();
}

int hq_batched_fill_uniform_dev(int count, int nb, const int *m, const int *n, const uint64_t *seeds,
                                double *const *A, void *stream) {
  return batched_layout_impl(0, count, nb, m, n, seeds, nullptr, nullptr, A, stream);
}

int hq_cm_to_tile_dev(int count, int nb, const int *m, const int *n, const double *const *C,
                      const int *ldc, double *const *A, void *stream) {
  return batched_layout_impl(1, count, nb, m, n, nullptr, C, ldc, A, stream);
}

int hq_cm32_to_tile_dev(int count, int nb, const int *m, const int *n, const float *const *C, const int *ldc,
                        double *const *A, void *stream) {
  return batched_layout_impl(3, count, nb, m, n, nullptr, (const double *const *)C, ldc, A, stream);
}

int hq_tile_to_cm32_dev(int count, int nb, const int *m, const int *n, const double *const *A, float *const *C,
                        const int *ldc, void *stream) {
  return batched_layout_impl(4, count, nb, m, n, nullptr, (const double *const *)C, ldc, (double *const *)A,
                             stream);
}

int hq_tile_to_cm_dev(int count, int nb, const int *m, const int *n, const double *const *A, double *const *C,
                      const int *ldc, void *stream) {
  return batched_layout_impl(2, count, nb, m, n, nullptr, (const double *const *)C, ldc,
                             (double *const *)A, stream);
}

int hq_fill_uniform_dev(int MT, int NT, int nb, uint64_t seed, double *A, void *stream) {
  if (MT < 1 || NT < 1 || nb < 1 || !A) return set_error(HQ_ERR_ARG, "bad fill arguments");
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  HQ_HIP(hq_launch_fill_uniform(A, MT, NT, nb, seed, pick_stream(c, stream)));
  return ok();
}

int hq_set_wait_limit(unsigned spins) {
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  c->wait_limit = spins ? spins : 1u << 22;
  return ok();
}

int hq_sync(void *stream) {
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  HQ_HIP(hipStreamSynchronize(pick_stream(c, stream)));
  // counter[1] is raised by the scheduler when a dependency wait timed out,
  // the fault word by a panel wait that gave up
  if (take_device_faults((int *)c->counter.p) != HQ_OK) return g_err;
  return ok();
}

void hq_finalize(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : g_ctx) {
    if (kv.second && kv.second->stream) (void)hipStreamSynchronize(kv.second->stream);
  }
  g_ctx.clear();
}

// ------------------------------------------------------------ host-pointer API
static int check_rank(int my_rank, int Proc_num) {
  if (my_rank != 0 || Proc_num != 1)
    return set_error(HQ_ERR_RANK,
                     "the engine runs on one GPU without ranks: my_rank=%d Proc_num=%d "
                     "(need 0 and 1)",
                     my_rank, Proc_num);
  return HQ_OK;
}

static int report(const char *fn) {
  if (g_err != HQ_OK) fprintf(stderr, "hybrid3: %s: %s\n", fn, g_msg);
  return g_err;
}

void Hybrid_tile_QR(mat_stat_t *status, double *A, double *TAU, int my_rank, int Proc_num,
                    int ib) {
  if (check_rank(my_rank, Proc_num) != HQ_OK) {
    report("Hybrid_tile_QR");
    return;
  }
  if (!status || !A || !TAU || status->tile_M != status->tile_N) {
    set_error(HQ_ERR_ARG, "need square tiles and non-null A / TAU");
    report("Hybrid_tile_QR");
    return;
  }
  const int MT = (int)status->glob_tile_M, NT = (int)status->glob_tile_N;
  const int nb = (int)status->tile_M;
  if (!geometry_ok(MT, NT, nb, ib)) {
    report("Hybrid_tile_QR");
    return;
  }
  Context *c;
  if (get_ctx(&c) != HQ_OK) {
    report("Hybrid_tile_QR");
    return;
  }
  const size_t nA = (size_t)MT * NT * nb * nb, nT = (size_t)MT * NT * nb;
  std::lock_guard<std::mutex> lk(c->host_mu);
  auto run = [&]() -> int {
    if (c->A.reserve(nA * sizeof(double)) != HQ_OK) return g_err;
    if (c->TAU.reserve(nT * sizeof(double)) != HQ_OK) return g_err;
    HQ_HIP(hipMemcpyAsync(c->A.p, A, nA * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HQ_HIP(hipMemcpyAsync(c->TAU.p, TAU, nT * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (tile_qr_dev_impl(MT, NT, nb, ib, (double *)c->A.p, (double *)c->TAU.p, nullptr, c->stream,
                         false) != HQ_OK)
      return g_err;
    HQ_HIP(hipMemcpyAsync(A, c->A.p, nA * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HQ_HIP(hipMemcpyAsync(TAU, c->TAU.p, nT * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HQ_HIP(hipStreamSynchronize(c->stream));
    if (take_device_faults((int *)c->counter.p) != HQ_OK) return g_err;
    return ok();
  };
  run();
  report("Hybrid_tile_QR");
}

}  // extern "C"

// Shared body of the host-pointer Q entry points: copies RY, TAU and X to
// the device, runs `fn`, copies X (and RY when `ry_out`) back.
template <typename Fn>
static int host_q_call(mat_stat_t *status, double *X, double *RY, double *TAU, int my_rank,
                       int Proc_num, bool x_in, bool ry_out, Fn &&fn) {
  if (check_rank(my_rank, Proc_num) != HQ_OK) return g_err;
  if (!status || !X || !RY || !TAU || status->tile_M != status->tile_N)
    return set_error(HQ_ERR_ARG, "need square tiles and non-null arrays");
  const int MT = (int)status->glob_tile_M, NT = (int)status->glob_tile_N;
  const int nb = (int)status->tile_M;
  const int ib = std::min(nb, HQ_IBMAX);
  if (!geometry_ok(MT, NT, nb, ib)) return g_err;
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  const size_t nA = (size_t)MT * NT * nb * nb, nT = (size_t)MT * NT * nb;
  std::lock_guard<std::mutex> lk(c->host_mu);
  if (c->A.reserve(nA * sizeof(double)) != HQ_OK) return g_err;
  if (c->TAU.reserve(nT * sizeof(double)) != HQ_OK) return g_err;
  if (c->X.reserve(nA * sizeof(double)) != HQ_OK) return g_err;
  double *dRY = (double *)c->A.p, *dTAU = (double *)c->TAU.p, *dX = (double *)c->X.p;
  HQ_HIP(hipMemcpyAsync(dRY, RY, nA * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HQ_HIP(hipMemcpyAsync(dTAU, TAU, nT * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (x_in) HQ_HIP(hipMemcpyAsync(dX, X, nA * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (fn(MT, NT, nb, ib, dRY, dTAU, dX, c->stream) != HQ_OK) return g_err;
  HQ_HIP(hipMemcpyAsync(X, dX, nA * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (ry_out) HQ_HIP(hipMemcpyAsync(RY, dRY, nA * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HQ_HIP(hipStreamSynchronize(c->stream));
  if (take_device_faults((int *)c->counter.p) != HQ_OK) return g_err;
  return ok();
}

extern "C" {

void Hybrid_dorgqr(mat_stat_t *status, double *Q, double *RY, double *TAU, int my_rank,
                   int Proc_num) {
  host_q_call(status, Q, RY, TAU, my_rank, Proc_num, false, false,
              [](int MT, int NT, int nb, int ib, double *dRY, double *dTAU, double *dX,
                 hipStream_t st) {
                return hq_tile_orgqr_dev(MT, NT, nb, ib, dRY, dTAU, nullptr, dX, st);
              });
  report("Hybrid_dorgqr");
}

void Hybrid_dormqr(mat_stat_t *status, double *RY, double *TAU, double *C, int my_rank,
                   int Proc_num) {
  host_q_call(status, C, RY, TAU, my_rank, Proc_num, true, false,
              [](int MT, int NT, int nb, int ib, double *dRY, double *dTAU, double *dX,
                 hipStream_t st) {
                return hq_tile_ormqr_dev(MT, NT, nb, ib, dRY, dTAU, nullptr, dX, NT, st);
              });
  report("Hybrid_dormqr");
}

void Hybrid_make_Q(mat_stat_t *status, double *Q, double *RY, double *TAU, int my_rank,
                   int Proc_num, int ib) {
  (void)ib;  // the reference applies full-width T here (ib unused, Hybrid_make_Q.c:124-178)
  host_q_call(status, Q, RY, TAU, my_rank, Proc_num, true, true,
              [](int MT, int NT, int nb, int ib2, double *dRY, double *dTAU, double *dX,
                 hipStream_t st) {
                return hq_tile_make_q_dev(MT, NT, nb, ib2, dRY, dTAU, dX, st);
              });
  report("Hybrid_make_Q");
}

}  // extern "C"

// =====================================================================
// Homegrown tile_kernel.h API on host pointers (tile_kernel.h:28-162)
// =====================================================================
namespace {

// One host array of a homegrown call: `rows` x `cols` with leading dimension
// `ld` (only (cols-1)*ld + rows entries are touched), copied in and/or out.
struct HArr {
  double *h;
  int rows, cols, ld;
  bool in, out;
  double *d = nullptr;
  size_t n() const { return cols <= 0 || rows <= 0 ? 0 : (size_t)(cols - 1) * ld + rows; }
};

// Lane pool of a context: take an idle lane (or make one), give it back.
Lane *take_lane(Context *c) {
  {
    std::lock_guard<std::mutex> lk(c->lane_mu);
    if (!c->lanes.empty()) {
      Lane *l = c->lanes.back().release();
      c->lanes.pop_back();
      return l;
    }
  }
  std::unique_ptr<Lane> l(new Lane());
  if (hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    set_error(HQ_ERR_DEVICE, "hipStreamCreateWithFlags failed");
    return nullptr;
  }
  return l.release();
}
void give_lane(Context *c, Lane *l) {
  std::lock_guard<std::mutex> lk(c->lane_mu);
  c->lanes.emplace_back(l);
}

template <typename Fn>
int homegrown_call(const char *fn_name, std::vector<HArr> arrs, size_t work_elems, Fn &&fn) {
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  Lane *lane = take_lane(c);
  if (!lane) return g_err;
  struct Back {  // the lane returns to the pool on every exit path
    Context *c;
    Lane *l;
    ~Back() { give_lane(c, l); }
  } back{c, lane};
  size_t total = work_elems;
  for (auto &a : arrs) total += (a.n() + 31) & ~(size_t)31;
  if (lane->buf.reserve(total * sizeof(double)) != HQ_OK) return g_err;
  double *base = (double *)lane->buf.p;
  hipStream_t st = lane->stream;
  size_t off = 0;
  for (auto &a : arrs) {
    a.d = base + off;
    off += (a.n() + 31) & ~(size_t)31;
    if (a.in && a.n())
      HQ_HIP(hipMemcpyAsync(a.d, a.h, a.n() * sizeof(double), hipMemcpyHostToDevice, st));
  }
  double *work = base + off;
  if constexpr (std::is_invocable_v<Fn, std::vector<HArr> &, double *, hipStream_t, Lane *>)
    HQ_HIP(fn(arrs, work, st, lane));
  else
    HQ_HIP(fn(arrs, work, st));
  for (auto &a : arrs)
    if (a.out && a.n())
      HQ_HIP(hipMemcpyAsync(a.h, a.d, a.n() * sizeof(double), hipMemcpyDeviceToHost, st));
  HQ_HIP(hipStreamSynchronize(st));
  (void)fn_name;
  return ok();
}

// The lane's rocBLAS handle on the lane's stream (nullptr if it cannot be made).
rocblas_handle lane_blas(Lane *l) {
  if (!l->blas && rocblas_create_handle(&l->blas) != rocblas_status_success) l->blas = nullptr;
  if (l->blas && rocblas_set_stream(l->blas, l->stream) != rocblas_status_success) return nullptr;
  return l->blas;
}

bool hg_dims_ok(int row, int col, int ld) {
  if (row < 1 || col < 1 || col > 4096 || ld < row)
    return set_error(HQ_ERR_ARG, "tile kernel: need 1 <= col <= 4096, row >= 1, ld >= row"), false;
  return true;
}

size_t hg_work(int row, int col) {
  size_t m = (size_t)(row > col ? row : col);
  return 3 * m * m + 64;
}

}  // namespace

extern "C" {

void *tile_init(char type, int tile_row, int tile_col) {
  (void)type;  // the reference allocates doubles whatever the type (tile_kernel.c:19-25)
  return calloc((size_t)tile_row * tile_col * 3, sizeof(double));
}

void tile_dgeqrt(const int row, const int col, const int block, double *A, const int lda,
                 double *TAU, double *Y, double *T, const int ldt, double *work) {
  (void)work;
  if (!hg_dims_ok(row, col, lda) || block < 1 || ldt < col) {
    if (g_err == HQ_OK) set_error(HQ_ERR_ARG, "tile_dgeqrt: bad block / ldt");
    report("tile_dgeqrt");
    return;
  }
  homegrown_call("tile_dgeqrt",
                 {{A, row, col, lda, true, true}, {TAU, col, 1, col, true, true},
                  {Y, row, col, lda, true, true}, {T, col, col, ldt, true, true}},
                 hg_work(row, col), [&](std::vector<HArr> &a, double *w, hipStream_t st) {
                   return hqh_launch_dgeqrt(row, col, block, a[0].d, lda, a[1].d, a[2].d, a[3].d,
                                            ldt, w, st);
                 });
  report("tile_dgeqrt");
}

void tile_dtsqrt(const int row, const int col, const int block, double *A1, const int lda1,
                 double *A2, const int lda2, double *TAU, double *T, const int ldt, double *work) {
  (void)work;
  if (!hg_dims_ok(row, col, lda2) || lda1 < col || block < 1 || ldt < col) {
    if (g_err == HQ_OK) set_error(HQ_ERR_ARG, "tile_dtsqrt: bad block / lda1 / ldt");
    report("tile_dtsqrt");
    return;
  }
  homegrown_call("tile_dtsqrt",
                 {{A1, col, col, lda1, true, true}, {A2, row, col, lda2, true, true},
                  {TAU, col, 1, col, true, true}, {T, col, col, ldt, true, true}},
                 hg_work(row, col), [&](std::vector<HArr> &a, double *w, hipStream_t st) {
                   return hqh_launch_dtsqrt(row, col, block, a[0].d, lda1, a[1].d, lda2, a[2].d,
                                            a[3].d, ldt, w, st);
                 });
  report("tile_dtsqrt");
}

}  // extern "C"

// tile_dlarfb on square tiles and tile_dssrfb: the reference's BLAS call
// sequence (tile_kernel.c:273-298 with row = col, and tile_sub_dssrfb,
// tile_kernel.c:313-363, with over_row = under_row = row, Y_col = A_col =
// col) as rocBLAS products on the lane's stream, so any T the caller passes
// is applied as written (the products are the formula, not a compact-WY
// shortcut), and rows of A1 past col come back zero as in the reference.
hipError_t hg_dlarfb_blas(Lane *ln, int n, double *A, int lda, const double *Y, const double *T, int ldt,
                          double *w) {
  rocblas_handle h = lane_blas(ln);
  if (!h) return hipErrorUnknown;
  const double one = 1.0, mone = -1.0;
  double *D = w, *D2 = w + (size_t)n * n;
  // D = tril(Y)^T A; D2 = tril(T) D; A -= Y D2
  if (rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                    rocblas_diagonal_non_unit, n, n, &one, Y, lda, A, lda, D, n) != rocblas_status_success ||
      rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit,
                    n, n, &one, T, ldt, D, n, D2, n) != rocblas_status_success ||
      rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &mone, Y, lda, D2, n, &one, A,
                    lda) != rocblas_status_success)
    return hipErrorUnknown;
  return hipSuccess;
}

hipError_t hg_dssrfb_blas(Lane *ln, int row, int col, double *A1, double *A2, int lda, const double *Y,
                          const double *T, int ldt, double *w, hipStream_t st) {
  rocblas_handle h = lane_blas(ln);
  if (!h) return hipErrorUnknown;
  const double one = 1.0, mone = -1.0;
  double *D = w, *D2 = w + (size_t)row * col;
  // D = A1 (row rows); D[:col] += Y^T A2; D2 = tril(T) D[:col]
  hipError_t e = hipMemcpy2DAsync(D, sizeof(double) * row, A1, sizeof(double) * lda, sizeof(double) * row, col,
                                  hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return e;
  if (rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, col, col, row, &one, Y, lda, A2, lda,
                    &one, D, row) != rocblas_status_success ||
      rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit,
                    col, col, &one, T, ldt, D, row, D2, col) != rocblas_status_success)
    return hipErrorUnknown;
  // A1 -= D: rows < col take D2, rows col..row-1 (D = A1 there) become zero
  if (rocblas_dgeam(h, rocblas_operation_none, rocblas_operation_none, col, col, &one, A1, lda, &mone, D2, col, A1,
                    lda) != rocblas_status_success)
    return hipErrorUnknown;
  if (row > col && (e = hipMemset2DAsync(A1 + col, sizeof(double) * lda, 0, sizeof(double) * (row - col), col,
                                         st)) != hipSuccess)
    return e;
  // A2 -= Y D2
  if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, row, col, col, &mone, Y, lda, D2, col, &one,
                    A2, lda) != rocblas_status_success)
    return hipErrorUnknown;
  return hipSuccess;
}

extern "C" {

void tile_dlarfb(const int row, const int col, double *A, const int lda, const double *Y,
                 const double *T, const int ldt, double *work) {
  (void)work;
  if (!hg_dims_ok(row, col, lda) || ldt < col) {
    if (g_err == HQ_OK) set_error(HQ_ERR_ARG, "tile_dlarfb: bad ldt");
    report("tile_dlarfb");
    return;
  }
  if (row == col) {
    homegrown_call("tile_dlarfb",
                   {{A, row, col, lda, true, true}, {(double *)Y, row, col, lda, true, false},
                    {(double *)T, col, col, ldt, true, false}},
                   2 * (size_t)col * col, [&](std::vector<HArr> &a, double *w, hipStream_t, Lane *ln) {
                     return hg_dlarfb_blas(ln, col, a[0].d, lda, a[1].d, a[2].d, ldt, w);
                   });
    report("tile_dlarfb");
    return;
  }
  homegrown_call("tile_dlarfb",
                 {{A, row, col, lda, true, true}, {(double *)Y, row, col, lda, true, false},
                  {(double *)T, col, col, ldt, true, false}},
                 hg_work(row, col), [&](std::vector<HArr> &a, double *w, hipStream_t st) {
                   return hqh_launch_dlarfb(row, col, a[0].d, lda, a[1].d, a[2].d, ldt, w, st);
                 });
  report("tile_dlarfb");
}

void tile_dssrfb(const int row, const int col, double *A1, double *A2, const int lda,
                 const double *Y, const double *T, const int ldt, double *work) {
  (void)work;
  if (!hg_dims_ok(row, col, lda) || ldt < col) {
    if (g_err == HQ_OK) set_error(HQ_ERR_ARG, "tile_dssrfb: bad ldt");
    report("tile_dssrfb");
    return;
  }
  homegrown_call("tile_dssrfb",
                 {{A1, row, col, lda, true, true}, {A2, row, col, lda, true, true},
                  {(double *)Y, row, col, lda, true, false}, {(double *)T, col, col, ldt, true, false}},
                 (size_t)row * col + (size_t)col * col, [&](std::vector<HArr> &a, double *w, hipStream_t st, Lane *ln) {
                   return hg_dssrfb_blas(ln, row, col, a[0].d, a[1].d, lda, a[2].d, a[3].d, ldt, w, st);
                 });
  report("tile_dssrfb");
}

void tile_dlarft(const int row, const int col, const int block, const double *Y, const int lda,
                 const double *TAU, double *T, const int ldt, double *work) {
  (void)work;
  if (!hg_dims_ok(row, col, lda) || ldt < col) {
    if (g_err == HQ_OK) set_error(HQ_ERR_ARG, "tile_dlarft: bad ldt");
    report("tile_dlarft");
    return;
  }
  homegrown_call("tile_dlarft",
                 {{(double *)Y, row, col, lda, true, false},
                  {(double *)TAU, col, 1, col, true, false}, {T, col, col, ldt, true, true}},
                 hg_work(row, col), [&](std::vector<HArr> &a, double *w, hipStream_t st) {
                   return hqh_launch_dlarft(row, col, block, a[0].d, lda, a[1].d, a[2].d, ldt, w,
                                            st);
                 });
  report("tile_dlarft");
}

// Hybrid_make_I.c:20-53 (one rank): Id <- identity in tile layout.
void Hybrid_make_I(mat_stat_t *status, double *Id, int my_rank, int Proc_num) {
  if (check_rank(my_rank, Proc_num) != HQ_OK || !status || !Id) {
    if (g_err == HQ_OK) set_error(HQ_ERR_ARG, "Hybrid_make_I: null argument");
    report("Hybrid_make_I");
    return;
  }
  const size_t MT = status->my_tile_M, NT = status->my_tile_N, nb = status->tile_M;
  memset(Id, 0, sizeof(double) * MT * NT * nb * status->tile_N);
  for (size_t i = 0; i < NT && i < MT; i++)
    for (size_t r = 0; r < nb; r++) Id[(i + i * MT) * nb * nb + r + r * nb] = 1.0;
  ok();
}

// Hybrid_dgemm.c:16-70 (one rank, square tile grids): C <- A^T B, on the GPU.
void Hybrid_dgemm(mat_stat_t *status, double *A, double *B, double *Cm, int my_rank,
                  int Proc_num) {
  if (check_rank(my_rank, Proc_num) != HQ_OK || !status || !A || !B || !Cm ||
      status->tile_M != status->tile_N) {
    if (g_err == HQ_OK) set_error(HQ_ERR_ARG, "Hybrid_dgemm: need square tiles, non-null arrays");
    report("Hybrid_dgemm");
    return;
  }
  const int MT = (int)status->glob_tile_M, NT = (int)status->glob_tile_N, nb = (int)status->tile_M;
  const int n = (int)((size_t)MT * NT * nb * nb);
  homegrown_call("Hybrid_dgemm",
                 {{A, n, 1, n, true, false}, {B, n, 1, n, true, false}, {Cm, n, 1, n, false, true}},
                 0, [&](std::vector<HArr> &a, double *, hipStream_t st) {
                   return hqh_launch_tile_gemm_tn(a[0].d, a[1].d, a[2].d, MT, NT, nb, st);
                 });
  report("Hybrid_dgemm");
}

}  // extern "C"

// =====================================================================
// PLASMA core_blas shim: CORE_dgeqrt / CORE_dtsqrt / CORE_dormqr / CORE_dtsmqr
// with PLASMA 2.x argument lists, as the reference calls them
// (Hybrid_tile_QR.c:985-990,1135-1141,1276-1283,1356-1363), on the GPU.
// Host pointers with any leading dimension; every array is packed to the
// kernels' layout (ld = nb, T ld = ib) on a lane's stream and unpacked after.
// Square nb x nb tiles (nb <= 512), K = nb, side = Left,
// trans = Trans (the only forms the reference uses).  Return values follow
// PLASMA: 0 on success, -i when argument i is illegal or outside what this
// shim supports (message in hq_last_error_message()).
// =====================================================================
namespace {

struct PArr {
  double *h;
  int rows, cols, ld;
  bool in, out;
  double *d = nullptr;
};

template <typename Fn>
int packed_call(std::vector<PArr> arrs, Fn &&fn) {
  Context *c;
  if (get_ctx(&c) != HQ_OK) return g_err;
  Lane *lane = take_lane(c);
  if (!lane) return g_err;
  struct Back {
    Context *c;
    Lane *l;
    ~Back() { give_lane(c, l); }
  } back{c, lane};
  size_t total = 0;
  for (auto &a : arrs) total += ((size_t)a.rows * a.cols + 31) & ~(size_t)31;
  if (lane->buf.reserve(total * sizeof(double)) != HQ_OK) return g_err;
  hipStream_t st = lane->stream;
  size_t off = 0;
  for (auto &a : arrs) {
    a.d = (double *)lane->buf.p + off;
    off += ((size_t)a.rows * a.cols + 31) & ~(size_t)31;
    if (a.in)
      HQ_HIP(hipMemcpy2DAsync(a.d, sizeof(double) * a.rows, a.h, sizeof(double) * a.ld, sizeof(double) * a.rows,
                              a.cols, hipMemcpyHostToDevice, st));
  }
  if (fn(arrs, st) != HQ_OK) return g_err;
  for (auto &a : arrs)
    if (a.out)
      HQ_HIP(hipMemcpy2DAsync(a.h, sizeof(double) * a.ld, a.d, sizeof(double) * a.rows, sizeof(double) * a.rows,
                              a.cols, hipMemcpyDeviceToHost, st));
  HQ_HIP(hipStreamSynchronize(st));
  if (take_device_faults((int *)c->counter.p) != HQ_OK) return g_err;
  return ok();
}

// -i for the first failing check, with the message kept for hq_last_error
int core_bad(int i, const char *fn, const char *what) {
  set_error(HQ_ERR_ARG, "%s: argument %d: %s", fn, i, what);
  return -i;
}

}  // namespace

extern "C" {

int CORE_dgeqrt(int M, int N, int IB, double *A, int LDA, double *T, int LDT, double *TAU, double *WORK) {
  (void)WORK;
  const char *fn = "CORE_dgeqrt";
  if (M < 0) return core_bad(1, fn, "M < 0");
  if (N < 0) return core_bad(2, fn, "N < 0");
  if (IB < 0 || (IB == 0 && M > 0 && N > 0)) return core_bad(3, fn, "illegal IB");
  if (LDA < (M > 1 ? M : 1)) return core_bad(5, fn, "LDA < max(1, M)");
  if (LDT < (IB > 1 ? IB : 1)) return core_bad(7, fn, "LDT < max(1, IB)");
  if (M == 0 || N == 0 || IB == 0) return (ok(), 0);
  if (M != N || !nb_ok(M)) return core_bad(1, fn, "square tiles of at most 512 only");
  if (IB > N) return core_bad(3, fn, "IB > N");
  const int nb = M;
  const int rc = packed_call({{A, nb, nb, LDA, true, true}, {T, IB, nb, LDT, false, true}, {TAU, nb, 1, nb, false, true}},
                             [&](std::vector<PArr> &a, hipStream_t st) {
                               return hq_dgeqrt_dev(nb, IB, a[0].d, a[1].d, a[2].d, st);
                             });
  return rc == HQ_OK ? 0 : -1000 - rc;
}

int CORE_dtsqrt(int M, int N, int IB, double *A1, int LDA1, double *A2, int LDA2, double *T, int LDT, double *TAU,
                double *WORK) {
  (void)WORK;
  const char *fn = "CORE_dtsqrt";
  if (M < 0) return core_bad(1, fn, "M < 0");
  if (N < 0) return core_bad(2, fn, "N < 0");
  if (IB < 0) return core_bad(3, fn, "IB < 0");
  if (LDA1 < (N > 1 ? N : 1)) return core_bad(5, fn, "LDA1 < max(1, N)");
  if (LDA2 < (M > 1 ? M : 1)) return core_bad(7, fn, "LDA2 < max(1, M)");
  if (LDT < (IB > 1 ? IB : 1)) return core_bad(9, fn, "LDT < max(1, IB)");
  if (M == 0 || N == 0 || IB == 0) return (ok(), 0);
  if (M != N || !nb_ok(M)) return core_bad(1, fn, "square tiles of at most 512 only");
  if (IB > N) return core_bad(3, fn, "IB > N");
  const int nb = M;
  const int rc = packed_call({{A1, nb, nb, LDA1, true, true}, {A2, nb, nb, LDA2, true, true},
                              {T, IB, nb, LDT, false, true}, {TAU, nb, 1, nb, false, true}},
                             [&](std::vector<PArr> &a, hipStream_t st) {
                               return hq_dtsqrt_dev(nb, IB, a[0].d, a[1].d, a[2].d, a[3].d, st);
                             });
  return rc == HQ_OK ? 0 : -1000 - rc;
}

int CORE_dormqr(int side, int trans, int M, int N, int K, int IB, const double *A, int LDA, const double *T, int LDT,
                double *C, int LDC, double *WORK, int LDWORK) {
  (void)WORK;
  (void)LDWORK;
  const char *fn = "CORE_dormqr";
  if (side != HQ_PLASMA_LEFT) return core_bad(1, fn, "side must be PlasmaLeft");
  if (trans != HQ_PLASMA_TRANS) return core_bad(2, fn, "trans must be PlasmaTrans");
  if (M < 0) return core_bad(3, fn, "M < 0");
  if (N < 0) return core_bad(4, fn, "N < 0");
  if (K < 0 || K > M) return core_bad(5, fn, "K outside 0..M");
  if (IB < 0) return core_bad(6, fn, "IB < 0");
  if (LDA < (M > 1 ? M : 1)) return core_bad(8, fn, "LDA < max(1, M)");
  if (LDT < (IB > 1 ? IB : 1)) return core_bad(10, fn, "LDT < max(1, IB)");
  if (LDC < (M > 1 ? M : 1)) return core_bad(12, fn, "LDC < max(1, M)");
  if (M == 0 || N == 0 || K == 0) return (ok(), 0);
  if (M != N || K != M || !nb_ok(M)) return core_bad(3, fn, "M = N = K <= 512 only");
  if (IB == 0 || IB > K) return core_bad(6, fn, "IB outside 1..K");
  if (M < 128 && IB > HQ_IBMAX) return core_bad(6, fn, "nb = 64 takes IB <= 32");
  const int nb = M;
  const int rc = packed_call({{(double *)A, nb, nb, LDA, true, false}, {(double *)T, IB, nb, LDT, true, false},
                              {C, nb, nb, LDC, true, true}},
                             [&](std::vector<PArr> &a, hipStream_t st) {
                               return hq_dormqr_dev(nb, IB, a[0].d, a[1].d, a[2].d, st);
                             });
  return rc == HQ_OK ? 0 : -1000 - rc;
}

int CORE_dtsmqr(int side, int trans, int M1, int N1, int M2, int N2, int K, int IB, double *A1, int LDA1,
                double *A2, int LDA2, const double *V, int LDV, const double *T, int LDT, double *WORK,
                int LDWORK) {
  (void)WORK;
  (void)LDWORK;
  const char *fn = "CORE_dtsmqr";
  if (side != HQ_PLASMA_LEFT) return core_bad(1, fn, "side must be PlasmaLeft");
  if (trans != HQ_PLASMA_TRANS) return core_bad(2, fn, "trans must be PlasmaTrans");
  if (M1 < 0) return core_bad(3, fn, "M1 < 0");
  if (N1 < 0) return core_bad(4, fn, "N1 < 0");
  if (M2 < 0) return core_bad(5, fn, "M2 < 0");
  if (N2 < 0 || N2 != N1) return core_bad(6, fn, "N2 != N1");
  if (K < 0) return core_bad(7, fn, "K < 0");
  if (IB < 0) return core_bad(8, fn, "IB < 0");
  if (LDA1 < (M1 > 1 ? M1 : 1)) return core_bad(10, fn, "LDA1 < max(1, M1)");
  if (LDA2 < (M2 > 1 ? M2 : 1)) return core_bad(12, fn, "LDA2 < max(1, M2)");
  if (LDV < (M2 > 1 ? M2 : 1)) return core_bad(14, fn, "LDV < max(1, M2)");
  if (LDT < (IB > 1 ? IB : 1)) return core_bad(16, fn, "LDT < max(1, IB)");
  if (M1 == 0 || N1 == 0 || M2 == 0 || K == 0) return (ok(), 0);
  if (M1 != N1 || M2 != M1 || K != M1 || !nb_ok(M1)) return core_bad(3, fn, "M1 = N1 = M2 = N2 = K <= 512 only");
  if (IB == 0 || IB > K) return core_bad(8, fn, "IB outside 1..K");
  if (M1 < 128 && IB > HQ_IBMAX) return core_bad(8, fn, "nb = 64 takes IB <= 32");
  const int nb = M1;
  const int rc = packed_call({{A1, nb, nb, LDA1, true, true}, {A2, nb, nb, LDA2, true, true},
                              {(double *)V, nb, nb, LDV, true, false}, {(double *)T, IB, nb, LDT, true, false}},
                             [&](std::vector<PArr> &a, hipStream_t st) {
                               return hq_dtsmqr_dev(nb, IB, a[0].d, a[1].d, a[2].d, a[3].d, st);
                             });
  return rc == HQ_OK ? 0 : -1000 - rc;
}

}  // extern "C"
