// Batched multigrid-preconditioned CG with per-row convergence flags
// (torch_fem.pcg, torch_fem.py:201-343, with torch_mg.VCycle as M).
//
// Every batch row carries a device-side `active` flag.  A row that reaches rtol is
// frozen at the iteration it converged (its x, r, p are never touched again), which
// is the reference's row retirement without compacting the batch: workgroups of a
// retired row exit on their first instruction.  The host reads one integer (the
// active count) every `check_every` iterations instead of syncing on every
// iteration.  All inner products and norms are reduced in fp64, in a fixed order
// (per-tile partials, then one workgroup per row), so results are deterministic.
#include <utility>
#include <vector>

#include <algorithm>
#include <memory>
#include <vector>

#include <hip/hip_ext.h>

#include "stt_kernels.h"

namespace stt {

// ------------------------------------------------------------ scratch map ---

struct Scratch {
  double* part[5];  // [rows][tiles] each: pAp, rr, rz, rzo, bb
  double* bsafe;    // [rows]
  double* rz;       // [rows]
  double* alpha;    // [2][rows]: alpha of iteration k in slot k & 1
  double* beta;     // [rows]
  int* active;      // [rows]
  int* counter;     // [1]
  unsigned* bottom;  // the persistent bottom pass's counters (bottom_sync_words(rows))
  int* bottom_err;   // [1] set when a wait of that pass gave up
};

static long long align256(long long b) { return (b + 255) & ~255LL; }

static long long scratch_bytes(int rows, int tiles) {
  long long b = 0;
  b += 5 * align256(8LL * rows * tiles);
  b += 3 * align256(8LL * rows) + align256(16LL * rows);
  b += align256(4LL * rows);
  b += align256(16);
  b += align256(4LL * (bottom_sync_words(rows) + 1));
  return b;
}

// Partial-sum slots per row: the level-0 tiles, or the two-sweep passes' output tiles
// when there are more of them.
static int partial_slots(const OpDesc& op) {
  const int t = op.tiles(), p = pair_tiles(op);
  return t > p ? t : p;
}

static Scratch carve(void* base, int rows, int tiles) {
  Scratch s;
  char* p = static_cast<char*>(base);
  for (int i = 0; i < 5; ++i) {
    s.part[i] = reinterpret_cast<double*>(p);
    p += align256(8LL * rows * tiles);
  }
  s.bsafe = reinterpret_cast<double*>(p);
  p += align256(8LL * rows);
  s.rz = reinterpret_cast<double*>(p);
  p += align256(8LL * rows);
  s.alpha = reinterpret_cast<double*>(p);
  p += align256(16LL * rows);
  s.beta = reinterpret_cast<double*>(p);
  p += align256(8LL * rows);
  s.active = reinterpret_cast<int*>(p);
  p += align256(4LL * rows);
  s.counter = reinterpret_cast<int*>(p);
  p += align256(16);
  s.bottom = reinterpret_cast<unsigned*>(p);
  s.bottom_err = reinterpret_cast<int*>(s.bottom + bottom_sync_words(rows));
  return s;
}

// Zero the persistent bottom pass's counters and error word (before a solve or cycle).
static int clear_bottom(const Scratch& sc, int rows, hipStream_t s) {
  const hipError_t e =
      hipMemsetAsync(sc.bottom, 0, 4 * (size_t)(bottom_sync_words(rows) + 1), s);
  return e == hipSuccess ? STTOPT_OK : hip_status(e, "hipMemsetAsync(bottom counters)");
}

// Levels of the V-cycle's bottom that run as one persistent launch (vcycle_bottom): at
// most this many two-pass levels above the coarsest; 0 runs them launch by launch.
static int g_bottom_levels = 0;

// --------------------------------------------------------- tile kernels -----
// Elementwise PCG kernels walk the same 64 x 16 node tiles as the level-0 stencil
// (4 node rows per thread), so every fp64 partial-sum array has one slot per
// (row, tile).

struct TileGeom {
  int nnx, nny, ntx, tiles, total, d;
};

// Row, tile and first node of this thread; nodes j = 0..3 follow at +j*nnx.
__device__ __forceinline__ void tile_origin(const TileGeom& g, int& row, int& tt, int& c,
                                           int& r0) {
  const int t = xcd_remap(blockIdx.x, g.total);
  row_tile(t, g.tiles, g.total, row, tt);
  const int by = tt / g.ntx, bx = tt - by * g.ntx;
  c = bx * kTileX + (threadIdx.x & (kTileX - 1));
  r0 = by * kTileY + (threadIdx.x / kTileX) * kRowsPerThread;
}

// r -= alpha Ap; partial r.r (and r.z for the flexible beta).  x += alpha p is left to
// the direction update, which reads p anyway.  Each thread moves its nodes' D dofs as
// one vector.
template <int D, typename ZT>
__global__ __launch_bounds__(kBlock) void pcg_update_kernel(
    TileGeom g, long long n, const double* __restrict__ alpha, double* __restrict__ r,
    const double* __restrict__ Ap, const ZT* __restrict__ z, const int* __restrict__ active,
    double* __restrict__ prr, double* __restrict__ prz) {
  __shared__ double red[kBlock / kWave];
  int row, tt, c, r0;
  tile_origin(g, row, tt, c, r0);
  if (active[row] == 0) return;
  double rr = 0.0, rz = 0.0;
  const double a = alpha[row];
  double* rrow = r + row * n;
  const double* apr = Ap + row * n;
  const ZT* zr = z != nullptr ? z + row * n : nullptr;
  if (c < g.nnx) {
#pragma unroll
    for (int j = 0; j < kRowsPerThread; ++j) {
      if (r0 + j < g.nny) {
        const long long node = (long long)(r0 + j) * g.nnx + c;
        Vec<double, D> rv = load_node<double, D>(rrow, node);
        const Vec<double, D> av = load_node<double, D>(apr, node);
        Vec<ZT, D> zv;
#pragma unroll
        for (int k = 0; k < D; ++k) zv.v[k] = ZT(0);
        if (zr != nullptr) zv = load_node<ZT, D>(zr, node);
#pragma unroll
        for (int k = 0; k < D; ++k) {
          rv.v[k] = fma(-a, av.v[k], rv.v[k]);
          rr += rv.v[k] * rv.v[k];
          rz += rv.v[k] * (double)zv.v[k];
        }
        store_node<double, D>(rrow, node, rv);
      }
    }
  }
  double s = block_sum(rr, red);
  if (threadIdx.x == 0) prr[(long long)row * g.tiles + tt] = s;
  if (z != nullptr) {
    __syncthreads();
    s = block_sum(rz, red);
    if (threadIdx.x == 0) prz[(long long)row * g.tiles + tt] = s;
  }
}

// For the rows that took iteration k (iters == k; k = *kdev when kdev is given): x +=
// alpha p (after x += alpha_prev p_prev when p_prev is given: the update of iteration
// k - 1 that the fused direction update held back), then, for those still active and
// when update_p, p = z + beta p.
template <int D, typename ZT>
__global__ __launch_bounds__(kBlock) void pcg_pupdate_kernel(
    TileGeom g, long long n, const double* __restrict__ alpha, const double* __restrict__ beta,
    double* __restrict__ x, long long xs, double* __restrict__ p, const ZT* __restrict__ z,
    const int* __restrict__ active, const int* __restrict__ iters, int k, int update_p,
    const int* __restrict__ kdev, const double* __restrict__ alpha_prev,
    const double* __restrict__ p_prev) {
  int row, tt, c, r0;
  tile_origin(g, row, tt, c, r0);
  if (kdev != nullptr) k = *kdev;  // the iteration fin_alpha advanced it to
  if (iters[row] != k || c >= g.nnx) return;
  const bool dir = update_p && active[row] != 0;
  const double a = alpha[row];
  const double b = dir ? beta[row] : 0.0;
  const double ap = p_prev != nullptr ? alpha_prev[row] : 0.0;
  double* xr = x + row * xs;  // the caller's solution rows (stride xs)
  double* pr = p + row * n;
  const ZT* zr = z + row * n;
#pragma unroll
  for (int j = 0; j < kRowsPerThread; ++j) {
    if (r0 + j < g.nny) {
      const long long node = (long long)(r0 + j) * g.nnx + c;
      Vec<double, D> pv = load_node<double, D>(pr, node);
      Vec<double, D> xv = load_node<double, D>(xr, node);
      if (p_prev != nullptr) {
        const Vec<double, D> pq = load_node<double, D>(p_prev + row * n, node);
#pragma unroll
        for (int q = 0; q < D; ++q) xv.v[q] = fma(ap, pq.v[q], xv.v[q]);
      }
#pragma unroll
      for (int q = 0; q < D; ++q) xv.v[q] = fma(a, pv.v[q], xv.v[q]);
      store_node<double, D>(xr, node, xv);
      if (dir) {
        const Vec<ZT, D> zv = load_node<ZT, D>(zr, node);
#pragma unroll
        for (int q = 0; q < D; ++q) pv.v[q] = (double)zv.v[q] + b * pv.v[q];
        store_node<double, D>(pr, node, pv);
      }
    }
  }
}

// Cold start (x0 = 0): x = 0, r = P b and the partials of r.r over free dofs (= b.b
// there), summed as dot_kernel sums them.
template <int D, typename Unused>
__global__ __launch_bounds__(kBlock) void cold_init_kernel(TileGeom g, long long n,
                                                           const double* b, long long sb,
                                                           const uint32_t* mask, double* x,
                                                           long long sx, double* r,
                                                           double* part) {
  __shared__ double red[kBlock / kWave];
  int row, tt, c, r0;
  tile_origin(g, row, tt, c, r0);
  double acc = 0.0;
  if (c < g.nnx) {
#pragma unroll
    for (int j = 0; j < kRowsPerThread; ++j) {
      if (r0 + j < g.nny) {
        const long long node = (long long)(r0 + j) * g.nnx + c;
        const uint32_t word = mask != nullptr ? mask[(node * D) >> 5] : ~0u;
        const Vec<double, D> bv = load_node<double, D>(b + row * sb, node);
        Vec<double, D> rv, zero;
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const bool fr = (word >> ((node * D + k) & 31)) & 1u;
          rv.v[k] = fr ? bv.v[k] : 0.0;
          zero.v[k] = 0.0;
          if (fr) acc += bv.v[k] * bv.v[k];
        }
        store_node<double, D>(r + row * n, node, rv);
        store_node<double, D>(x + row * sx, node, zero);
      }
    }
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) part[(long long)row * g.tiles + tt] = s;
}

// partial sums of a.b over free dofs (b == nullptr: a.a), optionally only of
// active rows.
template <int D, typename BT>
__global__ __launch_bounds__(kBlock) void dot_kernel(TileGeom g, long long n, const double* a,
                                                     long long sa, const BT* b,
                                                     long long sb, const uint32_t* mask,
                                                     const int* active, double* part) {
  __shared__ double red[kBlock / kWave];
  int row, tt, c, r0;
  tile_origin(g, row, tt, c, r0);
  if (active != nullptr && active[row] == 0) return;
  double acc = 0.0;
  if (c < g.nnx) {
#pragma unroll
    for (int j = 0; j < kRowsPerThread; ++j) {
      if (r0 + j < g.nny) {
        const long long node = (long long)(r0 + j) * g.nnx + c;
        const uint32_t word = mask != nullptr ? mask[(node * D) >> 5] : ~0u;
        const Vec<double, D> av = load_node<double, D>(a + row * sa, node);
        Vec<double, D> bv = av;
        if (b != nullptr) {
          const Vec<BT, D> bt = load_node<BT, D>(b + row * sb, node);
#pragma unroll
          for (int k = 0; k < D; ++k) bv.v[k] = (double)bt.v[k];
        }
#pragma unroll
        for (int k = 0; k < D; ++k)
          if ((word >> ((node * D + k) & 31)) & 1u) acc += av.v[k] * bv.v[k];
      }
    }
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) part[(long long)row * g.tiles + tt] = s;
}

// Launch a TileGeom kernel for the geometry's d (1 or 2; validated by the callers).
#define STT_TILE_LAUNCH(kernel, TA, g, s, ...)                                           \
  do {                                                                                   \
    if ((g).d == 2)                                                                      \
      hipLaunchKernelGGL((kernel<2, TA>), dim3((g).total), dim3(kBlock), 0, s, g, __VA_ARGS__); \
    else                                                                                 \
      hipLaunchKernelGGL((kernel<1, TA>), dim3((g).total), dim3(kBlock), 0, s, g, __VA_ARGS__); \
  } while (0)

// ---------------------------------------------------- per-row finalizers ----

__device__ double row_sum(const double* part, int tiles, int row, double* red) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < tiles; i += blockDim.x) acc += part[(long long)row * tiles + i];
  return block_sum(acc, red);
}

__device__ __forceinline__ double safe_div(double num, double den) {
  return den != 0.0 ? num / den : 0.0;
}

// Initial residual: b norm, relative residual, active flags.
__global__ __launch_bounds__(kBlock) void fin_init_kernel(int tiles, const double* pbb,
                                                          const double* prr, double rtol,
                                                          double* bsafe, double* resid,
                                                          int* active, int* iters,
                                                          int* counter) {
  __shared__ double red[kBlock / kWave];
  const int row = blockIdx.x;
  const double bb = row_sum(pbb, tiles, row, red);
  __syncthreads();
  const double rr = row_sum(prr, tiles, row, red);
  if (threadIdx.x == 0) {
    const double bn = sqrt(bb);
    const double bs = bn > 0.0 ? bn : 1.0;
    const double rel = sqrt(rr) / bs;
    bsafe[row] = bs;
    resid[row] = rel;
    const int act = (rel <= rtol) ? 0 : 1;
    active[row] = act;
    iters[row] = 0;
    if (act) atomicAdd(counter, 1);
  }
}

// fin_rz and fin_beta also zero the active-row counter for the next fin_resid (the
// host's copy of the count, if any, is queued before them).
__global__ __launch_bounds__(kBlock) void fin_rz_kernel(int tiles, const double* prz,
                                                        const int* active, double* rz,
                                                        int* counter) {
  __shared__ double red[kBlock / kWave];
  const int row = blockIdx.x;
  if (row == 0 && threadIdx.x == 0) *counter = 0;
  if (active[row] == 0) return;
  const double s = row_sum(prz, tiles, row, red);
  if (threadIdx.x == 0) rz[row] = s;
}

// kdev: the device's copy of the PCG iteration number, advanced here once per
// iteration (read by the direction update, which may run from a captured graph).
__global__ __launch_bounds__(kBlock) void fin_alpha_kernel(int tiles, const double* ppap,
                                                           const int* active,
                                                           const double* rz, double* alpha,
                                                           int* kdev) {
  __shared__ double red[kBlock / kWave];
  const int row = blockIdx.x;
  if (row == 0 && threadIdx.x == 0) kdev[0] += 1;
  if (active[row] == 0) return;
  const double pap = row_sum(ppap, tiles, row, red);
  if (threadIdx.x == 0) alpha[row] = safe_div(rz[row], pap);
}

// host_flag (pinned host memory, or null): every row still active after this check
// stores check_id there (a vector store at system scope), so the host finds check_id
// in it exactly when a row is left; the check ids are never reused.
__global__ __launch_bounds__(kBlock) void fin_resid_kernel(int tiles, const double* prr,
                                                           double rtol, int k,
                                                           const double* bsafe,
                                                           double* resid, int* active,
                                                           int* iters, int* counter,
                                                           int* host_flag, int check_id) {
  __shared__ double red[kBlock / kWave];
  const int row = blockIdx.x;
  if (active[row] == 0) return;
  const double rr = row_sum(prr, tiles, row, red);
  if (threadIdx.x == 0) {
    const double rel = sqrt(rr) / bsafe[row];
    resid[row] = rel;
    iters[row] = k;
    if (rel <= rtol) {
      active[row] = 0;
    } else {
      atomicAdd(counter, 1);
      if (host_flag != nullptr)
        __hip_atomic_store(host_flag, check_id, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

__global__ __launch_bounds__(kBlock) void fin_beta_kernel(int tiles, const double* prz,
                                                          int tiles_o, const double* przo,
                                                          int flexible,
                                                          const int* active, double* rz,
                                                          double* beta, int* counter) {
  __shared__ double red[kBlock / kWave];
  const int row = blockIdx.x;
  if (row == 0 && threadIdx.x == 0) *counter = 0;
  if (active[row] == 0) return;
  const double rzn = row_sum(prz, tiles, row, red);
  double rzo = 0.0;
  if (flexible) {
    __syncthreads();
    rzo = row_sum(przo, tiles_o, row, red);
  }
  if (threadIdx.x == 0) {
    beta[row] = safe_div(rzn - rzo, rz[row]);
    rz[row] = rzn;
  }
}

// ------------------------------------------------------------- V-cycle ------

template <typename T>
struct VCycle {
  const sttopt_mg* mg;
  hipStream_t s;
  const int* active;
  // the persistent bottom pass's counters and error word (null: launch by launch), and
  // whether a cycle has used it
  unsigned* bsync = nullptr;
  int* berr = nullptr;
  bool bottom_used = false;
  // the coarsest inverse's completion event (queued on another stream), waited for once
  // before the first launch that reads the inverse; null once waited for
  hipEvent_t cinv_ready = nullptr;
  int rows;
  OpDesc op[STTOPT_MAX_LEVELS];
  long long n[STTOPT_MAX_LEVELS];

  VCycle(const sttopt_mg* m, hipStream_t st, const int* act) : mg(m), s(st), active(act) {
    rows = m->rows;
    cinv_ready = static_cast<hipEvent_t>(m->cinv_ready);
    for (int l = 0; l < m->nlevels; ++l) {
      const sttopt_level& L = m->levels[l];
      OpDesc& o = op[l];
      o.d = m->d;
      o.nelx = L.nelx;
      o.nely = L.nely;
      o.mask = L.mask;
      if (l == 0) {
        o.rho = L.density;
        o.sr = L.density_stride;
        o.rdiag = L.rdiag;
        o.srd = L.rdiag_stride;
      } else if (L.ptable != nullptr) {
        o.ptab = L.ptable;
        o.pkx = L.pkx;
        o.pky = L.pky;
        o.fnelx = m->levels[0].nelx;
        o.prho = m->levels[0].density;
        o.spr = m->levels[0].density_stride;
      } else {
        o.keff = L.keff;
        o.sk = L.keff_stride;
        o.keff_upper = L.keff_layout == STTOPT_KEFF_UPPER;
        o.keff_node = L.keff_layout == STTOPT_KEFF_NODE;
      }
      for (int i = 0; i < 64; ++i) o.ke[i] = m->ke[i];
      n[l] = o.ndof();
    }
  }

  // Level-0 smoothing in two-sweep passes: the fp32 cycle with two sweeps and one
  // coarse visit over a 2 x 2 coarsening of the density operator.
  bool pair() const {
    const sttopt_level& L = mg->levels[0];
    return sizeof(T) == sizeof(float) && mg->n_smooth == 2 && mg->gamma == 1 &&
           mg->nlevels > 1 && L.kx == 2 && L.ky == 2 && op[0].rho != nullptr &&
           L.mask != nullptr &&
           op[0].ndof() * (long long)sizeof(double) < (1LL << 32) &&
           (L.flags & STTOPT_SEPARATE_SWEEPS) == 0;
  }
  // reduction partials per row of the r.z sums written by run()
  int rz_tiles() const { return pair() ? pair_tiles(op[0]) : op[0].tiles(); }

  // Orders the inverse's launches before what s launches next (the first cycle runs on
  // the solve's own stream, outside any graph capture, so later cycles need no wait).
  int wait_cinv() {
    if (cinv_ready == nullptr) return STTOPT_OK;
    hipError_t e = hipStreamWaitEvent(s, cinv_ready, 0);
    if (e != hipSuccess) return hip_status(e, "hipStreamWaitEvent(coarsest inverse)");
    cinv_ready = nullptr;
    return STTOPT_OK;
  }

  const T* invd(int l) const { return static_cast<const T*>(mg->levels[l].invdiag); }
  long long sd(int l) const { return mg->levels[l].invdiag_stride; }
  T* buf(void* p) const { return static_cast<T*>(p); }

  // n_smooth damped-Jacobi sweeps x <- x + w D^-1 (b - A x) starting from x;
  // returns the buffer holding the result (x or alt).
  int sweeps(int l, int count, T*& x, T*& alt, const T* b, long long sb) {
    for (int k = 0; k < count; ++k) {
      if (int e = stencil<T, T, T>(op[l], rows, kSmooth, false, kNoDot, x, n[l], b, sb,
                                   invd(l), sd(l), alt, n[l], mg->omega, active, nullptr, s))
        return e;
      T* t = x;
      x = alt;
      alt = t;
    }
    return STTOPT_OK;
  }

  // The number of two-pass node-stencil levels from level l to the one above the
  // coarsest when the rest of the cycle from l can run as one persistent launch
  // (vcycle_bottom), else 0.
  int bottom_levels(int l) const {
    if constexpr (sizeof(T) != sizeof(float)) {
      return 0;
    } else {
      const int last = mg->nlevels - 1, nb = last - l;
      if (bsync == nullptr || berr == nullptr || nb < 1 || nb > g_bottom_levels ||
          nb > kMaxBottomLevels || mg->n_smooth != 2 || mg->gamma != 1 ||
          (mg->levels[0].flags & STTOPT_SEPARATE_SWEEPS) != 0)
        return 0;
      for (int i = l; i < last; ++i) {
        const sttopt_level& Li = mg->levels[i];
        if (!op[i].keff_node || Li.kx != 2 || Li.ky != 2 || Li.mask == nullptr ||
            (Li.flags & STTOPT_LEVEL_SEPARATE) != 0)
          return 0;
      }
      const sttopt_level& C = mg->levels[last];
      if (C.cinv == nullptr || n[last] % 2 != 0 || C.cinv_stride % 2 != 0 ||
          (reinterpret_cast<uintptr_t>(C.cinv) & 15) != 0)
        return 0;
      return nb;
    }
  }

  // Cycle on level l >= 1 with right-hand side b; *out receives the result buffer.
  int cycle(int l, const T* b, T** out) {
    const sttopt_level& L = mg->levels[l];
    const long long nl = n[l];
    if constexpr (sizeof(T) == sizeof(float)) {
      // the two passes of levels l .. last - 1 and the coarsest solve in one launch: the
      // buffers of the launch-by-launch path below (x2 in vx, the post result in vx2,
      // the restricted residuals in the coarser levels' vb, the coarsest result in vx)
      if (const int nb = bottom_levels(l)) {
        const int last = mg->nlevels - 1;
        BottomCall c{};
        c.nlev = nb;
        for (int i = 0; i < nb; ++i) {
          const int li = l + i;
          const sttopt_level& Li = mg->levels[li];
          const sttopt_level& Ci = mg->levels[li + 1];
          PairLevel& P = c.lv[i];
          P.op = op[li];
          P.rows = rows;
          P.b = i == 0 ? b : buf(Li.vb);
          P.sb = n[li];
          P.invd = invd(li);
          P.sd = sd(li);
          P.x2 = buf(Li.vx);
          P.sx2 = n[li];
          P.cmask = Ci.mask;
          P.bc = buf(Ci.vb);
          P.sbc = n[li + 1];
          P.ec = li + 1 < last ? buf(Ci.vx2) : buf(Ci.vx);
          P.sec = n[li + 1];
          P.x4 = buf(Li.vx2);
          P.sx4 = n[li];
          P.omega = mg->omega;
          P.active = active;
        }
        const sttopt_level& C = mg->levels[last];
        c.n = n[last];
        c.cinv = C.cinv;
        c.scinv = C.cinv_stride;
        c.cb = buf(C.vb);
        c.scb = n[last];
        c.cmask = C.mask;
        c.cx = buf(C.vx);
        c.scx = n[last];
        if (int e = wait_cinv()) return e;
        if (int e = vcycle_bottom(c, bsync, berr, s)) return e;
        bottom_used = true;
        *out = buf(L.vx2);
        return STTOPT_OK;
      }
    }
    if (l == mg->nlevels - 1) {
      T* x = buf(L.vx);
      if (int e = wait_cinv()) return e;
      if (int e = coarse_solve<T>(nl, rows, L.cinv, L.cinv_stride, b, nl, L.mask, x, nl,
                                  active, s))
        return e;
      *out = x;
      return STTOPT_OK;
    }
    T* x = buf(L.vx);
    T* alt = buf(L.vx2);
    const int ns = mg->n_smooth;
    if constexpr (sizeof(T) == sizeof(float)) {
      // node-stencil levels at 2 x 2 with two sweeps: pre-smoothing, residual and
      // restriction in one pass, prolongation and post-smoothing in another
      const sttopt_level& C = mg->levels[l + 1];
      if (op[l].keff_node && L.kx == 2 && L.ky == 2 && ns == 2 && mg->gamma == 1 &&
          L.mask != nullptr && (mg->levels[0].flags & STTOPT_SEPARATE_SWEEPS) == 0 &&
          (L.flags & STTOPT_LEVEL_SEPARATE) == 0) {
        const long long nc = n[l + 1];
        const int variant = (L.flags >> STTOPT_TILE_VARIANT_SHIFT) & 15;
        if (int e = coarse_pair_pre(op[l], rows, b, nl, invd(l), sd(l), x, nl, C.mask,
                                    buf(C.vb), nc, mg->omega, active, s, variant))
          return e;
        T* ec = nullptr;
        if (int e = cycle(l + 1, buf(C.vb), &ec)) return e;
        if (int e = coarse_pair_post(op[l], rows, x, nl, ec, nc, b, nl, invd(l), sd(l), alt,
                                     nl, mg->omega, active, s, variant))
          return e;
        *out = alt;
        return STTOPT_OK;
      }
    }
    // the first two sweeps in one pass (the first one, from zero, formed where the
    // second reads it): on the LDS tile of the fp32 level 1 at 2 x 2, and on the
    // one-node-per-thread kernel for small levels, where the launch it saves costs
    // more than gathering b and D^-1 instead of x at the 9 neighbours (on a 256 x 128
    // level the separate first sweep and pass take 5 + 17 us, the fused pass 28 us)
    constexpr long long kFuseFirstMaxNodes = 16384;
    const bool tile_level = sizeof(T) == sizeof(float) && op[l].ptab != nullptr &&
                            op[l].pkx == 2 && op[l].pky == 2;
    const bool fuse_first = ns >= 2 &&
                            (mg->levels[0].flags & STTOPT_SEPARATE_SWEEPS) == 0 &&
                            (tile_level || op[l].nnodes() <= kFuseFirstMaxNodes);
    if (fuse_first) {
      if (int e = stencil<T, T, T>(op[l], rows, kSmoothFirst, false, kNoDot,
                                   static_cast<const T*>(nullptr), 0, b, nl, invd(l), sd(l), x,
                                   nl, mg->omega, active, nullptr, s))
        return e;
      if (int e = sweeps(l, ns - 2, x, alt, b, nl)) return e;
    } else if (ns >= 1) {
      if (int e = jacobi_first<T, T>(nl, rows, b, nl, invd(l), sd(l), mg->omega, L.mask, x,
                                     nl, nullptr, 0, active, s))
        return e;
      if (int e = sweeps(l, ns - 1, x, alt, b, nl)) return e;
    } else if (int e = fill<T>(nl, rows, T(0), x, nl, active, s)) {
      return e;
    }
    if constexpr (sizeof(T) == sizeof(float)) {
      // the prolongation folded into the first post sweep (its staging adds P e) on the
      // LDS tile of the fp32 level 1 at 2 x 2 over a 2 x 2 coarsening
      if (tile_level && ns >= 1 && mg->gamma == 1 && L.kx == 2 && L.ky == 2 &&
          (mg->levels[0].flags & STTOPT_SEPARATE_SWEEPS) == 0) {
        const sttopt_level& C = mg->levels[l + 1];
        T* ec = nullptr;
        if (int e = coarse_visit(l, x, b, nl, &ec)) return e;
        if (int e = smooth_prolong(op[l], rows, x, nl, ec, n[l + 1], C.nelx + 1, C.nely + 1,
                                   b, nl, invd(l), sd(l), alt, nl, mg->omega, active, s))
          return e;
        T* t = x;
        x = alt;
        alt = t;
        if (int e = sweeps(l, ns - 1, x, alt, b, nl)) return e;
        *out = x;
        return STTOPT_OK;
      }
    }
    if (int e = coarse_correct(l, x, b, nl)) return e;
    if (int e = sweeps(l, ns, x, alt, b, nl)) return e;
    *out = x;
    return STTOPT_OK;
  }

  // x += P (cycle(R (b - A x))), with gamma coarse visits.
  int coarse_correct(int l, T* x, const T* b, long long sb) {
    const sttopt_level& L = mg->levels[l];
    const sttopt_level& C = mg->levels[l + 1];
    const long long nl = n[l], nc = n[l + 1];
    T* ec = nullptr;
    if (int e = coarse_visit(l, x, b, sb, &ec)) return e;
    return prolong_add<T>(mg->d, C.nelx, C.nely, L.kx, L.ky, rows, ec, nc, L.mask, x, nl,
                          active, s);
  }

  // *ec = cycle(R (b - A x)) on level l + 1, with gamma coarse visits (the correction
  // that coarse_correct prolongates).
  int coarse_visit(int l, const T* x, const T* b, long long sb, T** ec_out) {
    const sttopt_level& L = mg->levels[l];
    const sttopt_level& C = mg->levels[l + 1];
    const long long nl = n[l], nc = n[l + 1];
    T* r = buf(L.vr);
    if (int e = stencil<T, T, T>(op[l], rows, kResid, false, kNoDot, x, nl, b, sb, nullptr,
                                 0, r, nl, 0.0, active, nullptr, s))
      return e;
    const int cnelx = C.nelx, cnely = C.nely;
    if (int e = restrict_to<T>(mg->d, cnelx, cnely, L.kx, L.ky, rows, r, nl, C.mask,
                               buf(C.vb), nc, active, s))
      return e;
    T* ec = nullptr;
    if (int e = cycle(l + 1, buf(C.vb), &ec)) return e;
    if (mg->gamma > 1) {
      T* acc = buf(C.ve);
      if (int e = convert<T, T>(nc, rows, ec, nc, acc, nc, active, s)) return e;
      for (int g = 1; g < mg->gamma; ++g) {
        if (int e = stencil<T, T, T>(op[l + 1], rows, kResid, false, kNoDot, acc, nc,
                                     buf(C.vb), nc, nullptr, 0, buf(C.vw), nc, 0.0, active,
                                     nullptr, s))
          return e;
        T* e2 = nullptr;
        if (int e = cycle(l + 1, buf(C.vw), &e2)) return e;
        if (int e = axpy<T>(nc, rows, e2, nc, acc, nc, active, s)) return e;
      }
      ec = acc;
    }
    *ec_out = ec;
    return STTOPT_OK;
  }

  // The pair cycle split at the PCG's convergence check.  pair_pre_update: the residual
  // update rnew = r - alpha Ap (per-tile sums of its squares in rr, and of rnew.zold in
  // rzo when zold is given) in the pre-smoothing pass; pair_rest: the coarse visit and
  // the post pass on that residual, z and the r.z partials.  Only for pair() (the fp32
  // cycle).
  int pair_pre_update(const double* r, const double* Ap, const double* alpha, double* rnew,
                      double* rr, const T* zold, double* rzo) {
    if constexpr (sizeof(T) == sizeof(float)) {
      const sttopt_level& L = mg->levels[0];
      const long long nl = n[0];
      return smooth_pair_pre(op[0], rows, r, nl, invd(0), sd(0), buf(L.vx), nl, buf(L.vr),
                             nl, mg->omega, active, s, Ap, nl, alpha, rnew, nl, rr, zold,
                             nl, rzo);
    }
    set_error("vcycle: the fused residual update needs the fp32 pair cycle");
    return STTOPT_EINVAL;
  }
  template <typename ZT>
  int pair_rest(const double* r, long long rs, ZT* z, long long zs, double* part) {
    if constexpr (sizeof(T) == sizeof(float)) {
      const sttopt_level& L = mg->levels[0];
      const sttopt_level& C = mg->levels[1];
      const long long nl = n[0], nc = n[1];
      if (int e = restrict_to<T>(mg->d, C.nelx, C.nely, 2, 2, rows, buf(L.vr), nl, C.mask,
                                 buf(C.vb), nc, active, s))
        return e;
      T* ec = nullptr;
      if (int e = cycle(1, buf(C.vb), &ec)) return e;
      return smooth_pair_post<ZT>(op[0], rows, buf(L.vx), nl, ec, nc, C.nelx + 1, r, rs,
                                  invd(0), sd(0), z, zs, mg->omega, active, part, s);
    }
    set_error("vcycle: the pair cycle needs the fp32 V-cycle");
    return STTOPT_EINVAL;
  }

  // z = M(r) on level 0.  r fp64; z fp64 or in the V-cycle precision (ZT = T: the
  // cycle's last sweep computes in T, so storing z in T loses nothing).  When part !=
  // nullptr, writes the tile partials of r.z (free dofs) for the PCG.  tmp64: rows x
  // ndof doubles of scratch, needed when ZT is not double and there is one level.
  template <typename ZT>
  int run(const double* r, long long rs, ZT* z, long long zs, double* part,
          double* tmp64 = nullptr) {
    const sttopt_level& L = mg->levels[0];
    const long long nl = n[0];
    const int ns = mg->n_smooth;
    if (mg->nlevels == 1) {
      if (int e = wait_cinv()) return e;
      if constexpr (sizeof(ZT) == sizeof(double)) {
        if (int e = coarse_solve<double>(nl, rows, L.cinv, L.cinv_stride, r, rs, L.mask, z, zs,
                                         active, s))
          return e;
      } else {
        if (tmp64 == nullptr) {
          set_error("vcycle: a one-level cycle with a reduced-precision z needs scratch");
          return STTOPT_EINVAL;
        }
        if (int e = coarse_solve<double>(nl, rows, L.cinv, L.cinv_stride, r, rs, L.mask,
                                         tmp64, nl, active, s))
          return e;
        if (int e = convert<double, ZT>(nl, rows, tmp64, nl, z, zs, active, s)) return e;
      }
      return part ? dot_rz(r, rs, z, zs, part) : STTOPT_OK;
    }
    if constexpr (sizeof(T) == sizeof(float)) {
      if (pair()) {
        if (int e = smooth_pair_pre(op[0], rows, r, rs, invd(0), sd(0), buf(L.vx), nl,
                                    buf(L.vr), nl, mg->omega, active, s))
          return e;
        return pair_rest(r, rs, z, zs, part);
      }
    }
    // fine-level right-hand side in the V-cycle precision
    const T* b;
    long long sb;
    T* x = buf(L.vx);
    T* alt = buf(L.vx2);
    if constexpr (sizeof(T) == sizeof(double)) {
      b = reinterpret_cast<const T*>(r);
      sb = rs;
    } else {
      b = buf(L.vb);
      sb = nl;
    }
    if (ns >= 2) {
      // the first two sweeps from x = 0 in one pass (x0 = w D^-1 r staged in the tile);
      // the fp32 cycle gets its copy of r from the same pass
      T* bcopy = sizeof(T) == sizeof(double) ? nullptr : buf(L.vb);
      if (int e = stencil<T, double, T>(op[0], rows, kSmoothFirst, false, kNoDot,
                                        static_cast<const T*>(nullptr), 0, r, rs, invd(0),
                                        sd(0), x, nl, mg->omega, active, nullptr, s, bcopy, nl))
        return e;
      if (int e = sweeps(0, ns - 2, x, alt, b, sb)) return e;
    } else if (ns == 1) {
      if constexpr (sizeof(T) == sizeof(double)) {
        if (int e = jacobi_first<T, T>(nl, rows, b, sb, invd(0), sd(0), mg->omega, L.mask, x,
                                       nl, nullptr, 0, active, s))
          return e;
      } else if (int e = jacobi_first<T, double>(nl, rows, r, rs, invd(0), sd(0), mg->omega,
                                                 L.mask, x, nl, buf(L.vb), nl, active, s)) {
        return e;
      }
    } else {
      if constexpr (sizeof(T) != sizeof(double)) {
        if (int e = convert<double, T>(nl, rows, r, rs, buf(L.vb), nl, active, s)) return e;
      }
      if (int e = fill<T>(nl, rows, T(0), x, nl, active, s)) return e;
    }
    if (int e = coarse_correct(0, x, b, sb)) return e;
    if (ns >= 1) {
      if (int e = sweeps(0, ns - 1, x, alt, b, sb)) return e;
      // last sweep writes z and the r.z partials in the same pass
      return stencil<T, double, ZT>(op[0], rows, kSmooth, false, part ? kDotB : kNoDot, x,
                                    nl, r, rs, invd(0), sd(0), z, zs, mg->omega, active, part,
                                    s);
    }
    if (int e = convert<T, ZT>(nl, rows, x, nl, z, zs, active, s)) return e;
    return part ? dot_rz(r, rs, z, zs, part) : STTOPT_OK;
  }

  template <typename ZT>
  int dot_rz(const double* r, long long rs, const ZT* z, long long zs, double* part) {
    TileGeom g{op[0].nnx(), op[0].nny(), op[0].ntx(), op[0].tiles(),
               op[0].tiles() * rows, mg->d};
    STT_TILE_LAUNCH(dot_kernel, ZT, g, s, n[0], r, rs, z, zs,
                       op[0].mask, active, part);
    STT_CHECK_LAUNCH("dot_kernel");
    return STTOPT_OK;
  }
};

static int validate(const sttopt_mg* mg) {
  if (mg == nullptr || mg->rows < 0 || (mg->d != 1 && mg->d != 2) || mg->nlevels < 1 ||
      mg->nlevels > STTOPT_MAX_LEVELS || mg->n_smooth < 0 || mg->gamma < 1 ||
      (mg->vdtype != STTOPT_F64 && mg->vdtype != STTOPT_F32)) {
    set_error("sttopt_mg: invalid hierarchy header");
    return STTOPT_EINVAL;
  }
  for (int l = 0; l < mg->nlevels; ++l) {
    const sttopt_level& L = mg->levels[l];
    if (L.nelx <= 0 || L.nely <= 0) {
      set_error("sttopt_mg: level with an empty grid");
      return STTOPT_EINVAL;
    }
    const bool last = l == mg->nlevels - 1;
    if (last && L.cinv == nullptr) {
      set_error("sttopt_mg: coarsest level needs its dense inverse");
      return STTOPT_EINVAL;
    }
    if (!last) {
      const sttopt_level& C = mg->levels[l + 1];
      if (L.kx < 1 || L.ky < 1 || L.nelx != C.nelx * L.kx || L.nely != C.nely * L.ky) {
        set_error("sttopt_mg: level grids do not match their coarsening factors");
        return STTOPT_EINVAL;
      }
      if (L.invdiag == nullptr || L.vx == nullptr || L.vx2 == nullptr || L.vr == nullptr) {
        set_error("sttopt_mg: level is missing its diagonal or work vectors");
        return STTOPT_EINVAL;
      }
      if (l == 0 && L.density == nullptr) {
        set_error("sttopt_mg: level 0 needs the density in the V-cycle precision");
        return STTOPT_EINVAL;
      }
      if (mg->gamma > 1 && (C.ve == nullptr || C.vw == nullptr)) {
        set_error("sttopt_mg: gamma > 1 needs ve/vw work vectors");
        return STTOPT_EINVAL;
      }
    } else if (l > 0 && L.vx == nullptr) {
      set_error("sttopt_mg: coarsest level is missing its work vector");
      return STTOPT_EINVAL;
    }
    if (l > 0 && (L.vb == nullptr || (!last && L.keff == nullptr && L.ptable == nullptr))) {
      set_error("sttopt_mg: coarse level is missing keff or its right-hand-side vector");
      return STTOPT_EINVAL;
    }
    if (l > 0 && L.ptable != nullptr &&
        (mg->levels[0].density == nullptr || L.pkx < 1 || L.pky < 1 ||
         L.nelx * L.pkx != mg->levels[0].nelx || L.nely * L.pky != mg->levels[0].nely)) {
      set_error("sttopt_mg: parameterized level does not tile level 0");
      return STTOPT_EINVAL;
    }
    if (l == 0 && mg->vdtype == STTOPT_F32 && mg->nlevels > 1 && L.vb == nullptr) {
      set_error("sttopt_mg: an fp32 V-cycle needs the level-0 vb vector");
      return STTOPT_EINVAL;
    }
  }
  return STTOPT_OK;
}

// Per-host-thread readback resources: 16 words of pinned, device-mapped, coherent host
// memory (word 0 for synchronous readbacks, words 1-2 for the marks of the lagged
// convergence check) and the two events of that check (one per pinned slot).  Created
// on first use and released when the thread ends, so short-lived worker threads leave
// nothing behind.
struct ThreadReadback {
  int* pinned = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~ThreadReadback() {
    for (hipEvent_t& e : ev)
      if (e != nullptr) (void)hipEventDestroy(e);
    if (pinned != nullptr) (void)hipHostFree(pinned);
  }
};
static thread_local ThreadReadback t_readback;

static int* pinned_counter() {
  int*& p = t_readback.pinned;
  if (p == nullptr) {
    if (hipHostMalloc(reinterpret_cast<void**>(&p), 64,
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
      p = nullptr;
    } else {
      for (int i = 0; i < 16; ++i) p[i] = 0;  // no check id is 0
    }
  }
  return p;
}

static hipEvent_t* lag_events() {
  hipEvent_t* ev = t_readback.ev;
  for (int i = 0; i < 2; ++i) {
    if (ev[i] == nullptr && hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess)
      ev[i] = nullptr;
  }
  return (ev[0] != nullptr && ev[1] != nullptr) ? ev : nullptr;
}

// Ids of the lagged convergence checks: unique per host thread (and so per pinned
// slot pair), so a slot never holds a matching id from an earlier check or solve.
static int next_check_id() {
  static thread_local int id = 0;
  id = id == 0x7fffffff ? 1 : id + 1;
  return id;
}

int read_device_int(const int* dev, hipStream_t s, int* out) {
  int* h = pinned_counter();
  if (h == nullptr) {
    set_error("hipHostMalloc of the readback word failed");
    return STTOPT_EHIP;
  }
  hipError_t e = hipMemcpyAsync(h, dev, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return hip_status(e, "hipMemcpyAsync(readback word)");
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_status(e, "hipStreamSynchronize");
  *out = *h;
  return STTOPT_OK;
}

// The active-row count after fin_init (word 8 of the pinned block; the lagged checks use
// words 1 and 2).
static int read_init(const int* counter, hipStream_t s, int* remaining) {
  int* h = pinned_counter();
  if (h == nullptr) {
    set_error("hipHostMalloc of the readback word failed");
    return STTOPT_EHIP;
  }
  hipError_t e = hipMemcpyAsync(h + 8, counter, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return hip_status(e, "hipMemcpyAsync(readback word)");
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_status(e, "hipStreamSynchronize");
  *remaining = h[8];
  return STTOPT_OK;
}

// Queues the read of the coarsest inverse's status word into *out (0 without one), after
// the stream has waited for the inverse; *out is valid once the stream is synchronized.
static int queue_cinv_status(const sttopt_mg* mg, hipStream_t s, int* out) {
  *out = 0;
  if (mg->cinv_status == nullptr) return STTOPT_OK;
  hipError_t e = hipSuccess;
  if (mg->cinv_ready != nullptr) {
    e = hipStreamWaitEvent(s, static_cast<hipEvent_t>(mg->cinv_ready), 0);
    if (e != hipSuccess) return hip_status(e, "hipStreamWaitEvent(coarsest inverse)");
  }
  e = hipMemcpyAsync(out, mg->cinv_status, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return hip_status(e, "hipMemcpyAsync(coarsest inverse status)");
  return STTOPT_OK;
}

static int cinv_status_error(int status) {
  if (status == 0) return STTOPT_OK;
  set_error("sttopt_mgpcg: the coarsest operator is not positive definite");
  return STTOPT_ENOTSPD;
}

// The status alone (a solve with no row to iterate).
static int read_cinv_status(const sttopt_mg* mg, hipStream_t s) {
  int st = 0;
  if (int e = queue_cinv_status(mg, s, &st)) return e;
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_status(e, "hipStreamSynchronize");
  return cinv_status_error(st);
}

// Optional HIP-event timing of the outer iteration's K(rho)p launches (process-wide, off
// by default): one event pair per launch, resolved after the solve's final readback.
// Bytes per launch: 8 (2 ndof + nel) per row still active in that iteration.
struct ApplyProfile {
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<int> iter_of;  // iteration of event pair i (events 2i, 2i + 1)
  double seconds = 0.0, bytes = 0.0;
  long long launches = 0;
};
static ApplyProfile g_apply_prof;

static int prof_events(size_t pairs) {
  while (g_apply_prof.pool.size() < 2 * pairs) {
    // timing only: no system-scope fence (cache writeback and invalidation) at each
    // record, which would stall the launches that follow
    hipEvent_t e;
    hipError_t st = hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
    if (st != hipSuccess) return hip_status(st, "hipEventCreateWithFlags");
    g_apply_prof.pool.push_back(e);
  }
  return STTOPT_OK;
}

// Process-wide switch of the PCG's captured second half-iteration (on by default).
static int g_pcg_graphs = 1;

// Executable graphs of the PCG's second half-iteration, one per residual-buffer
// parity, kept per host thread and refreshed in place (hipGraphExecUpdate) by each
// solve's capture; destroyed when the thread ends.  A solve synchronizes its stream
// before it returns, so no launch of a kept graph is in flight when the next solve on
// the thread updates it.
// The graphs are recorded on a private stream of the thread (a caller's stream may be
// the null stream, which cannot be captured); recording runs nothing, and the graph is
// then launched on the solve's stream.
struct ThreadGraphs {
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipStream_t cap = nullptr;
  int capture_stream(hipStream_t* out) {
    if (cap == nullptr) {
      hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
      if (e != hipSuccess) {
        cap = nullptr;
        return hip_status(e, "hipStreamCreateWithFlags(capture stream)");
      }
    }
    *out = cap;
    return STTOPT_OK;
  }
  ~ThreadGraphs() {
    for (hipGraphExec_t& g : exec)
      if (g != nullptr) (void)hipGraphExecDestroy(g);
    if (cap != nullptr) (void)hipStreamDestroy(cap);
  }
};
static thread_local ThreadGraphs t_graphs;

// Captures the launches `body` makes on stream s into *slot: an update of the graph
// already there when the topology matches, else a new instantiation.
template <typename Body>
static int capture_graph(hipStream_t s, Body body, hipGraphExec_t* slot) {
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_status(e, "hipStreamBeginCapture");
  const int st = body();
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(s, &graph);
  if (st != STTOPT_OK) {
    if (graph != nullptr) (void)hipGraphDestroy(graph);
    return st;
  }
  if (e != hipSuccess) return hip_status(e, "hipStreamEndCapture");
  if (*slot != nullptr) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult res;
    if (hipGraphExecUpdate(*slot, graph, &bad, &res) == hipSuccess &&
        res == hipGraphExecUpdateSuccess) {
      (void)hipGraphDestroy(graph);
      return STTOPT_OK;
    }
    (void)hipGetLastError();
    (void)hipGraphExecDestroy(*slot);
    *slot = nullptr;
  }
  e = hipGraphInstantiate(slot, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    *slot = nullptr;
    return hip_status(e, "hipGraphInstantiate");
  }
  return STTOPT_OK;
}

template <typename T>
static int mgpcg_impl(const sttopt_mg* mg, const double* b, long long bs, double* x,
                      long long xs, double rtol, int max_iter, int flexible, int check_every,
                      int cold, double* work, int* iters, double* resid, int* n_iter,
                      hipStream_t s) {
  const int rows = mg->rows;
  OpDesc op;
  op.d = mg->d;
  op.nelx = mg->levels[0].nelx;
  op.nely = mg->levels[0].nely;
  op.mask = mg->levels[0].mask;
  op.rho = mg->density64;
  op.sr = mg->density64_stride;
  op.rdiag = mg->rdiag64;
  op.srd = mg->rdiag64_stride;
  for (int i = 0; i < 64; ++i) op.ke[i] = mg->ke[i];
  const long long n = op.ndof();
  const int tiles = op.tiles();
  Scratch sc = carve(mg->scratch, rows, partial_slots(op));
  double* r = work;
  T* z = reinterpret_cast<T*>(work + rows * n);  // the preconditioned residual, in T
  double* p = work + 2 * rows * n;
  double* Ap = work + 3 * rows * n;
  double* r2 = work + 4 * rows * n;  // the other residual buffer of the fused update
  // the search direction of iteration k is in pbuf[k % 3]: the fused direction update
  // (operator_dir) reads p_k on every tile's halo while writing p_{k+1}, and every other
  // launch also reads p_{k-1}, whose x update the launch before held back (one pass over
  // x per two iterations); alpha of iteration k is in slot k & 1
  double* pbuf[3] = {work + 5 * rows * n, p, work + 6 * rows * n};
  auto alpha_of = [&](int k) { return sc.alpha + (k & 1) * rows; };
  TileGeom g{op.nnx(), op.nny(), op.ntx(), tiles, tiles * rows, mg->d};
  const dim3 tblock(kBlock), rgrid(rows);
  *n_iter = 0;

  // x <- P x0;  r = P b - A x;  relative residual; active flags.  A cold start
  // (x0 = 0) writes x = 0 and r = P b in one pass, whose r.r partials are b.b's too.
  if (cold) {
    STT_TILE_LAUNCH(cold_init_kernel, double, g, s, n, b, bs, op.mask, x, xs, r, sc.part[4]);
  } else {
    if (int e = project_f64(n, rows, op.mask, x, xs, s)) return e;
    if (int e = stencil<double, double, double>(op, rows, kResid, false, kNoDot, x, xs, b, bs,
                                                nullptr, 0, r, n, 0.0, nullptr, nullptr, s))
      return e;
    STT_TILE_LAUNCH(dot_kernel, double, g, s, n, b, bs, static_cast<const double*>(nullptr), 0,
                    op.mask, nullptr, sc.part[4]);
    STT_TILE_LAUNCH(dot_kernel, double, g, s, n, r, n, static_cast<const double*>(nullptr), 0,
                    op.mask, nullptr, sc.part[1]);
  }
  // counter[0]: active rows; counter[1]: the device's iteration number
  if (int e = fill<int>(2, 1, 0, sc.counter, 2, nullptr, s)) return e;
  hipLaunchKernelGGL(fin_init_kernel, rgrid, tblock, 0, s, tiles, sc.part[4],
                     cold ? sc.part[4] : sc.part[1], rtol, sc.bsafe, resid, sc.active, iters,
                     sc.counter);
  STT_CHECK_LAUNCH("pcg init");
  int remaining = 0;
  if (int e = read_init(sc.counter, s, &remaining)) return e;
  if (remaining == 0) return read_cinv_status(mg, s);

  VCycle<T> vc(mg, s, sc.active);
  if (g_bottom_levels > 0) {
    if (int e = clear_bottom(sc, rows, s)) return e;
    vc.bsync = sc.bottom;
    vc.berr = sc.bottom_err;
  }
  const int rz_tiles = vc.rz_tiles();
  if (int e = vc.run(r, n, z, n, sc.part[2], Ap)) return e;
  hipLaunchKernelGGL(fin_rz_kernel, rgrid, tblock, 0, s, rz_tiles, sc.part[2], sc.active,
                     sc.rz, sc.counter);
  if (int e = convert<T, double>(n, rows, z, n, pbuf[1], n, sc.active, s)) return e;
  STT_CHECK_LAUNCH("pcg first preconditioner");

  // The fp32 pair cycle takes the residual update r <- r - alpha Ap in its pre-smoothing
  // pass (reading r and Ap at the staged nodes, writing the new residual to the other
  // buffer), so the convergence check follows that pass; the rows that converge there
  // skip the rest of the cycle.
  const bool fused = vc.pair();
  const int fused_tiles = pair_tiles(op);
  if (check_every < 1) check_every = 1;
  int* lag_host = pinned_counter();
  hipEvent_t* lag_ev = lag_events();
  int* lag_dev = nullptr;  // the device's address of lag_host
  if (lag_host == nullptr || lag_ev == nullptr ||
      hipHostGetDevicePointer(reinterpret_cast<void**>(&lag_dev), lag_host, 0) != hipSuccess) {
    set_error("sttopt_mgpcg: pinned readback words or events unavailable");
    return STTOPT_EHIP;
  }
  int checks = 0;  // lagged checks issued so far
  int check_ids[2] = {0, 0};  // the id each pinned slot is expected to hold
  bool done = false;
  int k_last = 0;  // iteration whose x += alpha p is still pending
  int* kdev = sc.counter + 1;  // the device's iteration number (fin_alpha advances it)
  // The residual buffers: the fused update of iteration k reads rbuf[(k - 1) & 1] and
  // writes rbuf[k & 1]; without it the residual stays in rbuf[0].
  double* rbuf[2] = {r, r2};
  // The second half of an iteration -- the coarse visit and post-smoothing (or the
  // whole cycle) and fin_beta -- is the same launch sequence every iteration up to the
  // residual buffer, so it is captured once per buffer parity into a HIP graph and
  // replayed (about 25 launches in one submission).  The direction update p = z + beta p
  // and the x += alpha p it carries run in the next iteration's apply (operator_dir).
  // Launches on stream st (the solve's stream, or the capture stream).
  auto rest = [&](int parity, hipStream_t st) -> int {
    struct Restore {
      VCycle<T>& v;
      hipStream_t s0;
      ~Restore() { v.s = s0; }
    } restore{vc, vc.s};
    vc.s = st;
    if (fused) {
      if (int e = vc.pair_rest(rbuf[parity], n, z, n, sc.part[2])) return e;
    } else if (int e = vc.run(rbuf[0], n, z, n, sc.part[2], Ap)) {
      return e;
    }
    hipLaunchKernelGGL(fin_beta_kernel, rgrid, tblock, 0, st, rz_tiles, sc.part[2],
                       fused ? fused_tiles : tiles, sc.part[3], flexible, sc.active, sc.rz,
                       sc.beta, sc.counter);
    STT_CHECK_LAUNCH("pcg beta");
    return STTOPT_OK;
  };
  const bool graphs = g_pcg_graphs && max_iter > 1;
  bool captured[2] = {false, false};
  for (int k = 1; k <= max_iter; ++k) {
    if (k == 1) {
      if (int e = stencil<double, double, double>(op, rows, kOperator, false, kDotU, pbuf[1],
                                                  n, (const double*)nullptr, 0, nullptr, 0, Ap,
                                                  n, 0.0, sc.active, sc.part[0], s))
        return e;
    } else {
      // the profiled launches record their events as part of the dispatch
      hipEvent_t t0 = nullptr, t1 = nullptr;
      if (g_apply_prof.on) {
        const size_t pi = g_apply_prof.iter_of.size();
        if (int e = prof_events(pi + 1)) return e;
        t0 = g_apply_prof.pool[2 * pi];
        t1 = g_apply_prof.pool[2 * pi + 1];
        g_apply_prof.iter_of.push_back(k);
      }
      // iteration k - 1's x update: held back when k - 1 is odd, taken with the held-back
      // one of k - 2 when it is even
      const bool two = ((k - 1) & 1) == 0;
      if (int e = operator_dir(op, rows, pbuf[(k - 1) % 3], pbuf[k % 3], z,
                               sizeof(T) == sizeof(float), x, xs, alpha_of(k - 1), sc.beta, iters,
                               k - 1, Ap, sc.active, sc.part[0], s, t0, t1,
                               two ? pbuf[(k - 2) % 3] : nullptr,
                               two ? alpha_of(k - 2) : nullptr, two ? 0 : 1))
        return e;
    }
    hipLaunchKernelGGL(fin_alpha_kernel, rgrid, tblock, 0, s, tiles, sc.part[0], sc.active,
                       sc.rz, alpha_of(k), kdev);
    const bool fuse_now = fused && k < max_iter;
    const int parity = fused ? (k & 1) : 0;
    double* rcur = fused ? rbuf[(k - 1) & 1] : rbuf[0];  // the residual entering iteration k
    if (fuse_now) {
      if (int e = vc.pair_pre_update(rcur, Ap, alpha_of(k), rbuf[k & 1], sc.part[1],
                                     flexible ? z : static_cast<T*>(nullptr), sc.part[3]))
        return e;
    } else {
      STT_TILE_LAUNCH(pcg_update_kernel, T, g, s, n, alpha_of(k), rcur, Ap,
                      flexible ? z : static_cast<T*>(nullptr), sc.active, sc.part[1],
                      sc.part[3]);
    }
    // the counter was zeroed by fin_rz (first iteration) or by the last fin_beta
    const bool check = k < max_iter && k % check_every == 0;
    const int cur = checks & 1, prev = cur ^ 1;
    if (check) check_ids[cur] = next_check_id();
    // a check's event is recorded by the fin_resid launch itself (no marker of its own
    // between the launches)
    hipExtLaunchKernelGGL(fin_resid_kernel, rgrid, tblock, 0, s, nullptr,
                          check ? lag_ev[cur] : nullptr, 0, fuse_now ? fused_tiles : tiles,
                          (const double*)sc.part[1], rtol, k, (const double*)sc.bsafe, resid,
                          sc.active, iters, sc.counter, check ? lag_dev + 1 + cur : nullptr,
                          check ? check_ids[cur] : 0);
    STT_CHECK_LAUNCH("pcg update");
    k_last = k;
    if (k == max_iter) {
      if (int e = read_device_int(sc.counter, s, &remaining)) return e;
      done = remaining == 0;
      break;
    }
    if (check) {
      // fin_resid marks this check in pinned host memory if a row is left, without a
      // wait; the host reads the previous check's mark, by which time this iteration
      // is queued behind it, so the device never drains.  An unmarked check means
      // every row converged there: this iteration's launches found no active row, and
      // the converged rows' last x += alpha p was in this iteration's fused apply (the
      // direction update of the iteration after the check).  Checks alternate between
      // two pinned slots and events.
      if (checks > 0) {
        hipError_t e = hipEventSynchronize(lag_ev[prev]);
        if (e != hipSuccess) return hip_status(e, "hipEventSynchronize");
        if (reinterpret_cast<volatile int*>(lag_host)[1 + prev] != check_ids[prev]) {
          done = true;
          k_last = 0;
          break;
        }
      }
      ++checks;
    }
    if (graphs) {
      if (!captured[parity]) {
        hipStream_t cap = nullptr;
        if (int e = t_graphs.capture_stream(&cap)) return e;
        if (int e = capture_graph(cap, [&] { return rest(parity, cap); },
                                  &t_graphs.exec[parity]))
          return e;
        captured[parity] = true;
      }
      hipError_t e = hipGraphLaunch(t_graphs.exec[parity], s);
      if (e != hipSuccess) return hip_status(e, "hipGraphLaunch");
    } else if (int e = rest(parity, s)) {
      return e;
    }
    k_last = 0;
  }
  if (k_last > 0) {
    // the last iteration's x += alpha p (no direction update follows it)
    // (after the one of k_last - 1 when the last fused apply held that back)
    const bool two = k_last >= 2 && (k_last & 1) == 0;
    STT_TILE_LAUNCH(pcg_pupdate_kernel, T, g, s, n, alpha_of(k_last), sc.beta, x, xs,
                    pbuf[k_last % 3], z, sc.active, iters, k_last, 0,
                    static_cast<const int*>(nullptr),
                    two ? static_cast<const double*>(alpha_of(k_last - 1)) : nullptr,
                    two ? static_cast<const double*>(pbuf[(k_last - 1) % 3]) : nullptr);
    STT_CHECK_LAUNCH("pcg last x update");
  }
  // n_iter: the slowest row's count (and whether a wait of the persistent bottom pass
  // gave up, read with it)
  std::vector<int> h(rows > 0 ? rows : 1);
  int bottom_err = 0, cinv_st = 0;
  hipError_t e = hipMemcpyAsync(h.data(), iters, sizeof(int) * rows, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return hip_status(e, "hipMemcpyAsync(iters)");
  if (int ec = queue_cinv_status(mg, s, &cinv_st)) return ec;
  if (vc.bottom_used) {
    e = hipMemcpyAsync(&bottom_err, sc.bottom_err, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_status(e, "hipMemcpyAsync(bottom error)");
  }
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_status(e, "hipStreamSynchronize");
  if (int ec = cinv_status_error(cinv_st)) return ec;
  if (bottom_err != 0) {
    set_error("sttopt_mgpcg: a wait of the persistent V-cycle bottom pass gave up");
    return STTOPT_EHIP;
  }
  int mx = 0;
  for (int i = 0; i < rows; ++i) mx = h[i] > mx ? h[i] : mx;
  if (g_apply_prof.on && !g_apply_prof.iter_of.empty()) {
    // the fused launches (iteration 2 on; iteration 1's plain apply is not counted):
    // algorithmic bytes per row still iterating: p_old, z and rho read, p_new and Ap
    // written, and in the launches that take two iterations' x updates (k - 1 even) also
    // x and p_prev read and x written; per row that retired in the iteration before:
    // p_old (and p_prev) and x read, x written
    const double nd = (double)n, ne = (double)op.nel();
    for (size_t i = 0; i < g_apply_prof.iter_of.size(); ++i) {
      float ms = 0.0f;
      e = hipEventElapsedTime(&ms, g_apply_prof.pool[2 * i], g_apply_prof.pool[2 * i + 1]);
      if (e != hipSuccess) return hip_status(e, "hipEventElapsedTime");
      const int k = g_apply_prof.iter_of[i];
      if (k == 1) continue;
      const bool two = ((k - 1) & 1) == 0;
      const double fused = 8.0 * ((two ? 6.0 : 3.0) * nd + ne) + (double)sizeof(T) * nd;
      const double retiring = 8.0 * (two ? 4.0 : 3.0) * nd;
      int act = 0, ret = 0;
      for (int r = 0; r < rows; ++r) {
        act += h[r] >= k ? 1 : 0;
        ret += h[r] == k - 1 ? 1 : 0;
      }
      if (act == 0) continue;  // the lagged check's launch after convergence: no rows
      g_apply_prof.seconds += ms * 1e-3;
      g_apply_prof.bytes += fused * act + retiring * ret;
      g_apply_prof.launches += 1;
    }
    g_apply_prof.iter_of.clear();
  }
  *n_iter = done ? mx : max_iter;
  if (!done) {
    set_error("sttopt_mgpcg: rows above rtol after max_iter iterations");
    return STTOPT_ENOCONV;
  }
  return STTOPT_OK;
}

// ------------------------------------------------------------ uneven batch ----
// One batched MG-PCG over hierarchies of different grids (sttopt_mgpcg_batch): every
// pass of an iteration is one launch over all hierarchies (kMaxBatchGroups per launch),
// each hierarchy taking its own range of workgroups with its own arguments; the V-cycle
// visits level l of every hierarchy that has it in the same launch, and the coarsest
// solves of all hierarchies (at their own depths) in one launch.  Rows converge and
// retire per row, as in sttopt_mgpcg, whose arithmetic every pass repeats.

// Finalizers and the direction update over the rows of several hierarchies: workgroup
// b is row b - first[i] of hierarchy i.
struct FinCall {
  int tiles, tiles_o, flexible, k;
  const double* p0;  // alpha: p.Ap; resid: r.r; beta: r.z
  const double* p1;  // beta: r.z_old
  const double* bsafe;
  const int* active_c;
  int* active;
  double *rz, *alpha, *beta, *resid;
  int *iters, *counter, *kdev;
};
struct FinGroups {
  int n;
  int first[kMaxBatchGroups + 1];
  FinCall a[kMaxBatchGroups];
};

__device__ __forceinline__ int fin_group(const FinGroups& g, int b) {
  int i = 0;
  while (i + 1 < g.n && b >= g.first[i + 1]) ++i;
  return i;
}

__global__ __launch_bounds__(kBlock) void batch_alpha_kernel(FinGroups g) {
  __shared__ double red[kBlock / kWave];
  const int i = fin_group(g, blockIdx.x);
  const FinCall& c = g.a[i];
  const int row = blockIdx.x - g.first[i];
  if (row == 0 && threadIdx.x == 0) c.kdev[0] += 1;
  if (c.active[row] == 0) return;
  const double pap = row_sum(c.p0, c.tiles, row, red);
  if (threadIdx.x == 0) c.alpha[row] = safe_div(c.rz[row], pap);
}

__global__ __launch_bounds__(kBlock) void batch_resid_kernel(FinGroups g, double rtol,
                                                             int* host_flag, int check_id) {
  __shared__ double red[kBlock / kWave];
  const int i = fin_group(g, blockIdx.x);
  const FinCall& c = g.a[i];
  const int row = blockIdx.x - g.first[i];
  if (c.active[row] == 0) return;
  const double rr = row_sum(c.p0, c.tiles, row, red);
  if (threadIdx.x == 0) {
    const double rel = sqrt(rr) / c.bsafe[row];
    c.resid[row] = rel;
    c.iters[row] = c.k;
    if (rel <= rtol) {
      c.active[row] = 0;
    } else {
      atomicAdd(c.counter, 1);
      if (host_flag != nullptr)
        __hip_atomic_store(host_flag, check_id, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

__global__ __launch_bounds__(kBlock) void batch_beta_kernel(FinGroups g) {
  __shared__ double red[kBlock / kWave];
  const int i = fin_group(g, blockIdx.x);
  const FinCall& c = g.a[i];
  const int row = blockIdx.x - g.first[i];
  if (row == 0 && threadIdx.x == 0) *c.counter = 0;
  if (c.active[row] == 0) return;
  const double rzn = row_sum(c.p0, c.tiles, row, red);
  double rzo = 0.0;
  if (c.flexible) {
    __syncthreads();
    rzo = row_sum(c.p1, c.tiles_o, row, red);
  }
  if (threadIdx.x == 0) {
    c.beta[row] = safe_div(rzn - rzo, c.rz[row]);
    c.rz[row] = rzn;
  }
}

// x += alpha p and p = z + beta p for the rows that took iteration *kdev (the body of
// pcg_pupdate_kernel with update_p, tile origin from the group-local workgroup).
struct PupdateCall {
  TileGeom g;
  long long n, xs;
  const double *alpha, *beta;
  double *x, *p;
  const float* z;
  const int *active, *iters, *kdev;
};
struct PupdateGroups {
  int n;
  int first[kMaxBatchGroups + 1];
  PupdateCall a[kMaxBatchGroups];
};

template <int D>
__global__ __launch_bounds__(kBlock) void batch_pupdate_kernel(PupdateGroups gs) {
  int gi = 0;
  while (gi + 1 < gs.n && (int)blockIdx.x >= gs.first[gi + 1]) ++gi;
  const PupdateCall& c = gs.a[gi];
  const TileGeom& g = c.g;
  const int t = xcd_remap(blockIdx.x - gs.first[gi], g.total);
  int row, tt;
  row_tile(t, g.tiles, g.total, row, tt);
  const int by = tt / g.ntx, bx = tt - by * g.ntx;
  const int cc = bx * kTileX + (threadIdx.x & (kTileX - 1));
  const int r0 = by * kTileY + (threadIdx.x / kTileX) * kRowsPerThread;
  const int k = *c.kdev;
  if (c.iters[row] != k || cc >= g.nnx) return;
  const bool dir = c.active[row] != 0;
  const double a = c.alpha[row];
  const double b = dir ? c.beta[row] : 0.0;
  double* xr = c.x + row * c.xs;
  double* pr = c.p + row * c.n;
  const float* zr = c.z + row * c.n;
#pragma unroll
  for (int j = 0; j < kRowsPerThread; ++j) {
    if (r0 + j < g.nny) {
      const long long node = (long long)(r0 + j) * g.nnx + cc;
      Vec<double, D> pv = load_node<double, D>(pr, node);
      Vec<double, D> xv = load_node<double, D>(xr, node);
#pragma unroll
      for (int q = 0; q < D; ++q) xv.v[q] = fma(a, pv.v[q], xv.v[q]);
      store_node<double, D>(xr, node, xv);
      if (dir) {
        const Vec<float, D> zv = load_node<float, D>(zr, node);
#pragma unroll
        for (int q = 0; q < D; ++q) pv.v[q] = (double)zv.v[q] + b * pv.v[q];
        store_node<double, D>(pr, node, pv);
      }
    }
  }
}

// One hierarchy of the batch.
struct BatchGroup {
  const sttopt_mg* mg = nullptr;
  int rows = 0, L = 0;
  OpDesc op;  // the fp64 outer operator
  long long n = 0, xs = 0, bs = 0;
  int tiles = 0, ptiles = 0;
  Scratch sc{};
  double *x = nullptr, *p = nullptr, *Ap = nullptr;
  double* rbuf[2] = {nullptr, nullptr};
  float* z = nullptr;
  const double* b = nullptr;
  int* iters = nullptr;
  double* resid = nullptr;
  VCycle<float>* vc = nullptr;
  TileGeom g{};

  float* lbuf(int l, void* sttopt_level::*f) const {
    return static_cast<float*>(mg->levels[l].*f);
  }
  // where the V-cycle leaves level l's correction: vx at the coarsest level and at the
  // tile-smoothed level 1, vx2 at a node-stencil level (its post pass writes there)
  float* result(int l) const { return (l == L - 1 || l == 1) ? lbuf(l, &sttopt_level::vx) : lbuf(l, &sttopt_level::vx2); }
};

// Calls fn(chunk of at most kMaxBatchGroups entries) over a list.
template <typename C, typename F>
static int in_chunks(const std::vector<C>& calls, F fn) {
  for (size_t i = 0; i < calls.size(); i += kMaxBatchGroups) {
    const int m = (int)std::min<size_t>(kMaxBatchGroups, calls.size() - i);
    if (int e = fn(calls.data() + i, m)) return e;
  }
  return STTOPT_OK;
}

static int launch_fin(int which, const std::vector<FinCall>& calls, const std::vector<int>& rows,
                      double rtol, int* host_flag, int check_id, hipStream_t s) {
  for (size_t i0 = 0; i0 < calls.size(); i0 += kMaxBatchGroups) {
    FinGroups g{};
    int total = 0;
    for (size_t i = i0; i < calls.size() && g.n < kMaxBatchGroups; ++i) {
      g.a[g.n] = calls[i];
      g.first[g.n] = total;
      total += rows[i];
      ++g.n;
    }
    g.first[g.n] = total;
    if (which == 0)
      hipLaunchKernelGGL(batch_alpha_kernel, dim3(total), dim3(kBlock), 0, s, g);
    else if (which == 1)
      hipLaunchKernelGGL(batch_resid_kernel, dim3(total), dim3(kBlock), 0, s, g, rtol,
                         host_flag, check_id);
    else
      hipLaunchKernelGGL(batch_beta_kernel, dim3(total), dim3(kBlock), 0, s, g);
    STT_CHECK_LAUNCH("batch finalizer");
  }
  return STTOPT_OK;
}

static int batch_validate(const BatchGroup& G, int d) {
  const sttopt_mg* mg = G.mg;
  if (mg->vdtype != STTOPT_F32 || mg->n_smooth != 2 || mg->gamma != 1 || mg->nlevels < 2 ||
      mg->d != d || !G.vc->pair()) {
    set_error("sttopt_mgpcg_batch: every hierarchy needs the fp32 two-sweep cycle coarsened "
              "2 x 2 from level 0, at least two levels and the same dofs per node");
    return STTOPT_EINVAL;
  }
  for (int l = 1; l < mg->nlevels - 1; ++l) {
    const sttopt_level& Lv = mg->levels[l];
    const OpDesc& o = G.vc->op[l];
    const bool ok = Lv.kx == 2 && Lv.ky == 2 &&
                    ((l == 1 && o.ptab != nullptr && o.pkx == 2 && o.pky == 2) ||
                     (l >= 2 && o.keff_node));
    if (!ok) {
      set_error("sttopt_mgpcg_batch: level 1 must be the density-parameterized 2 x 2 level "
                "and deeper levels node stencils, each coarsened 2 x 2");
      return STTOPT_EINVAL;
    }
  }
  return STTOPT_OK;
}

// The second half of an iteration for every hierarchy: the coarse visit below level 0,
// the level-0 post-smoothing, the beta finalizer and the direction update.
static int batch_rest(std::vector<BatchGroup>& G, int parity, double omega, hipStream_t s) {
  const int ng = (int)G.size();
  int Lmax = 0;
  for (auto& g : G) Lmax = std::max(Lmax, g.L);
  // restriction from level 0
  {
    std::vector<TransferCall> c;
    for (auto& g : G) {
      const sttopt_level& L1 = g.mg->levels[1];
      c.push_back({g.mg->d, L1.nelx, L1.nely, g.rows, g.lbuf(0, &sttopt_level::vr), g.n,
                   L1.mask, g.lbuf(1, &sttopt_level::vb), g.vc->n[1], g.sc.active});
    }
    if (int e = in_chunks(c, [&](const TransferCall* p, int m) {
          return transfer22_groups(p, m, true, s);
        }))
      return e;
  }
  // down: level l of every hierarchy that has l above its coarsest level
  for (int l = 1; l <= Lmax - 2; ++l) {
    if (l == 1) {
      std::vector<StencilCall<float>> first, resid;
      std::vector<TransferCall> down;
      for (auto& g : G) {
        if (g.L - 1 <= 1) continue;
        const VCycle<float>& v = *g.vc;
        const long long n1 = v.n[1];
        float* vb = g.lbuf(1, &sttopt_level::vb);
        float* vx = g.lbuf(1, &sttopt_level::vx);
        float* vr = g.lbuf(1, &sttopt_level::vr);
        first.push_back({v.op[1], g.rows, kSmoothFirst, kNoDot, nullptr, 0, vb, n1, v.invd(1),
                         v.sd(1), vx, n1, omega, g.sc.active, nullptr});
        resid.push_back({v.op[1], g.rows, kResid, kNoDot, vx, n1, vb, n1, nullptr, 0, vr, n1,
                         0.0, g.sc.active, nullptr});
        const sttopt_level& L2 = g.mg->levels[2];
        down.push_back({g.mg->d, L2.nelx, L2.nely, g.rows, vr, n1, L2.mask,
                        g.lbuf(2, &sttopt_level::vb), v.n[2], g.sc.active});
      }
      if (first.empty()) continue;
      auto st = [&](const StencilCall<float>* p, int m) { return stencil_tile_groups<float>(p, m, s); };
      if (int e = in_chunks(first, st)) return e;
      if (int e = in_chunks(resid, st)) return e;
      if (int e = in_chunks(down, [&](const TransferCall* p, int m) {
            return transfer22_groups(p, m, true, s);
          }))
        return e;
    } else {
      std::vector<PairLevel> pre;
      int variant = -1;
      for (auto& g : G) {
        if (g.L - 1 <= l) continue;
        const VCycle<float>& v = *g.vc;
        PairLevel P{};
        P.op = v.op[l];
        P.rows = g.rows;
        P.b = g.lbuf(l, &sttopt_level::vb);
        P.sb = v.n[l];
        P.invd = v.invd(l);
        P.sd = v.sd(l);
        P.x2 = g.lbuf(l, &sttopt_level::vx);
        P.sx2 = v.n[l];
        P.cmask = g.mg->levels[l + 1].mask;
        P.bc = g.lbuf(l + 1, &sttopt_level::vb);
        P.sbc = v.n[l + 1];
        P.omega = omega;
        P.active = g.sc.active;
        pre.push_back(P);
        if (variant < 0) variant = (g.mg->levels[l].flags >> STTOPT_TILE_VARIANT_SHIFT) & 15;
      }
      if (pre.empty()) continue;
      if (int e = in_chunks(pre, [&](const PairLevel* p, int m) {
            return coarse_pair_groups(p, m, true, variant, s);
          }))
        return e;
    }
  }
  // the coarsest solves, each hierarchy at its own depth
  {
    std::vector<CoarseSolveCall> c;
    for (auto& g : G) {
      const int lc = g.L - 1;
      const sttopt_level& Lc = g.mg->levels[lc];
      const long long nc = g.vc->n[lc];
      c.push_back({nc, g.rows, Lc.cinv, Lc.cinv_stride, g.lbuf(lc, &sttopt_level::vb), nc,
                   Lc.mask, g.lbuf(lc, &sttopt_level::vx), nc, g.sc.active});
    }
    if (int e = in_chunks(c, [&](const CoarseSolveCall* p, int m) {
          return coarse_solve_groups(p, m, s);
        }))
      return e;
  }
  // up
  for (int l = Lmax - 2; l >= 1; --l) {
    if (l >= 2) {
      std::vector<PairLevel> post;
      int variant = -1;
      for (auto& g : G) {
        if (g.L - 1 <= l) continue;
        const VCycle<float>& v = *g.vc;
        PairLevel P{};
        P.op = v.op[l];
        P.rows = g.rows;
        P.x2 = g.lbuf(l, &sttopt_level::vx);
        P.sx2 = v.n[l];
        P.ec = g.result(l + 1);
        P.sec = v.n[l + 1];
        P.b = g.lbuf(l, &sttopt_level::vb);
        P.sb = v.n[l];
        P.invd = v.invd(l);
        P.sd = v.sd(l);
        P.x4 = g.lbuf(l, &sttopt_level::vx2);
        P.sx4 = v.n[l];
        P.omega = omega;
        P.active = g.sc.active;
        post.push_back(P);
        if (variant < 0) variant = (g.mg->levels[l].flags >> STTOPT_TILE_VARIANT_SHIFT) & 15;
      }
      if (post.empty()) continue;
      if (int e = in_chunks(post, [&](const PairLevel* p, int m) {
            return coarse_pair_groups(p, m, false, variant, s);
          }))
        return e;
    } else {
      std::vector<TransferCall> up;
      std::vector<StencilCall<float>> sw1, sw2;
      for (auto& g : G) {
        if (g.L - 1 <= 1) continue;
        const VCycle<float>& v = *g.vc;
        const long long n1 = v.n[1];
        const sttopt_level& L2 = g.mg->levels[2];
        float* vb = g.lbuf(1, &sttopt_level::vb);
        float* vx = g.lbuf(1, &sttopt_level::vx);
        float* vx2 = g.lbuf(1, &sttopt_level::vx2);
        up.push_back({g.mg->d, L2.nelx, L2.nely, g.rows, g.result(2), v.n[2],
                      g.mg->levels[1].mask, vx, n1, g.sc.active});
        sw1.push_back({v.op[1], g.rows, kSmooth, kNoDot, vx, n1, vb, n1, v.invd(1), v.sd(1),
                       vx2, n1, omega, g.sc.active, nullptr});
        sw2.push_back({v.op[1], g.rows, kSmooth, kNoDot, vx2, n1, vb, n1, v.invd(1), v.sd(1),
                       vx, n1, omega, g.sc.active, nullptr});
      }
      if (up.empty()) continue;
      if (int e = in_chunks(up, [&](const TransferCall* p, int m) {
            return transfer22_groups(p, m, false, s);
          }))
        return e;
      auto st = [&](const StencilCall<float>* p, int m) { return stencil_tile_groups<float>(p, m, s); };
      if (int e = in_chunks(sw1, st)) return e;
      if (int e = in_chunks(sw2, st)) return e;
    }
  }
  // level-0 post-smoothing: z and the r.z partials
  {
    std::vector<PairPostCall<float>> c;
    for (auto& g : G) {
      const VCycle<float>& v = *g.vc;
      c.push_back({v.op[0], g.rows, g.lbuf(0, &sttopt_level::vx), g.n, g.result(1), v.n[1],
                   g.mg->levels[1].nelx + 1, g.rbuf[parity], g.n, v.invd(0), v.sd(0), g.z,
                   g.n, omega, g.sc.active, g.sc.part[2]});
    }
    if (int e = in_chunks(c, [&](const PairPostCall<float>* p, int m) {
          return smooth_pair_post_groups(p, m, s);
        }))
      return e;
  }
  // beta, then the direction update
  std::vector<FinCall> fb;
  std::vector<int> rows;
  std::vector<PupdateCall> pu;
  for (auto& g : G) {
    FinCall f{};
    f.tiles = g.ptiles;
    f.tiles_o = g.ptiles;
    f.flexible = 1;
    f.p0 = g.sc.part[2];
    f.p1 = g.sc.part[3];
    f.active = g.sc.active;
    f.rz = g.sc.rz;
    f.beta = g.sc.beta;
    f.counter = g.sc.counter;
    fb.push_back(f);
    rows.push_back(g.rows);
    pu.push_back({g.g, g.n, g.xs, g.sc.alpha, g.sc.beta, g.x, g.p, g.z, g.sc.active, g.iters,
                  g.sc.counter + 1});
  }
  if (int e = launch_fin(2, fb, rows, 0.0, nullptr, 0, s)) return e;
  for (size_t i0 = 0; i0 < pu.size(); i0 += kMaxBatchGroups) {
    PupdateGroups pg{};
    int total = 0;
    for (size_t i = i0; i < pu.size() && pg.n < kMaxBatchGroups; ++i) {
      pg.a[pg.n] = pu[i];
      pg.first[pg.n] = total;
      total += pu[i].g.total;
      ++pg.n;
    }
    pg.first[pg.n] = total;
    if (G[0].mg->d == 2)
      hipLaunchKernelGGL(batch_pupdate_kernel<2>, dim3(total), dim3(kBlock), 0, s, pg);
    else
      hipLaunchKernelGGL(batch_pupdate_kernel<1>, dim3(total), dim3(kBlock), 0, s, pg);
    STT_CHECK_LAUNCH("batch_pupdate_kernel");
  }
  (void)ng;
  return STTOPT_OK;
}

static thread_local ThreadGraphs t_batch_graphs;

static int mgpcg_batch_impl(const sttopt_batch_group* in, int ng, double rtol, int max_iter,
                            int check_every, int cold, int* n_iter, hipStream_t s) {
  std::vector<BatchGroup> G(ng);
  std::vector<std::unique_ptr<VCycle<float>>> vcs;
  const int d = in[0].mg->d;
  const double omega = in[0].mg->omega;
  for (int i = 0; i < ng; ++i) {
    BatchGroup& g = G[i];
    const sttopt_mg* mg = in[i].mg;
    g.mg = mg;
    g.rows = mg->rows;
    g.L = mg->nlevels;
    g.op.d = mg->d;
    g.op.nelx = mg->levels[0].nelx;
    g.op.nely = mg->levels[0].nely;
    g.op.mask = mg->levels[0].mask;
    g.op.rho = mg->density64;
    g.op.sr = mg->density64_stride;
    g.op.rdiag = mg->rdiag64;
    g.op.srd = mg->rdiag64_stride;
    for (int q = 0; q < 64; ++q) g.op.ke[q] = mg->ke[q];
    g.n = g.op.ndof();
    g.tiles = g.op.tiles();
    g.ptiles = pair_tiles(g.op);
    g.sc = carve(mg->scratch, g.rows, partial_slots(g.op));
    double* w = in[i].work;
    g.rbuf[0] = w;
    g.z = reinterpret_cast<float*>(w + g.rows * g.n);
    g.p = w + 2 * g.rows * g.n;
    g.Ap = w + 3 * g.rows * g.n;
    g.rbuf[1] = w + 4 * g.rows * g.n;
    g.x = in[i].x;
    g.xs = in[i].x_stride;
    g.b = in[i].b;
    g.bs = in[i].b_stride;
    g.iters = in[i].iters;
    g.resid = in[i].resid;
    g.g = TileGeom{g.op.nnx(), g.op.nny(), g.op.ntx(), g.tiles, g.tiles * g.rows, d};
    vcs.emplace_back(new VCycle<float>(mg, s, g.sc.active));
    g.vc = vcs.back().get();
    if (int e = batch_validate(g, d)) return e;
    if (mg->omega != omega) {
      set_error("sttopt_mgpcg_batch: one omega for all hierarchies");
      return STTOPT_EINVAL;
    }
  }
  *n_iter = 0;
  const dim3 tblock(kBlock);
  // initial residuals, per hierarchy (once per solve)
  int remaining = 0;
  for (auto& g : G) {
    const TileGeom& tg = g.g;
    if (cold) {
      STT_TILE_LAUNCH(cold_init_kernel, double, tg, s, g.n, g.b, g.bs, g.op.mask, g.x, g.xs,
                      g.rbuf[0], g.sc.part[4]);
    } else {
      if (int e = project_f64(g.n, g.rows, g.op.mask, g.x, g.xs, s)) return e;
      if (int e = stencil<double, double, double>(g.op, g.rows, kResid, false, kNoDot, g.x,
                                                  g.xs, g.b, g.bs, nullptr, 0, g.rbuf[0], g.n,
                                                  0.0, nullptr, nullptr, s))
        return e;
      STT_TILE_LAUNCH(dot_kernel, double, tg, s, g.n, g.b, g.bs,
                      static_cast<const double*>(nullptr), 0, g.op.mask, nullptr, g.sc.part[4]);
      STT_TILE_LAUNCH(dot_kernel, double, tg, s, g.n, g.rbuf[0], g.n,
                      static_cast<const double*>(nullptr), 0, g.op.mask, nullptr, g.sc.part[1]);
    }
    if (int e = fill<int>(2, 1, 0, g.sc.counter, 2, nullptr, s)) return e;
    hipLaunchKernelGGL(fin_init_kernel, dim3(g.rows), tblock, 0, s, g.tiles, g.sc.part[4],
                       cold ? g.sc.part[4] : g.sc.part[1], rtol, g.sc.bsafe, g.resid,
                       g.sc.active, g.iters, g.sc.counter);
    STT_CHECK_LAUNCH("batch init");
  }
  for (auto& g : G) {
    int rem = 0;
    if (int e = read_init(g.sc.counter, s, &rem)) return e;
    remaining += rem;
  }
  if (remaining == 0) {
    for (auto& g : G)
      if (int e = read_cinv_status(g.mg, s)) return e;
    return STTOPT_OK;
  }
  for (auto& g : G) {
    if (int e = g.vc->run(g.rbuf[0], g.n, g.z, g.n, g.sc.part[2], g.Ap)) return e;
    hipLaunchKernelGGL(fin_rz_kernel, dim3(g.rows), tblock, 0, s, g.vc->rz_tiles(),
                       g.sc.part[2], g.sc.active, g.sc.rz, g.sc.counter);
    if (int e = convert<float, double>(g.n, g.rows, g.z, g.n, g.p, g.n, g.sc.active, s))
      return e;
    STT_CHECK_LAUNCH("batch first preconditioner");
  }
  if (check_every < 1) check_every = 1;
  int* lag_host = pinned_counter();
  hipEvent_t* lag_ev = lag_events();
  int* lag_dev = nullptr;
  if (lag_host == nullptr || lag_ev == nullptr ||
      hipHostGetDevicePointer(reinterpret_cast<void**>(&lag_dev), lag_host, 0) != hipSuccess) {
    set_error("sttopt_mgpcg_batch: pinned readback words or events unavailable");
    return STTOPT_EHIP;
  }
  std::vector<int> rows;
  for (auto& g : G) rows.push_back(g.rows);
  int checks = 0;
  int check_ids[2] = {0, 0};
  bool done = false;
  int k_last = 0;
  bool captured[2] = {false, false};
  const bool graphs = g_pcg_graphs && max_iter > 1;
  for (int k = 1; k <= max_iter; ++k) {
    // outer apply with the p.Ap partials, alpha
    std::vector<StencilCall<double>> ap;
    std::vector<FinCall> fa, fr;
    for (auto& g : G) {
      ap.push_back({g.op, g.rows, kOperator, kDotU, g.p, g.n, nullptr, 0, nullptr, 0, g.Ap,
                    g.n, 0.0, g.sc.active, g.sc.part[0]});
      FinCall f{};
      f.tiles = g.tiles;
      f.p0 = g.sc.part[0];
      f.active = g.sc.active;
      f.rz = g.sc.rz;
      f.alpha = g.sc.alpha;
      f.kdev = g.sc.counter + 1;
      fa.push_back(f);
    }
    if (int e = in_chunks(ap, [&](const StencilCall<double>* p, int m) {
          return stencil_tile_groups<double>(p, m, s);
        }))
      return e;
    if (int e = launch_fin(0, fa, rows, 0.0, nullptr, 0, s)) return e;
    const bool last = k == max_iter;
    // the residual update (fused into level 0's pre-smoothing except at the last
    // iteration)
    if (!last) {
      std::vector<PairPreCall> pre;
      for (auto& g : G) {
        const VCycle<float>& v = *g.vc;
        pre.push_back({v.op[0], g.rows, g.rbuf[(k - 1) & 1], g.n, v.invd(0), v.sd(0),
                       g.lbuf(0, &sttopt_level::vx), g.n, g.lbuf(0, &sttopt_level::vr), g.n,
                       omega, g.sc.active, g.Ap, g.n, g.sc.alpha, g.rbuf[k & 1], g.n,
                       g.sc.part[1], g.z, g.n, g.sc.part[3]});
      }
      if (int e = in_chunks(pre, [&](const PairPreCall* p, int m) {
            return smooth_pair_pre_groups(p, m, s);
          }))
        return e;
    } else {
      for (auto& g : G) {
        STT_TILE_LAUNCH(pcg_update_kernel, float, g.g, s, g.n, g.sc.alpha, g.rbuf[(k - 1) & 1],
                        g.Ap, g.z, g.sc.active, g.sc.part[1], g.sc.part[3]);
      }
    }
    const bool check = !last && k % check_every == 0;
    const int cur = checks & 1, prev = cur ^ 1;
    if (check) check_ids[cur] = next_check_id();
    for (auto& g : G) {
      FinCall f{};
      f.tiles = last ? g.tiles : g.ptiles;
      f.k = k;
      f.p0 = g.sc.part[1];
      f.bsafe = g.sc.bsafe;
      f.resid = g.resid;
      f.active = g.sc.active;
      f.iters = g.iters;
      f.counter = g.sc.counter;
      fr.push_back(f);
    }
    if (int e = launch_fin(1, fr, rows, rtol, check ? lag_dev + 1 + cur : nullptr,
                           check ? check_ids[cur] : 0, s))
      return e;
    k_last = k;
    if (last) {
      remaining = 0;
      for (auto& g : G) {
        int rem = 0;
        if (int e = read_device_int(g.sc.counter, s, &rem)) return e;
        remaining += rem;
      }
      done = remaining == 0;
      break;
    }
    if (check) {
      hipError_t e = hipEventRecord(lag_ev[cur], s);
      if (e != hipSuccess) return hip_status(e, "hipEventRecord");
      if (checks > 0) {
        e = hipEventSynchronize(lag_ev[prev]);
        if (e != hipSuccess) return hip_status(e, "hipEventSynchronize");
        if (reinterpret_cast<volatile int*>(lag_host)[1 + prev] != check_ids[prev]) {
          done = true;
          k_last = 0;
          break;
        }
      }
      ++checks;
    }
    const int parity = k & 1;
    if (graphs) {
      if (!captured[parity]) {
        hipStream_t cap = nullptr;
        if (int e = t_batch_graphs.capture_stream(&cap)) return e;
        if (int e = capture_graph(cap, [&] {
              for (auto& g : G) g.vc->s = cap;
              const int st = batch_rest(G, parity, omega, cap);
              for (auto& g : G) g.vc->s = s;
              return st;
            }, &t_batch_graphs.exec[parity]))
          return e;
        captured[parity] = true;
      }
      hipError_t e = hipGraphLaunch(t_batch_graphs.exec[parity], s);
      if (e != hipSuccess) return hip_status(e, "hipGraphLaunch");
    } else if (int e = batch_rest(G, parity, omega, s)) {
      return e;
    }
    k_last = 0;
  }
  if (k_last > 0) {
    for (auto& g : G) {
      STT_TILE_LAUNCH(pcg_pupdate_kernel, float, g.g, s, g.n, g.sc.alpha, g.sc.beta, g.x, g.xs,
                      g.p, g.z, g.sc.active, g.iters, k_last, 0, static_cast<const int*>(nullptr),
                      static_cast<const double*>(nullptr), static_cast<const double*>(nullptr));
      STT_CHECK_LAUNCH("batch last x update");
    }
  }
  int mx = 0;
  for (auto& g : G) {
    std::vector<int> h(g.rows);
    int cinv_st = 0;
    hipError_t e = hipMemcpyAsync(h.data(), g.iters, sizeof(int) * g.rows,
                                  hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_status(e, "hipMemcpyAsync(iters)");
    if (int ec = queue_cinv_status(g.mg, s, &cinv_st)) return ec;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e, "hipStreamSynchronize");
    if (int ec = cinv_status_error(cinv_st)) return ec;
    for (int v : h) mx = std::max(mx, v);
  }
  *n_iter = done ? mx : max_iter;
  if (!done) {
    set_error("sttopt_mgpcg_batch: rows above rtol after max_iter iterations");
    return STTOPT_ENOCONV;
  }
  return STTOPT_OK;
}

}  // namespace stt

using namespace stt;

extern "C" {

int64_t sttopt_mg_scratch_bytes(const sttopt_mg* mg) {
  if (mg == nullptr || mg->nlevels < 1) return 0;
  OpDesc op;
  op.d = mg->d;
  op.nelx = mg->levels[0].nelx;
  op.nely = mg->levels[0].nely;
  return scratch_bytes(mg->rows, partial_slots(op));
}

int64_t sttopt_mgpcg_work_count(const sttopt_mg* mg) {
  if (mg == nullptr || mg->nlevels < 1) return 0;
  const long long n = (long long)(mg->levels[0].nelx + 1) * (mg->levels[0].nely + 1) * mg->d;
  // r, z, p, Ap, the second residual buffer of the fused update and the second and third
  // search direction buffers of the fused direction update
  return 7LL * mg->rows * n;
}

int sttopt_mgpcg_graphs(int32_t enable) {
  const int prev = g_pcg_graphs;
  if (enable >= 0) g_pcg_graphs = enable != 0;
  return prev;
}

int sttopt_profile_outer_apply(int32_t enable) {
  g_apply_prof.on = enable != 0;
  g_apply_prof.iter_of.clear();
  g_apply_prof.seconds = g_apply_prof.bytes = 0.0;
  g_apply_prof.launches = 0;
  return STTOPT_OK;
}

int sttopt_profile_outer_apply_read(double* seconds, double* bytes, int64_t* launches) {
  if (seconds == nullptr || bytes == nullptr || launches == nullptr) {
    set_error("sttopt_profile_outer_apply_read: null pointer");
    return STTOPT_EINVAL;
  }
  *seconds = g_apply_prof.seconds;
  *bytes = g_apply_prof.bytes;
  *launches = g_apply_prof.launches;
  return STTOPT_OK;
}

int sttopt_vcycle(const sttopt_mg* mg, const double* r, int64_t r_stride, double* z,
                  int64_t z_stride, void* stream) {
  if (int e = validate(mg)) return e;
  if (r == nullptr || z == nullptr) {
    set_error("sttopt_vcycle: null vector");
    return STTOPT_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  if (mg->vdtype == STTOPT_F64) {
    VCycle<double> vc(mg, s, nullptr);
    return vc.run(r, r_stride, z, z_stride, nullptr);
  }
  VCycle<float> vc(mg, s, nullptr);
  Scratch sc{};
  if (g_bottom_levels > 0 && mg->scratch != nullptr) {
    OpDesc op0;
    op0.d = mg->d;
    op0.nelx = mg->levels[0].nelx;
    op0.nely = mg->levels[0].nely;
    sc = carve(mg->scratch, mg->rows, partial_slots(op0));
    if (int e = clear_bottom(sc, mg->rows, s)) return e;
    vc.bsync = sc.bottom;
    vc.berr = sc.bottom_err;
  }
  if (int e = vc.run(r, r_stride, z, z_stride, nullptr)) return e;
  if (vc.bottom_used) {
    int bottom_err = 0;
    if (int e = read_device_int(sc.bottom_err, s, &bottom_err)) return e;
    if (bottom_err != 0) {
      set_error("sttopt_vcycle: a wait of the persistent bottom pass gave up");
      return STTOPT_EHIP;
    }
  }
  return STTOPT_OK;
}

int sttopt_vcycle_bottom(int32_t max_levels) {
  if (max_levels < 0 || max_levels > kMaxBottomLevels) {
    set_error("sttopt_vcycle_bottom: 0 to 3 levels");
    return STTOPT_EINVAL;
  }
  g_bottom_levels = max_levels;
  return STTOPT_OK;
}

int sttopt_level_apply(const sttopt_mg* mg, int32_t level, const void* v, int64_t v_stride,
                       void* y, int64_t y_stride, void* stream) {
  if (int e = validate(mg)) return e;
  if (level < 0 || level >= mg->nlevels || v == nullptr || y == nullptr) {
    set_error("sttopt_level_apply: bad level or null vector");
    return STTOPT_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  if (mg->vdtype == STTOPT_F64) {
    VCycle<double> vc(mg, s, nullptr);
    return stencil<double, double, double>(vc.op[level], mg->rows, kOperator, true, kNoDot,
                                           static_cast<const double*>(v), v_stride, nullptr,
                                           0, nullptr, 0, static_cast<double*>(y), y_stride,
                                           0.0, nullptr, nullptr, s);
  }
  VCycle<float> vc(mg, s, nullptr);
  return stencil<float, float, float>(vc.op[level], mg->rows, kOperator, true, kNoDot,
                                      static_cast<const float*>(v), v_stride, nullptr, 0,
                                      nullptr, 0, static_cast<float*>(y), y_stride, 0.0,
                                      nullptr, nullptr, s);
}

int sttopt_mgpcg(const sttopt_mg* mg, const double* b, int64_t b_stride, double* x,
                 int64_t x_stride, double rtol, int32_t max_iter, int32_t flexible,
                 int32_t check_every, int32_t cold_start, double* work, int32_t* iters,
                 double* resid, int32_t* n_iter, void* stream) {
  if (int e = validate(mg)) return e;
  if (b == nullptr || x == nullptr || work == nullptr || iters == nullptr ||
      resid == nullptr || n_iter == nullptr || mg->scratch == nullptr ||
      mg->density64 == nullptr || max_iter < 0) {
    set_error("sttopt_mgpcg: null pointer or negative max_iter");
    return STTOPT_EINVAL;
  }
  if (mg->d == 2 && ((reinterpret_cast<uintptr_t>(b) | reinterpret_cast<uintptr_t>(x) |
                       reinterpret_cast<uintptr_t>(work)) % 16 != 0 ||
                      b_stride % 2 != 0 || x_stride % 2 != 0)) {
    set_error("sttopt_mgpcg: d = 2 vectors need node-aligned pointers and even row strides");
    return STTOPT_EINVAL;
  }
  const long long ndof0 =
      (long long)mg->d * (mg->levels[0].nelx + 1) * (long long)(mg->levels[0].nely + 1);
  if (mg->rows > 1 && x_stride < ndof0) {
    set_error("sttopt_mgpcg: x rows overlap (x_stride below the level-0 dof count)");
    return STTOPT_EINVAL;
  }
  if (mg->rows == 0) {
    *n_iter = 0;
    return STTOPT_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  if (mg->vdtype == STTOPT_F64)
    return mgpcg_impl<double>(mg, b, b_stride, x, x_stride, rtol, max_iter, flexible,
                              check_every, cold_start, work, iters, resid, n_iter, s);
  return mgpcg_impl<float>(mg, b, b_stride, x, x_stride, rtol, max_iter, flexible,
                           check_every, cold_start, work, iters, resid, n_iter, s);
}

int sttopt_mgpcg_batch(const sttopt_batch_group* groups, int32_t ngroups, double rtol,
                       int32_t max_iter, int32_t check_every, int32_t cold_start,
                       int32_t* n_iter, void* stream) {
  if (groups == nullptr || ngroups < 1 || n_iter == nullptr || max_iter < 0) {
    set_error("sttopt_mgpcg_batch: no groups, null n_iter or negative max_iter");
    return STTOPT_EINVAL;
  }
  for (int i = 0; i < ngroups; ++i) {
    const sttopt_batch_group& g = groups[i];
    if (int e = validate(g.mg)) return e;
    if (g.b == nullptr || g.x == nullptr || g.work == nullptr || g.iters == nullptr ||
        g.resid == nullptr || g.mg->scratch == nullptr || g.mg->density64 == nullptr ||
        g.mg->rows < 1 || g.mg->d != groups[0].mg->d) {
      set_error("sttopt_mgpcg_batch: a group with null vectors, no rows or other dofs per node");
      return STTOPT_EINVAL;
    }
    const long long ndof0 = (long long)g.mg->d * (g.mg->levels[0].nelx + 1) *
                            (long long)(g.mg->levels[0].nely + 1);
    if ((g.mg->d == 2 && ((reinterpret_cast<uintptr_t>(g.b) | reinterpret_cast<uintptr_t>(g.x) |
                           reinterpret_cast<uintptr_t>(g.work)) % 16 != 0 ||
                          g.b_stride % 2 != 0 || g.x_stride % 2 != 0)) ||
        (g.mg->rows > 1 && g.x_stride < ndof0)) {
      set_error("sttopt_mgpcg_batch: node-aligned vectors, even strides, non-overlapping x rows");
      return STTOPT_EINVAL;
    }
  }
  return mgpcg_batch_impl(groups, ngroups, rtol, max_iter, check_every, cold_start, n_iter,
                          (hipStream_t)stream);
}

}  // extern "C"
