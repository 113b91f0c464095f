// Internal launch interface shared by the translation units of the library.
#pragma once

#include "stt_common.h"

namespace stt {

// Values per node of the STTOPT_KEFF_NODE layout: the own block's upper triangle padded
// to 4, then the four forward d x d blocks (20 at d = 2, 8 at d = 1; records and blocks
// start 16-byte aligned in fp32).
__host__ __device__ constexpr int node_stencil_stride(int d) { return 4 + 4 * d * d; }

// N consecutive values starting 16-byte aligned (fp32: N a multiple of 4 or N = 1;
// fp64: N even or N = 1), with vector loads.
template <typename T, int N>
__device__ __forceinline__ void load_record(const T* p, T (&out)[N]) {
  constexpr int W = 16 / (int)sizeof(T);  // values per 16-byte load
  if constexpr (N % W == 0) {
    using V = HIP_vector_type<T, W>;
    const V* pv = reinterpret_cast<const V*>(p);
#pragma unroll
    for (int q = 0; q < N / W; ++q) {
      const V v = pv[q];
      if constexpr (W == 4) {
        out[4 * q] = v.x;
        out[4 * q + 1] = v.y;
        out[4 * q + 2] = v.z;
        out[4 * q + 3] = v.w;
      } else {
        out[2 * q] = v.x;
        out[2 * q + 1] = v.y;
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < N; ++q) out[q] = p[q];
  }
}

// kSmoothFirst (level 0 only): two damped-Jacobi sweeps from x = 0 in one pass, the
// first one (x0 = w D^-1 b) formed while staging the tile; also copies b into bcopy in
// the stencil precision when given.
// kOperatorDir: the PCG's direction update fused into the level-0 fp64 operator (the
// staged vector is p_new = z + beta p_old; see operator_dir).
// kSmoothProlong (the fp32 parameterized level 1 at 2 x 2): a damped-Jacobi sweep from
// x + P e, the coarse correction e prolongated while staging (see smooth_prolong).
enum StencilMode {
  kMatvec = 0,
  kOperator = 1,
  kSmooth = 2,
  kResid = 3,
  kSmoothFirst = 4,
  kOperatorDir = 5,
  kSmoothProlong = 6
};
enum DotKind { kNoDot = 0, kDotU = 1, kDotB = 2 };

// One level's operator, type-erased.  Exactly one of rho (level 0: density * KE)
// or keff (coarser levels: entry-major element matrices) is set.
struct OpDesc {
  int d = 2, nelx = 0, nely = 0;
  const void* rho = nullptr;
  long long sr = 0;
  const void* keff = nullptr;
  long long sk = 0;
  int keff_upper = 0;  // keff holds the packed upper triangle (STTOPT_KEFF_UPPER)
  int keff_node = 0;   // keff holds the assembled node stencil (STTOPT_KEFF_NODE)
  const void* rdiag = nullptr;
  long long srd = 0;
  const uint32_t* mask = nullptr;
  double ke[64] = {};
  // Density-parameterized coarse level: keff_E = sum over the pkx x pky fine
  // elements g under E of prho_g * ptab[g] (ptab: (pkx*pky) x (4d)^2 matrices), with
  // prho the level-0 density on its fnelx-wide grid.
  const void* ptab = nullptr;
  int pkx = 1, pky = 1, fnelx = 0;
  const void* prho = nullptr;
  long long spr = 0;

  int nnx() const { return nelx + 1; }
  int nny() const { return nely + 1; }
  long long nnodes() const { return (long long)(nelx + 1) * (nely + 1); }
  long long ndof() const { return nnodes() * d; }
  long long nel() const { return (long long)nelx * nely; }
  int ntx() const { return ceil_div(nnx(), kTileX); }
  int nty() const { return ceil_div(nny(), kTileY); }
  int tiles() const { return ntx() * nty(); }  // level-0 tiles = reduction partials per row
  int nty_small() const { return ceil_div(nny(), kSmallTileY); }
  int tiles_small() const { return ntx() * nty_small(); }
};

// y = mode(u): see StencilMode.  Types: T is the stencil (storage and compute)
// precision, TB the type of b, TY of y.  proj_in masks the input vector (only
// needed when it may carry values at fixed dofs).  dot != kNoDot writes one fp64
// partial sum per tile and row to partial[row * tiles + tile].
template <typename T, typename TB, typename TY>
int stencil(const OpDesc& op, int rows, int mode, bool proj_in, int dot, const T* u,
            long long su, const TB* b, long long sb, const T* invd, long long sd,
            TY* y, long long sy, double omega, const int* active, double* partial,
            hipStream_t s, T* bcopy = nullptr, long long sbc = 0);

// One PCG step's direction update fused into the level-0 fp64 apply (d = 1 or 2,
// density operator): for the rows with iters == k, x += alpha p_old at every node; for
// the active rows also p_new = z + beta p_old (z fp32 when z_f32, else fp64), Ap = K
// p_new at free dofs and the per-tile p_new . Ap partials.  p_old and p_new are
// different buffers (each tile reads p_old on its halo), all rows with stride n.
// start / stop (optional): events that the launch itself records at the kernel's start
// and end (no marker of their own in the stream).  The x updates of two iterations can
// share one pass over x: with defer, the rows still iterating hold iteration k's update
// back (the rows that retired in k take it); the next launch passes p_prev = that p_old
// and alpha_prev = its alpha, and takes x += alpha_prev p_prev, then x += alpha p_old,
// as two fused multiply-adds, the same values as one update per launch.
int operator_dir(const OpDesc& op, int rows, const double* p_old, double* p_new,
                 const void* z, int z_f32, double* x, long long xs, const double* alpha,
                 const double* beta, const int* iters, int k, double* Ap,
                 const int* active, double* partial, hipStream_t s,
                 hipEvent_t start = nullptr, hipEvent_t stop = nullptr,
                 const double* p_prev = nullptr, const double* alpha_prev = nullptr,
                 int defer = 0);

// One damped-Jacobi sweep y = xp + w D^-1 (b - A xp) from xp = x + P e (P: the bilinear
// 2 x 2 prolongation of the coarse correction e, added at the free dofs, the
// arithmetic of prolong_add), on the fp32 level applied from the level-0 densities at
// 2 x 2 (op.ptab with pkx = pky = 2, float).  cnnx x cnny: e's grid, which must be this
// level's at 2 x 2.  x is read, not updated; y must not overlap x.
int smooth_prolong(const OpDesc& op, int rows, const float* x, long long sx, const float* e,
                   long long se, int cnnx, int cnny, const float* b, long long sb,
                   const float* invd, long long sd, float* y, long long sy, double omega,
                   const int* active, hipStream_t s);

// Coarsest solve x = cinv b, output dofs bx * (waves per workgroup) + wave of one row:
// one wavefront per output dof, the 64 lanes stream one row of the dense inverse
// (coalesced) and reduce in fp64.
template <typename T, bool PAIRS>
__device__ __forceinline__ void coarse_solve_body(long long n, const double* cinv,
                                                  long long scinv, const T* b, long long sb,
                                                  const uint32_t* mask, T* x, long long sx,
                                                  const int* active, const int bx,
                                                  const int row) {
  if (active != nullptr && active[row] == 0) return;
  const long long i = (long long)bx * (blockDim.x / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (i >= n) return;  // uniform per wave
  const double* Ai = cinv + row * scinv + i * n;
  const T* bb = b + row * sb;
  double acc = 0.0;
  if (PAIRS) {
    // 16-byte loads (n even and 16-byte aligned rows), several in flight per lane
    const long long n2 = n >> 1;
    const double2* A2 = reinterpret_cast<const double2*>(Ai);
#pragma unroll 8
    for (long long j = lane; j < n2; j += kWave) {
      const double2 av = A2[j];
      acc += av.x * (double)bb[2 * j] + av.y * (double)bb[2 * j + 1];
    }
  } else {
#pragma unroll 8
    for (long long j = lane; j < n; j += kWave) acc += Ai[j] * (double)bb[j];
  }
  acc = wave_sum(acc);
  if (lane == 0) x[row * sx + i] = mask_free(mask, i) ? static_cast<T>(acc) : T(0);
}

// ---- one launch over the same level of several hierarchies (the uneven batch) ----
// Each kernel below runs the same pass on up to kMaxBatchGroups hierarchies of different
// grids in one launch: hierarchy i takes a contiguous range of workgroups, with its own
// arguments; the arithmetic is that of the single-hierarchy pass.
constexpr int kMaxBatchGroups = 4;

// A tile stencil pass (level 0 with density * KE, or the parameterized 2 x 2 level 1 in
// fp32): the arguments of `stencil` with TB = TY = T and no input projection.
template <typename T>
struct StencilCall {
  OpDesc op;
  int rows, mode, dot;
  const T* u;
  long long su;
  const T* b;
  long long sb;
  const T* invd;
  long long sd;
  T* y;
  long long sy;
  double omega;
  const int* active;
  double* partial;
};
template <typename T>
int stencil_tile_groups(const StencilCall<T>* calls, int n, hipStream_t s);

// Node-centric diagonal of an operator (+ rdiag).  out_mode: 0 raw diagonal,
// 1 diagonal with 1 at fixed dofs, 2 inverse at free dofs and 0 at fixed.
template <typename TO>
int diagonal(const OpDesc& op, int rows, int out_mode, TO* out, long long so,
             hipStream_t s);

// Elementwise kernels over [rows][n] (n = ndof of a level).
template <typename T, typename TB>
int jacobi_first(long long n, int rows, const TB* b, long long sb, const T* invd,
                 long long sd, double omega, const uint32_t* mask, T* x, long long sx,
                 T* bcopy, long long sbc, const int* active, hipStream_t s);
template <typename TI, typename TO>
int convert(long long n, int rows, const TI* in, long long si, TO* out, long long so,
            const int* active, hipStream_t s);
template <typename T>
int fill(long long n, int rows, T value, T* x, long long sx, const int* active,
         hipStream_t s);
template <typename T>
int axpy(long long n, int rows, const T* xin, long long si, T* y, long long sy,
         const int* active, hipStream_t s);  // y += x
int project_f64(long long n, int rows, const uint32_t* mask, double* x, long long sx,
                hipStream_t s);

// Grid transfers between a level with cnelx x cnely elements and its fine level.
template <typename T>
int restrict_to(int d, int cnelx, int cnely, int kx, int ky, int rows, const T* fine,
                long long sf, const uint32_t* cmask, T* coarse, long long sc,
                const int* active, hipStream_t s);
template <typename T>
int prolong_add(int d, int cnelx, int cnely, int kx, int ky, int rows, const T* coarse,
                long long sc, const uint32_t* fmask, T* fine, long long sf,
                const int* active, hipStream_t s);

// Dense coarsest solve x = P (cinv b).
template <typename T>
int coarse_solve(long long n, int rows, const double* cinv, long long scinv,
                 const T* b, long long sb, const uint32_t* mask, T* x, long long sx,
                 const int* active, hipStream_t s);

// Level-0 smoothing of the fp32 V-cycle with two sweeps per pass (n_smooth = 2), on
// 62 x 14 node output tiles (pair_tiles per row).  pre: x = two sweeps from zero on
// b (fp64, rounded to fp32 as the cycle's copy of b would be) and r = b - A x.  post:
// z = two sweeps from x + P ec (ec on the 2 x 2 coarser grid with cnnx nodes per row);
// partial, when given, receives the per-tile r.z sums (free dofs).  Both need the
// level-0 free-dof mask and rows below 4 GiB (32-bit offsets).  pre with ap != nullptr
// first takes the PCG update: its b is b - alpha[row] ap, written to rnew (not b: the
// tiles read b's halo) with the per-tile sums of its squares in rr and, with zold,
// of its products with zold in rzo.
int pair_tiles(const OpDesc& op);
int smooth_pair_pre(const OpDesc& op, int rows, const double* b, long long sb,
                    const float* invd, long long sd, float* x, long long sx, float* r,
                    long long sr, double omega, const int* active, hipStream_t s,
                    const double* ap = nullptr, long long sap = 0,
                    const double* alpha = nullptr, double* rnew = nullptr, long long srn = 0,
                    double* rr = nullptr, const float* zold = nullptr, long long szo = 0,
                    double* rzo = nullptr);
template <typename ZT>
int smooth_pair_post(const OpDesc& op, int rows, const float* x, long long sx,
                     const float* ec, long long sec, int cnnx, const double* b, long long sb,
                     const float* invd, long long sd, ZT* z, long long sz, double omega,
                     const int* active, double* partial, hipStream_t s);

// Coarse levels of the fp32 V-cycle (node-stencil operator, 2 x 2 coarsening to the
// next level, two sweeps) in two passes.  pre: x2 = two sweeps from zero on b (written
// to x2) and the restricted residual of x2 on the coarser level (written to bc, masked
// by cmask).  post: x4 = two sweeps from x2 + P ec (written to x4, which must not alias
// x2).  Same results as the separate first sweeps, residual, restriction, prolongation
// and sweeps.  variant: the tile schedule (0: 32 x 4 coarse / 64 x 8 fine nodes per
// 512-thread workgroup; 1: 16 x 4 / 32 x 8 per 256 threads).
int coarse_pair_pre(const OpDesc& op, int rows, const float* b, long long sb,
                    const float* invd, long long sd, float* x2, long long sx2,
                    const uint32_t* cmask, float* bc, long long sbc, double omega,
                    const int* active, hipStream_t s, int variant = 0);
int coarse_pair_post(const OpDesc& op, int rows, const float* x2, long long sx2,
                     const float* ec, long long sec, const float* b, long long sb,
                     const float* invd, long long sd, float* x4, long long sx4, double omega,
                     const int* active, hipStream_t s, int variant = 0);

// The same passes over one level of each of up to 8 hierarchies in one launch (the
// uneven batch): lv[i] describes hierarchy i's level, its rows and vectors (pre: b,
// invd, x2 out, cmask, bc out; post: x2, ec, b, invd, x4 out).
struct PairLevel {
  OpDesc op;
  int rows;
  const float* b;
  long long sb;
  const float* invd;
  long long sd;
  const float* x2;
  long long sx2;
  const uint32_t* cmask;
  float* bc;
  long long sbc;
  const float* ec;
  long long sec;
  float* x4;
  long long sx4;
  double omega;
  const int* active;
};
int coarse_pair_groups(const PairLevel* lv, int n, bool pre, int variant, hipStream_t s);

// The bottom of the fp32 V-cycle in one persistent launch: the two-pass node-stencil
// levels lv[0 .. nlev - 1] (finest first, each 2 x 2 over the next; lv[i].bc is
// lv[i + 1].b and the coarsest level's right-hand side cb), the coarsest solve
// cx = cinv cb, and the post passes back up (lv[i].ec is lv[i + 1].x4, or cx).  The
// passes' tiles form one work list, phase by phase (pre passes, coarsest solve, post
// passes) and row by row; workgroups take items in order, and an item waits (bounded)
// until every item of the previous phase of its row is done, so no launch boundary
// separates the phases.  The arithmetic is that of the separate passes (the same
// bodies), so results are bit-identical.  sync: bottom_sync_words(rows) words that are
// zero before the launch (the launch leaves them at zero); *err is set if a wait gave
// up (the results are then not valid).
constexpr int kMaxBottomLevels = 3;
struct BottomCall {
  PairLevel lv[kMaxBottomLevels];
  int nlev;
  long long n;  // coarsest level dofs (even, 16-byte aligned rows)
  const double* cinv;
  long long scinv;
  const float* cb;
  long long scb;
  const uint32_t* cmask;
  float* cx;
  long long scx;
};
long long bottom_sync_words(int rows);
int vcycle_bottom(const BottomCall& c, unsigned* sync, int* err, hipStream_t s);

// The level-0 passes of several hierarchies in one launch (arguments of smooth_pair_pre
// and smooth_pair_post; all calls of a launch with the same dofs per node, reaction
// diagonal presence and, for pre, residual update presence).
struct PairPreCall {
  OpDesc op;
  int rows;
  const double* b;
  long long sb;
  const float* invd;
  long long sd;
  float* x;
  long long sx;
  float* r;
  long long sr;
  double omega;
  const int* active;
  const double* ap;
  long long sap;
  const double* alpha;
  double* rnew;
  long long srn;
  double* rr;
  const float* zold;
  long long szo;
  double* rzo;
};
template <typename ZT>
struct PairPostCall {
  OpDesc op;
  int rows;
  const float* x;
  long long sx;
  const float* ec;
  long long sec;
  int cnnx;
  const double* b;
  long long sb;
  const float* invd;
  long long sd;
  ZT* z;
  long long sz;
  double omega;
  const int* active;
  double* partial;
};
int smooth_pair_pre_groups(const PairPreCall* calls, int n, hipStream_t s);
int smooth_pair_post_groups(const PairPostCall<float>* calls, int n, hipStream_t s);

// 2 x 2 grid transfers of several hierarchies in one launch (fp32): restriction
// coarse = P_c (R src) with mask the coarse free dofs, or prolongation dst += P_f (P src)
// with mask the fine free dofs; cnelx x cnely is the coarse grid.
struct TransferCall {
  int d, cnelx, cnely, rows;
  const float* src;
  long long ss;
  const uint32_t* mask;
  float* dst;
  long long sd;
  const int* active;
};
int transfer22_groups(const TransferCall* calls, int n, bool restrict_, hipStream_t s);

// The coarsest dense solves x = P (cinv b) of several hierarchies in one launch (fp32
// vectors, even n).
struct CoarseSolveCall {
  long long n;
  int rows;
  const double* cinv;
  long long scinv;
  const float* b;
  long long sb;
  const uint32_t* mask;
  float* x;
  long long sx;
  const int* active;
};
int coarse_solve_groups(const CoarseSolveCall* calls, int n, hipStream_t s);

// Copy one int from device memory on stream s and wait for it (pinned readback word).
int read_device_int(const int* dev, hipStream_t s, int* out);

}  // namespace stt
