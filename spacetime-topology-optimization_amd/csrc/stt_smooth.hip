// Level-0 passes of the fp32 V-cycle that evaluate the stencil twice per tile
// (temporal blocking), for two damped-Jacobi pre- and post-smoothing sweeps
// (torch_mg.py:440-455 with n_smooth = 2):
//
//   pre:  x1 = w D^-1 b,  x2 = x1 + w D^-1 (b - A x1),  r = b - A x2
//   post: xp = x2 + P e,  x3 = xp + w D^-1 (b - A xp),  z = x3 + w D^-1 (b - A x3)
//
// A workgroup owns a 62 x 14 node output tile.  The first stencil runs on the
// surrounding 64 x 16 nodes (the output tile plus a one-node ring, one wavefront
// column per lane, 4 node rows per thread as in the single-pass tile kernel), its
// result goes to LDS, and the second stencil reads it from there.  So x1, xp and
// x3 never reach HBM: the pre pass reads b, D^-1 and the densities and writes x2
// and r; the post pass reads x2, the coarse correction e, b, D^-1 and the
// densities and writes z.  The arithmetic of each sweep is the one of the
// single-pass kernels (same operands, same order), so results are identical.
#include <cstddef>

#include "stt_kernels.h"

namespace stt {

namespace {

constexpr int kOutX = kTileX - 2, kOutY = kTileY - 2;  // output tile
constexpr int kHX = kTileX + 2, kHY = kTileY + 2;      // staged halo of the first stencil
constexpr int kEX = kTileX + 1, kEY = kTileY + 1;      // element densities

template <typename ZT>
struct PairArgs {
  int nelx, nely, nnx, nny, ntx, tiles, total;
  int cnnx;               // post: nodes per row of the coarse grid (2 x 2 coarsening)
  const float* x;         // post: x2, the result of the pre-smoothing
  long long sx;
  const float* ec;        // post: coarse-grid correction e
  long long sec;
  const double* b;        // right-hand side (the fp64 PCG residual)
  long long sb;
  const float* invd;
  long long sd;
  const float* rho;
  long long sr;
  const float* rdiag;
  long long srd;
  const uint32_t* mask;
  ZT* out;                // pre: x2; post: z
  long long so;
  float* res;             // pre: r = b - A x2
  long long sres;
  float omega;
  const int* active;
  double* partial;        // post: r.z per tile (may be null)
  // pre with the PCG residual update: the staged b is r - alpha Ap (r = b above, Ap
  // fp64, alpha per row); the output nodes' fp64 values go to rnew, the tile sum of
  // their squares to rr and, when zold is given (the flexible beta), the tile sum of
  // their products with zold to rzo
  const double* ap;
  long long sap;
  const double* alpha;
  double* rnew;
  long long srn;
  double* rr;
  const float* zold;
  long long szo;
  double* rzo;
  KE<float> ke;
};

using KePtr = const __attribute__((address_space(4))) float*;

// acc[j] = sum over the 4 elements around node row j of rho_e (KE u_e) for the NR node
// rows of one thread: W is the thread's (NR + 2) x 3 node window, R the (NR + 1) x 2
// element densities around it.  Element position outer, node row inner.
template <int D, int NR>
__device__ __forceinline__ void apply_window(const Vec<float, D> (&W)[NR + 2][3],
                                             const float (&R)[NR + 1][2], KePtr ke_args,
                                             float (&acc)[NR][D]) {
  constexpr int K = 4 * D;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
#pragma unroll
    for (int ii = 0; ii < D; ++ii) acc[j][ii] = 0.0f;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    constexpr int locs[4] = {1, 0, 2, 3};
    const int loc = locs[q];
    KePtr kq = ke_args + loc * D * K;
    float ke[D][K];
#pragma unroll
    for (int ii = 0; ii < D; ++ii) {
#pragma unroll
      for (int c = 0; c < K; ++c) ke[ii][c] = kq[ii * K + c];
    }
    const int qr = q >> 1, qc = q & 1;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      float part[D];
#pragma unroll
      for (int ii = 0; ii < D; ++ii) part[ii] = 0.0f;
#pragma unroll
      for (int jn = 0; jn < 4; ++jn) {
        const Vec<float, D>& un = W[j + qr + loc_dr(jn)][qc + loc_dc(jn)];
#pragma unroll
        for (int ii = 0; ii < D; ++ii) {
#pragma unroll
          for (int kk = 0; kk < D; ++kk) part[ii] += ke[ii][jn * D + kk] * un.v[kk];
        }
      }
      const float dens = R[j + qr][qc];
#pragma unroll
      for (int ii = 0; ii < D; ++ii) acc[j][ii] += dens * part[ii];
    }
  }
}

// Node access with a 32-bit byte offset from a wave-uniform row base (one row of a
// batched vector stays below 4 GiB), so that loads take the scalar-base form.
template <typename T, int D>
__device__ __forceinline__ Vec<T, D> ld(const T* base, uint32_t node) {
  return load_node<T, D>(reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                                    node * (uint32_t)(D * sizeof(T))),
                         0);
}
template <typename T, int D>
__device__ __forceinline__ void st(T* base, uint32_t node, const Vec<T, D>& v) {
  store_node<T, D>(reinterpret_cast<T*>(reinterpret_cast<char*>(base) +
                                        node * (uint32_t)(D * sizeof(T))),
                   0, v);
}

template <typename T, int D>
__device__ __forceinline__ Vec<T, D> zvec() {
  Vec<T, D> z;
#pragma unroll
  for (int k = 0; k < D; ++k) z.v[k] = T(0);
  return z;
}

// RD: the operator carries a lumped reaction diagonal.  UPD (pre only): the pass also
// takes the PCG's residual update r <- r - alpha Ap.
template <typename ZT, int D, bool POST, bool RD, bool UPD>
__device__ __forceinline__ void smooth_pair_body(const PairArgs<ZT>& a, const int bid,
                                                 const int karg_off) {
  static_assert(!(POST && UPD), "the residual update belongs to the pre pass");
  constexpr int NR = kRowsPerThread;
  // input of the first stencil, and its output with a one-node border (sX).  In the
  // pre pass sX reuses sU's LDS once every thread holds its window in registers (the
  // LDS of five workgroups per CU); the post pass, limited by its registers, keeps
  // its own.
  __shared__ Vec<float, D> sU[kHY][kHX];
  __shared__ Vec<float, D> sXp[POST ? kTileY + 2 : 1][POST ? kTileX + 2 : 1];
  static_assert(kHY == kTileY + 2 && kHX == kTileX + 2, "sX may alias sU");
  Vec<float, D>(*const sX)[kTileX + 2] =
      POST ? reinterpret_cast<Vec<float, D>(*)[kTileX + 2]>(&sXp[0][0]) : sU;
  __shared__ float sR[kEY][kEX];
  // pre: b and D^-1 of the first stencil's nodes, from the staging loads
  __shared__ Vec<float, D> sB[POST ? 1 : kTileY][POST ? 1 : kTileX];
  __shared__ Vec<float, D> sI[POST ? 1 : kTileY][POST ? 1 : kTileX];
  __shared__ double red[kWaves];
  const int t = xcd_remap(bid, a.total);
  int row, tt;
  row_tile(t, a.tiles, a.total, row, tt);
  if (a.active != nullptr && a.active[row] == 0) return;  // whole block: uniform
  const int by = tt / a.ntx, bx = tt - by * a.ntx;
  // origin of the first stencil's 64 x 16 nodes; the staged halo starts one node
  // further up and left
  const int x0 = bx * kOutX - 1, y0 = by * kOutY - 1;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int nnx = a.nnx, nny = a.nny;
  const uint32_t* __restrict__ mask = a.mask;
  const double* __restrict__ b = a.b + row * a.sb;
  const float* __restrict__ invd = a.invd + row * a.sd;
  const float* __restrict__ rho = a.rho + row * a.sr;
  const float* __restrict__ xin = POST ? a.x + row * a.sx : nullptr;
  const float* __restrict__ ec = POST ? a.ec + row * a.sec : nullptr;
  const float* __restrict__ rdg = RD ? a.rdiag + row * a.srd : nullptr;
  const double* __restrict__ apr = UPD ? a.ap + row * a.sap : nullptr;
  const double alpha = UPD ? a.alpha[row] : 0.0;
  const bool flex = UPD && a.zold != nullptr;  // uniform
  const float* __restrict__ zor = flex ? a.zold + row * a.szo : nullptr;

  // ---- staging: wavefront w takes halo rows w, w + 4, ...; lane l column x0 - 1 + l.
  // The two trailing node columns and the trailing element column are taken by the
  // block's first 2 * kHY + kEY threads.  Every staging load is issued first, from rows
  // and columns clamped into the grid (so none is conditional); the staged values are
  // then formed, zeroed outside the grid, and written to LDS.
  constexpr int kUR = (kHY + kWaves - 1) / kWaves;
  constexpr int kRR = (kEY + kWaves - 1) / kWaves;
  constexpr int kTailU = 2 * kHY, kTail = kTailU + kEY;
  const int cx = x0 - 1 + lane;
  const bool cnode = cx >= 0 && cx < nnx;
  const bool celem = cx >= 0 && cx < a.nelx;
  const uint32_t cxn = (uint32_t)min(max(cx, 0), nnx - 1);
  const uint32_t cxe = (uint32_t)min(max(cx, 0), a.nelx - 1);
  const int tix = threadIdx.x;
  const int ti = tix < kTailU ? (tix >> 1) : tix - kTailU;
  const int tk = tix < kTailU ? kTileX + (tix & 1) : kTileX;
  const int tgy = y0 - 1 + ti, tgx = x0 - 1 + tk;
  auto clampy = [&](int gy) -> uint32_t { return (uint32_t)min(max(gy, 0), nny - 1); };

  // post: the coarse correction at the tile's nodes of the 2 x 2 coarser grid (at most
  // kCY x kCX of them from (cy0, cx0), the coarse node at or before the halo's first
  // fine node), staged in LDS
  constexpr int kCY = kHY / 2 + 1, kCX = kHX / 2 + 1;
  __shared__ Vec<float, D> sE[POST ? kCY : 1][POST ? kCX : 1];
  const int cy0 = (y0 + 1) / 2 - 1, cx0 = (x0 + 1) / 2 - 1;  // floor((y0 - 1) / 2), y0 >= -1
  constexpr int kCE = (kCY * kCX + kBlock - 1) / kBlock;
  Vec<float, D> cev[POST ? kCE : 1];
  if constexpr (POST) {
    const int cnny = a.nely / 2 + 1;
#pragma unroll
    for (int m = 0; m < kCE; ++m) {
      const int e = tix + m * kBlock, cy = cy0 + e / kCX, cx = cx0 + e % kCX;
      cev[m] = zvec<float, D>();
      if (e < kCY * kCX && cy >= 0 && cy < cnny && cx >= 0 && cx < a.cnnx)
        cev[m] = ld<float, D>(ec, (uint32_t)cy * (uint32_t)a.cnnx + (uint32_t)cx);
    }
  }

  // loads of one staged node (gy, gx), both inside the grid
  struct Raw {
    Vec<float, D> x;  // post: x2
    uint32_t word;
    Vec<double, D> b;                      // pre: b and D^-1 (and Ap with UPD)
    Vec<float, D> dv;
    Vec<double, D> ap;
    Vec<float, D> zo;
  };
  auto load_raw = [&](uint32_t gy, uint32_t gx) -> Raw {
    Raw r;
    const uint32_t node = gy * (uint32_t)nnx + gx;
    if constexpr (POST) {
      r.x = ld<float, D>(xin, node);
      r.word = mask[(node * D) >> 5];
    } else {
      r.b = ld<double, D>(b, node);
      r.dv = ld<float, D>(invd, node);
      if constexpr (UPD) {
        r.ap = ld<double, D>(apr, node);
        r.zo = flex ? ld<float, D>(zor, node) : zvec<float, D>();
      }
    }
    return r;
  };
  // staged value (ok: the node is inside the grid, else zero): pre x1 = w D^-1 b, with
  // b and D^-1 also returned; post xp = x2 + P e at free dofs (bilinear weights 1, 0 at
  // even and 1/2, 1/2 at odd fine positions)
  struct Staged {
    Vec<float, D> u, bc, dv;
    Vec<double, D> b64;  // pre with UPD: the updated fp64 b
    Vec<float, D> zo;    // and zold
  };
  auto finish = [&](const Raw& r, uint32_t gy, uint32_t gx, bool ok) -> Staged {
    Staged s;
    if constexpr (POST) {
      const uint32_t node = gy * (uint32_t)nnx + gx;
      const uint32_t bits = (r.word >> ((node * D) & 31)) & ((1u << D) - 1u);
      const float wy1 = (gy & 1u) ? 0.5f : 0.0f, wy0 = 1.0f - wy1;
      const float wx1 = (gx & 1u) ? 0.5f : 0.0f, wx0 = 1.0f - wx1;
      // the 2 x 2 coarse nodes around; at an even fine position the second one of a
      // pair is the first one again (its weight is 0)
      const int ly = (int)(gy >> 1) - cy0, lx = (int)(gx >> 1) - cx0;
      const int ly1 = ly + (int)(gy & 1u), lx1 = lx + (int)(gx & 1u);
      const Vec<float, D> c00 = sE[ly][lx], c01 = sE[ly][lx1];
      const Vec<float, D> c10 = sE[ly1][lx], c11 = sE[ly1][lx1];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const float top = wx0 * c00.v[k] + wx1 * c01.v[k];
        const float bot = wx0 * c10.v[k] + wx1 * c11.v[k];
        float u = r.x.v[k];
        if ((bits >> k) & 1u) u += wy0 * top + wy1 * bot;
        s.u.v[k] = ok ? u : 0.0f;
      }
    } else {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        // the update of pcg_update_kernel, same expression
        const double bk = UPD ? fma(-alpha, r.ap.v[k], r.b.v[k]) : r.b.v[k];
        s.b64.v[k] = bk;
        s.zo.v[k] = r.zo.v[k];
        s.bc.v[k] = ok ? static_cast<float>(bk) : 0.0f;
        s.dv.v[k] = ok ? r.dv.v[k] : 0.0f;
        s.u.v[k] = a.omega * s.dv.v[k] * s.bc.v[k];
      }
    }
    return s;
  };
  Raw rw[kUR];
#pragma unroll
  for (int n = 0; n < kUR; ++n) rw[n] = load_raw(clampy(y0 - 1 + wave + n * kWaves), cxn);
  const uint32_t tgyc = clampy(tgy), tgxc = (uint32_t)min(tgx, nnx - 1);
  Raw rt;
  if (tix < kTailU) rt = load_raw(tgyc, tgxc);
  float wr[kRR];
#pragma unroll
  for (int n = 0; n < kRR; ++n) {
    const int ey = y0 - 1 + wave + n * kWaves;
    wr[n] = rho[(uint32_t)min(max(ey, 0), a.nely - 1) * (uint32_t)a.nelx + cxe];
  }
  float tr = 0.0f;
  if (tix >= kTailU && tix < kTail)
    tr = rho[(uint32_t)min(max(tgy, 0), a.nely - 1) * (uint32_t)a.nelx +
             (uint32_t)min(tgx, a.nelx - 1)];

  if constexpr (POST) {
#pragma unroll
    for (int m = 0; m < kCE; ++m) {
      const int e = tix + m * kBlock;
      if (e < kCY * kCX) sE[e / kCX][e % kCX] = cev[m];
    }
    __syncthreads();  // the prolongation below reads them
  }

  // ---- staged values to LDS
  // UPD: squares of the updated residual at this thread's output nodes, and its
  // products with zold
  double rr = 0.0, rzo = 0.0;
#pragma unroll
  for (int n = 0; n < kUR; ++n) {
    const int i = wave + n * kWaves, gy = y0 - 1 + i;
    if (i < kHY) {
      const bool ok = gy >= 0 && gy < nny && cnode;
      const Staged sg = finish(rw[n], clampy(gy), cxn, ok);
      sU[i][lane] = sg.u;
      if constexpr (!POST) {
        // halo node (i, lane) is node (i - 1, lane - 1) of the first stencil
        if (i >= 1 && i <= kTileY && lane >= 1) {
          sB[i - 1][lane - 1] = sg.bc;
          sI[i - 1][lane - 1] = sg.dv;
        }
      }
      if constexpr (UPD) {
        // output nodes (halo rows 2 .. kOutY + 1, columns 2 .. kOutX + 1): each is
        // written by exactly one tile
        if (ok && i >= 2 && i <= kOutY + 1 && lane >= 2 && lane <= kOutX + 1) {
          st<double, D>(a.rnew + row * a.srn, (uint32_t)gy * (uint32_t)nnx + cxn, sg.b64);
#pragma unroll
          for (int k = 0; k < D; ++k) {
            rr += sg.b64.v[k] * sg.b64.v[k];
            rzo += sg.b64.v[k] * (double)sg.zo.v[k];
          }
        }
      }
    }
  }
  if (tix < kTailU) {
    const Staged sg = finish(rt, tgyc, tgxc, tgy >= 0 && tgy < nny && tgx < nnx);
    sU[ti][tk] = sg.u;
    if constexpr (!POST) {
      // halo column 64 is the first stencil's last node column (63); 65 is outside it
      if (tk == kTileX && ti >= 1 && ti <= kTileY) {
        sB[ti - 1][kTileX - 1] = sg.bc;
        sI[ti - 1][kTileX - 1] = sg.dv;
      }
    }
  } else if (tix < kTail) {
    sR[ti][tk] = (tgy >= 0 && tgy < a.nely && tgx < a.nelx) ? tr : 0.0f;
  }
#pragma unroll
  for (int n = 0; n < kRR; ++n) {
    const int i = wave + n * kWaves, ey = y0 - 1 + i;
    if (i < kEY) sR[i][lane] = (ey >= 0 && ey < a.nely && celem) ? wr[n] : 0.0f;
  }

  // operands of this thread's first-stencil nodes (y0 + 4 wave + j, x0 + lane), loaded
  // now for the epilogues: mask words (0 outside the grid: those nodes stay zero),
  // reaction diagonal, and for post b and D^-1 (pre has them in LDS)
  const int gxo = x0 + lane, gyo = y0 + wave * NR;
  const bool oin = gxo >= 0 && gxo < nnx;
  const uint32_t gxc = (uint32_t)min(max(gxo, 0), nnx - 1);
  uint32_t ow[NR];
  Vec<double, D> ob[POST ? NR : 1];
  Vec<float, D> oi[POST ? NR : 1];
  Vec<float, D> ord[RD ? NR : 1];
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const int gy = gyo + j;
    const uint32_t node = clampy(gy) * (uint32_t)nnx + gxc;
    const uint32_t w = mask[(node * D) >> 5];
    ow[j] = (oin && gy >= 0 && gy < nny) ? w : 0u;
    if constexpr (RD) ord[j] = ld<float, D>(rdg, node);
    if constexpr (POST) {
      ob[j] = ld<double, D>(b, node);
      oi[j] = ld<float, D>(invd, node);
    }
  }

  __syncthreads();

  const int i0 = wave * NR;
  Vec<float, D> bc[NR], iv[NR];  // b (in fp32) and D^-1 of the first stencil's nodes
  if constexpr (!POST) {
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      bc[j] = sB[i0 + j][lane];
      iv[j] = sI[i0 + j][lane];
    }
  } else {
#pragma unroll
    for (int j = 0; j < NR; ++j) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        bc[j].v[k] = static_cast<float>(ob[j].v[k]);
        iv[j].v[k] = oi[j].v[k];
      }
    }
  }

  using Args = PairArgs<ZT>;
  const KePtr ke_args = reinterpret_cast<KePtr>(
      (__attribute__((address_space(4))) char*)__builtin_amdgcn_kernarg_segment_ptr() +
      karg_off + offsetof(Args, ke));  // karg_off: where `a` sits in the argument segment
  float R[NR + 1][2];
#pragma unroll
  for (int r = 0; r < NR + 1; ++r) {
    R[r][0] = sR[i0 + r][lane];
    R[r][1] = sR[i0 + r][lane + 1];
  }
  // free-dof bit of dof ii of node row j
  auto freebit = [&](int j, int ii) -> bool {
    const uint32_t node = (uint32_t)(gyo + j) * (uint32_t)nnx + (uint32_t)gxo;
    return (ow[j] >> ((node * D + ii) & 31)) & 1u;
  };

  // ---- first stencil on the 64 x 16 nodes: one damped-Jacobi sweep
  {
    Vec<float, D> W[NR + 2][3];
#pragma unroll
    for (int r = 0; r < NR + 2; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) W[r][c] = sU[i0 + r][lane + c];
    }
    if constexpr (!POST) __syncthreads();  // every window is read: sU becomes sX
    // border of sX: read only by nodes without output; kept at zero
    if (tix < kTileX + 2) {
      sX[0][tix] = zvec<float, D>();
      sX[kTileY + 1][tix] = zvec<float, D>();
    } else if (tix < kTileX + 2 + kTileY) {
      sX[tix - (kTileX + 2) + 1][0] = zvec<float, D>();
      sX[tix - (kTileX + 2) + 1][kTileX + 1] = zvec<float, D>();
    }
    float acc[NR][D];
    apply_window<D, NR>(W, R, ke_args, acc);
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const Vec<float, D>& uc = W[j + 1][1];
      if constexpr (RD) {
#pragma unroll
        for (int ii = 0; ii < D; ++ii) acc[j][ii] += ord[j].v[ii] * uc.v[ii];
      }
      Vec<float, D> o;
#pragma unroll
      for (int ii = 0; ii < D; ++ii)
        o.v[ii] = freebit(j, ii) ? uc.v[ii] + a.omega * iv[j].v[ii] * (bc[j].v[ii] - acc[j][ii])
                                 : 0.0f;
      sX[i0 + j + 1][lane + 1] = o;
      if constexpr (!POST) {
        // x2 of the output nodes
        const int oy = i0 + j;
        if (oy >= 1 && oy <= kOutY && lane >= 1 && lane <= kOutX && gyo + j < nny && gxo < nnx)
          st<float, D>(reinterpret_cast<float*>(a.out) + row * a.so,
                       (uint32_t)(gyo + j) * (uint32_t)nnx + (uint32_t)gxo, o);
      }
    }
  }
  __syncthreads();

  // ---- second stencil on the output nodes
  double dot = 0.0;
  {
    Vec<float, D> W[NR + 2][3];
#pragma unroll
    for (int r = 0; r < NR + 2; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) W[r][c] = sX[i0 + r][lane + c];
    }
    float acc[NR][D];
    apply_window<D, NR>(W, R, ke_args, acc);
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const int oy = i0 + j;
      const int gy = gyo + j;
      if (!(oy >= 1 && oy <= kOutY && lane >= 1 && lane <= kOutX && gy < nny && gxo < nnx))
        continue;
      const uint32_t node = (uint32_t)gy * (uint32_t)nnx + (uint32_t)gxo;
      const Vec<float, D>& uc = W[j + 1][1];
      if constexpr (RD) {
#pragma unroll
        for (int ii = 0; ii < D; ++ii) acc[j][ii] += ord[j].v[ii] * uc.v[ii];
      }
      if constexpr (POST) {
        Vec<ZT, D> zo;
#pragma unroll
        for (int ii = 0; ii < D; ++ii) {
          const float o = freebit(j, ii)
                              ? uc.v[ii] + a.omega * iv[j].v[ii] * (bc[j].v[ii] - acc[j][ii])
                              : 0.0f;
          zo.v[ii] = static_cast<ZT>(o);
          dot += ob[j].v[ii] * (double)zo.v[ii];
        }
        st<ZT, D>(a.out + row * a.so, node, zo);
      } else {
        Vec<float, D> ro;
#pragma unroll
        for (int ii = 0; ii < D; ++ii) ro.v[ii] = freebit(j, ii) ? bc[j].v[ii] - acc[j][ii] : 0.0f;
        st<float, D>(a.res + row * a.sres, node, ro);
      }
    }
  }
  if constexpr (POST) {
    if (a.partial != nullptr) {
      const double sum = block_sum(dot, red);
      if (threadIdx.x == 0) a.partial[(long long)row * a.tiles + tt] = sum;
    }
  }
  if constexpr (UPD) {
    const double sum = block_sum(rr, red);
    if (threadIdx.x == 0) a.rr[(long long)row * a.tiles + tt] = sum;
    if (flex) {
      __syncthreads();
      const double sz = block_sum(rzo, red);
      if (threadIdx.x == 0) a.rzo[(long long)row * a.tiles + tt] = sz;
    }
  }
}

// The pre pass runs at five waves per SIMD (at most 96 VGPRs; its LDS allows five
// workgroups per CU); the post pass keeps four, where five would spill.
template <int D, bool RD, bool UPD>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5)))
void smooth_pair_pre_kernel(PairArgs<float> a) {
  smooth_pair_body<float, D, false, RD, UPD>(a, blockIdx.x, 0);
}
template <typename ZT, int D, bool RD>
__global__ __launch_bounds__(kBlock) void smooth_pair_post_kernel(PairArgs<ZT> a) {
  smooth_pair_body<ZT, D, true, RD, false>(a, blockIdx.x, 0);
}

// The same passes over level 0 of several hierarchies in one launch: hierarchy i takes
// workgroups first[i] .. first[i + 1] - 1.
template <typename ZT>
struct PairArgGroups {
  int n;
  int first[kMaxBatchGroups + 1];
  PairArgs<ZT> a[kMaxBatchGroups];
};
__device__ __forceinline__ int pair_group(const int* first, int n, int b) {
  int i = 0;
  while (i + 1 < n && b >= first[i + 1]) ++i;
  return i;
}
template <int D, bool RD, bool UPD>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5)))
void smooth_pair_pre_groups_kernel(PairArgGroups<float> g) {
  const int i = pair_group(g.first, g.n, blockIdx.x);
  smooth_pair_body<float, D, false, RD, UPD>(
      g.a[i], blockIdx.x - g.first[i],
      (int)(offsetof(PairArgGroups<float>, a) + i * sizeof(PairArgs<float>)));
}
template <typename ZT, int D, bool RD>
__global__ __launch_bounds__(kBlock) void smooth_pair_post_groups_kernel(PairArgGroups<ZT> g) {
  const int i = pair_group(g.first, g.n, blockIdx.x);
  smooth_pair_body<ZT, D, true, RD, false>(
      g.a[i], blockIdx.x - g.first[i],
      (int)(offsetof(PairArgGroups<ZT>, a) + i * sizeof(PairArgs<ZT>)));
}

template <typename ZT>
PairArgs<ZT> pair_args(const OpDesc& op, int rows) {
  PairArgs<ZT> a{};
  a.nelx = op.nelx;
  a.nely = op.nely;
  a.nnx = op.nnx();
  a.nny = op.nny();
  a.ntx = ceil_div(op.nnx(), kOutX);
  a.tiles = pair_tiles(op);
  a.total = a.tiles * rows;
  a.rho = static_cast<const float*>(op.rho);
  a.sr = op.sr;
  a.rdiag = static_cast<const float*>(op.rdiag);
  a.srd = op.srd;
  a.mask = op.mask;
  a.ke = make_ke<float>(op.ke, op.d);
  return a;
}

template <bool UPD>
void launch_pre(const PairArgs<float>& a, int d, hipStream_t s) {
  const dim3 grid(a.total), block(kBlock);
  const bool rd = a.rdiag != nullptr;
  if (d == 2) {
    if (rd) hipLaunchKernelGGL((smooth_pair_pre_kernel<2, true, UPD>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((smooth_pair_pre_kernel<2, false, UPD>), grid, block, 0, s, a);
  } else {
    if (rd) hipLaunchKernelGGL((smooth_pair_pre_kernel<1, true, UPD>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((smooth_pair_pre_kernel<1, false, UPD>), grid, block, 0, s, a);
  }
}
template <typename ZT>
void launch_post(const PairArgs<ZT>& a, int d, hipStream_t s) {
  const dim3 grid(a.total), block(kBlock);
  const bool rd = a.rdiag != nullptr;
  if (d == 2) {
    if (rd) hipLaunchKernelGGL((smooth_pair_post_kernel<ZT, 2, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((smooth_pair_post_kernel<ZT, 2, false>), grid, block, 0, s, a);
  } else {
    if (rd) hipLaunchKernelGGL((smooth_pair_post_kernel<ZT, 1, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((smooth_pair_post_kernel<ZT, 1, false>), grid, block, 0, s, a);
  }
}

// One row of a batched level-0 vector must stay below 4 GiB (32-bit offsets).
bool fits_u32(const OpDesc& op) { return op.ndof() * (long long)sizeof(double) < (1LL << 32); }

}  // namespace

int pair_tiles(const OpDesc& op) {
  return ceil_div(op.nnx(), kOutX) * ceil_div(op.nny(), kOutY);
}

static int pre_args(const PairPreCall& c, PairArgs<float>& a) {
  const OpDesc& op = c.op;
  if (op.rho == nullptr || op.mask == nullptr || (op.d != 1 && op.d != 2) || !fits_u32(op)) {
    set_error("smooth_pair_pre: level-0 density operator with a free-dof mask, d = 1 or 2, "
              "rows below 4 GiB required");
    return STTOPT_EINVAL;
  }
  const bool upd = c.ap != nullptr;
  if (upd && (c.alpha == nullptr || c.rnew == nullptr || c.rr == nullptr || c.rnew == c.b ||
              (c.zold != nullptr && c.rzo == nullptr))) {
    set_error("smooth_pair_pre: the residual update needs alpha, a separate output and the "
              "partial sums");
    return STTOPT_EINVAL;
  }
  if (op.d == 2 && !(pair_aligned(c.b, c.sb, sizeof(double)) &&
                     pair_aligned(c.invd, c.sd, sizeof(float)) &&
                     pair_aligned(c.x, c.sx, sizeof(float)) && pair_aligned(c.r, c.sr, sizeof(float)) &&
                     pair_aligned(op.rdiag, op.srd, sizeof(float)) &&
                     pair_aligned(c.ap, c.sap, sizeof(double)) &&
                     pair_aligned(c.rnew, c.srn, sizeof(double)) &&
                     pair_aligned(c.zold, c.szo, sizeof(float)))) {
    set_error("smooth_pair_pre: d = 2 vectors need node-aligned pointers and even row strides");
    return STTOPT_EINVAL;
  }
  a = pair_args<float>(op, c.rows);
  a.b = c.b;
  a.sb = c.sb;
  a.invd = c.invd;
  a.sd = c.sd;
  a.out = c.x;
  a.so = c.sx;
  a.res = c.r;
  a.sres = c.sr;
  a.omega = static_cast<float>(c.omega);
  a.active = c.active;
  if (upd) {
    a.ap = c.ap;
    a.sap = c.sap;
    a.alpha = c.alpha;
    a.rnew = c.rnew;
    a.srn = c.srn;
    a.rr = c.rr;
    a.zold = c.zold;
    a.szo = c.szo;
    a.rzo = c.rzo;
  }
  return STTOPT_OK;
}

template <typename ZT>
static int post_args(const PairPostCall<ZT>& c, PairArgs<ZT>& a) {
  const OpDesc& op = c.op;
  if (op.rho == nullptr || op.mask == nullptr || (op.d != 1 && op.d != 2) ||
      op.nelx % 2 != 0 || op.nely % 2 != 0 || c.cnnx != op.nelx / 2 + 1 || !fits_u32(op)) {
    set_error("smooth_pair_post: level-0 density operator over a 2 x 2 coarsening required");
    return STTOPT_EINVAL;
  }
  if (op.d == 2 && !(pair_aligned(c.b, c.sb, sizeof(double)) &&
                     pair_aligned(c.invd, c.sd, sizeof(float)) &&
                     pair_aligned(c.x, c.sx, sizeof(float)) && pair_aligned(c.ec, c.sec, sizeof(float)) &&
                     pair_aligned(c.z, c.sz, sizeof(ZT)) &&
                     pair_aligned(op.rdiag, op.srd, sizeof(float)))) {
    set_error("smooth_pair_post: d = 2 vectors need node-aligned pointers and even row strides");
    return STTOPT_EINVAL;
  }
  a = pair_args<ZT>(op, c.rows);
  a.cnnx = c.cnnx;
  a.x = c.x;
  a.sx = c.sx;
  a.ec = c.ec;
  a.sec = c.sec;
  a.b = c.b;
  a.sb = c.sb;
  a.invd = c.invd;
  a.sd = c.sd;
  a.out = c.z;
  a.so = c.sz;
  a.omega = static_cast<float>(c.omega);
  a.active = c.active;
  a.partial = c.partial;
  return STTOPT_OK;
}

int smooth_pair_pre(const OpDesc& op, int rows, const double* b, long long sb,
                    const float* invd, long long sd, float* x, long long sx, float* r,
                    long long sr, double omega, const int* active, hipStream_t s,
                    const double* ap, long long sap, const double* alpha, double* rnew,
                    long long srn, double* rr, const float* zold, long long szo,
                    double* rzo) {
  if (rows <= 0) return STTOPT_OK;
  const PairPreCall c{op, rows, b, sb, invd, sd, x, sx, r, sr, omega, active, ap, sap, alpha,
                      rnew, srn, rr, zold, szo, rzo};
  PairArgs<float> a;
  if (int e = pre_args(c, a)) return e;
  if (ap != nullptr) launch_pre<true>(a, op.d, s);
  else launch_pre<false>(a, op.d, s);
  STT_CHECK_LAUNCH("smooth_pair_pre_kernel");
  return STTOPT_OK;
}

template <typename ZT>
int smooth_pair_post(const OpDesc& op, int rows, const float* x, long long sx,
                     const float* ec, long long sec, int cnnx, const double* b, long long sb,
                     const float* invd, long long sd, ZT* z, long long sz, double omega,
                     const int* active, double* partial, hipStream_t s) {
  if (rows <= 0) return STTOPT_OK;
  const PairPostCall<ZT> c{op, rows, x, sx, ec, sec, cnnx, b, sb, invd, sd, z, sz, omega,
                           active, partial};
  PairArgs<ZT> a;
  if (int e = post_args<ZT>(c, a)) return e;
  launch_post<ZT>(a, op.d, s);
  STT_CHECK_LAUNCH("smooth_pair_post_kernel");
  return STTOPT_OK;
}

int smooth_pair_pre_groups(const PairPreCall* calls, int n, hipStream_t s) {
  if (n < 1 || n > kMaxBatchGroups) {
    set_error("smooth_pair_pre groups: 1 to 4 hierarchies per launch");
    return STTOPT_EINVAL;
  }
  PairArgGroups<float> g{};
  int total = 0, d = 0, rd = -1, upd = -1;
  for (int i = 0; i < n; ++i) {
    if (calls[i].rows <= 0) continue;
    PairArgs<float> a;
    if (int e = pre_args(calls[i], a)) return e;
    const int crd = a.rdiag != nullptr, cupd = calls[i].ap != nullptr;
    if (rd >= 0 && (calls[i].op.d != d || crd != rd || cupd != upd)) {
      set_error("smooth_pair_pre groups: one kind of pass per launch");
      return STTOPT_EINVAL;
    }
    d = calls[i].op.d;
    rd = crd;
    upd = cupd;
    g.a[g.n] = a;
    g.first[g.n] = total;
    total += a.total;
    ++g.n;
  }
  g.first[g.n] = total;
  if (total == 0) return STTOPT_OK;
  const dim3 grid(total), block(kBlock);
#define STT_PRE(D, RD, UPD) \
  hipLaunchKernelGGL((smooth_pair_pre_groups_kernel<D, RD, UPD>), grid, block, 0, s, g)
  if (d == 2) {
    if (rd) { if (upd) STT_PRE(2, true, true); else STT_PRE(2, true, false); }
    else { if (upd) STT_PRE(2, false, true); else STT_PRE(2, false, false); }
  } else {
    if (rd) { if (upd) STT_PRE(1, true, true); else STT_PRE(1, true, false); }
    else { if (upd) STT_PRE(1, false, true); else STT_PRE(1, false, false); }
  }
#undef STT_PRE
  STT_CHECK_LAUNCH("smooth_pair_pre_groups_kernel");
  return STTOPT_OK;
}

int smooth_pair_post_groups(const PairPostCall<float>* calls, int n, hipStream_t s) {
  if (n < 1 || n > kMaxBatchGroups) {
    set_error("smooth_pair_post groups: 1 to 4 hierarchies per launch");
    return STTOPT_EINVAL;
  }
  PairArgGroups<float> g{};
  int total = 0, d = 0, rd = -1;
  for (int i = 0; i < n; ++i) {
    if (calls[i].rows <= 0) continue;
    PairArgs<float> a;
    if (int e = post_args<float>(calls[i], a)) return e;
    const int crd = a.rdiag != nullptr;
    if (rd >= 0 && (calls[i].op.d != d || crd != rd)) {
      set_error("smooth_pair_post groups: one kind of pass per launch");
      return STTOPT_EINVAL;
    }
    d = calls[i].op.d;
    rd = crd;
    g.a[g.n] = a;
    g.first[g.n] = total;
    total += a.total;
    ++g.n;
  }
  g.first[g.n] = total;
  if (total == 0) return STTOPT_OK;
  const dim3 grid(total), block(kBlock);
#define STT_POST(D, RD) \
  hipLaunchKernelGGL((smooth_pair_post_groups_kernel<float, D, RD>), grid, block, 0, s, g)
  if (d == 2) {
    if (rd) STT_POST(2, true); else STT_POST(2, false);
  } else {
    if (rd) STT_POST(1, true); else STT_POST(1, false);
  }
#undef STT_POST
  STT_CHECK_LAUNCH("smooth_pair_post_groups_kernel");
  return STTOPT_OK;
}

template int smooth_pair_post<float>(const OpDesc&, int, const float*, long long, const float*,
                                     long long, int, const double*, long long, const float*,
                                     long long, float*, long long, double, const int*, double*,
                                     hipStream_t);
template int smooth_pair_post<double>(const OpDesc&, int, const float*, long long,
                                      const float*, long long, int, const double*, long long,
                                      const float*, long long, double*, long long, double,
                                      const int*, double*, hipStream_t);

}  // namespace stt
