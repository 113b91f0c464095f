// Matrix-free Q4 operator kernels for gfx950.
//
// One thread owns one node and gathers the contributions of the (up to) four
// elements around it, so the scatter of torch_fem.matvec's index_add_
// (torch_fem.py:76-83) becomes a deterministic, atomic-free gather.  The element
// dof map is never read: node and element indices follow from the structured grid.
// A workgroup is a 64 x 4 node tile (one wavefront per node row); the 3 x 3 node
// neighbourhood each thread reads is shared with its neighbours through L1, and the
// tile order is remapped so that each XCD streams a contiguous band of the grid.
// SIMP scaling, the lumped reaction diagonal, boundary projection and the
// damped-Jacobi / residual epilogues of the V-cycle are fused into the same pass.
#include <cstddef>
#include <cstdio>
#include <mutex>

#include <hip/hip_ext.h>

#include "stt_kernels.h"

namespace stt {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_status(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return STTOPT_EHIP;
}

template <typename T, typename TB, typename TY>
struct StencilArgs {
  int nelx, nely, nnx, nny, ntx, tiles, total;
  int mode, proj_in, dot;
  const T* u;
  long long su;
  const T* rho;
  long long sr;
  const T* keff;
  long long sk;
  const T* rdiag;
  long long srd;
  const uint32_t* mask;
  const TB* b;
  long long sb;
  const T* invd;
  long long sd;
  TY* y;
  long long sy;
  T omega;
  const int* active;
  double* partial;
  const T* ptab;
  int pkx, pky, fnelx;
  const T* prho;
  long long spr;
  T* bcopy;
  long long sbc;
  // kOperatorDir: u holds p_old; the staged (and applied) vector is p_new = z + beta p_old;
  // the tile's own nodes write p_new to dp and, for rows with iters == dk, take
  // x += alpha p_old, preceded by x += alpha_prev p_prev when dpp is given (the update
  // of iteration dk - 1, held back by the previous launch); with ddefer the rows still
  // iterating hold theirs back instead (rows that retired in dk take it)
  double* dp;
  const void* dz;
  double* dx;
  long long sdx;
  const double* dalpha;
  const double* dbeta;
  const int* diters;
  int dk, dz_f32;
  const double* dpp;
  const double* dalpha_prev;
  int ddefer;
  // kSmoothProlong: the coarse correction (row stride spe) on a grid of pcnnx nodes per row
  const T* pe;
  long long spe;
  int pcnnx;
  KE<T> ke;
};

// SRC: 0 = density * KE (level 0), 1 = explicit entry-major keff, 2 = keff
// parameterized by the fine densities under each coarse element.
// 3 = explicit keff, packed upper triangle.  4 = assembled node stencil
// (STTOPT_KEFF_NODE).
enum { kSrcDensity = 0, kSrcKeff = 1, kSrcParam = 2, kSrcKeffUpper = 3, kSrcNode = 4 };

// Entry (r, c) of a K x K symmetric matrix in the packed upper-triangle order.
__host__ __device__ constexpr int upper_index(int r, int c, int K) {
  return r <= c ? r * K - r * (r - 1) / 2 + (c - r) : c * K - c * (c - 1) / 2 + (r - c);
}

// Coarse-level stencil (SRC 1 or 2): one node per thread on 64 x 4 node tiles.
// PKX, PKY > 0 fix the parameterized block at compile time (fully unrolled).
template <typename T, typename TB, typename TY, int D, int SRC, int PKX = 0, int PKY = 0>
__global__ __launch_bounds__(kBlock) void stencil_kernel(StencilArgs<T, TB, TY> a) {
  constexpr int K = 4 * D;
  __shared__ double red[kBlock / kWave];
  const int t = xcd_remap(blockIdx.x, a.total);
  int row, tt;
  row_tile(t, a.tiles, a.total, row, tt);
  if (a.active != nullptr && a.active[row] == 0) return;  // whole block: uniform
  const int by = tt / a.ntx, bx = tt - by * a.ntx;
  const int c = bx * kTileX + (threadIdx.x & (kTileX - 1));
  const int r = by * kSmallTileY + (threadIdx.x / kTileX);
  double dot = 0.0;

  if (c < a.nnx && r < a.nny) {
    const T* u = a.u + row * a.su;
    // kSmoothFirst: the input is the first sweep from zero, x1 = w D^-1 b (the
    // expression of jacobi_first_kernel), formed here at each neighbour; D^-1 is 0 at
    // fixed dofs, so x1 is zero there without the mask (a -0 at most, which sums and
    // compares as +0)
    const bool first = a.mode == kSmoothFirst;
    // 3 x 3 node neighbourhood, zero outside the grid (and at fixed dofs when the
    // input has to be projected).
    T U[3][3][D];
#pragma unroll
    for (int dr = 0; dr < 3; ++dr) {
#pragma unroll
      for (int dc = 0; dc < 3; ++dc) {
        const int rr = r + dr - 1, cc = c + dc - 1;
        const bool ok = rr >= 0 && rr < a.nny && cc >= 0 && cc < a.nnx;
        const long long node = (long long)rr * a.nnx + cc;
        Vec<T, D> w;
        if (ok && first) {
          const Vec<TB, D> bn = load_node<TB, D>(a.b + row * a.sb, node);
          const Vec<T, D> dn = load_node<T, D>(a.invd + row * a.sd, node);
#pragma unroll
          for (int k = 0; k < D; ++k) w.v[k] = a.omega * dn.v[k] * static_cast<T>(bn.v[k]);
        } else if (ok) {
          w = load_node<T, D>(u, node);
        } else {
#pragma unroll
          for (int k = 0; k < D; ++k) w.v[k] = T(0);
        }
        const uint32_t word =
            (a.proj_in && ok && a.mask != nullptr) ? a.mask[(node * D) >> 5] : ~0u;
#pragma unroll
        for (int k = 0; k < D; ++k)
          U[dr][dc][k] = ((word >> ((node * D + k) & 31)) & 1u) ? w.v[k] : T(0);
      }
    }

    T acc[D];
#pragma unroll
    for (int i = 0; i < D; ++i) acc[i] = T(0);
    if constexpr (SRC == kSrcNode) {
      // own block (upper triangle), the forward blocks E, NW, N, NE stored here, and the
      // backward neighbours' forward blocks toward this node, transposed; each record
      // and block is read with vector loads
      constexpr int NS = node_stencil_stride(D), ND = 4, B = D * D;
      constexpr int fdr[4] = {0, 1, 1, 1}, fdc[4] = {1, -1, 0, 1};
      const T* S = a.keff + row * a.sk;
      const long long node = (long long)r * a.nnx + c;
      T own[NS];
      load_record<T, NS>(S + node * NS, own);
      int idx = 0;
#pragma unroll
      for (int i = 0; i < D; ++i) {
#pragma unroll
        for (int k = i; k < D; ++k) {
          const T v = own[idx++];
          acc[i] += v * U[1][1][k];
          if (k != i) acc[k] += v * U[1][1][i];
        }
      }
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const int cc = c + fdc[f];
        if (r + fdr[f] < a.nny && cc >= 0 && cc < a.nnx) {
#pragma unroll
          for (int i = 0; i < D; ++i) {
#pragma unroll
            for (int k = 0; k < D; ++k)
              acc[i] += own[ND + (f * D + i) * D + k] * U[1 + fdr[f]][1 + fdc[f]][k];
          }
        }
      }
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const int rr = r - fdr[f], cc = c - fdc[f];
        if (rr >= 0 && cc >= 0 && cc < a.nnx) {
          T blk[B];
          load_record<T, B>(S + ((long long)rr * a.nnx + cc) * NS + ND + f * B, blk);
#pragma unroll
          for (int i = 0; i < D; ++i) {
#pragma unroll
            for (int k = 0; k < D; ++k) acc[i] += blk[k * D + i] * U[1 - fdr[f]][1 - fdc[f]][k];
          }
        }
      }
    }
    // q: element up-left, up-right, down-left, down-right of the node; the node is
    // that element's local BR, BL, TR, TL corner respectively.
#pragma unroll
    for (int q = 0; q < 4 && SRC != kSrcNode; ++q) {
      const int er = r - 1 + (q >> 1), ec = c - 1 + (q & 1);
      if (er < 0 || er >= a.nely || ec < 0 || ec >= a.nelx) continue;
      const long long e = (long long)er * a.nelx + ec;
      constexpr int locs[4] = {1, 0, 2, 3};
      const int loc = locs[q];
      if constexpr (SRC == kSrcDensity) {
        T part[D];
#pragma unroll
        for (int i = 0; i < D; ++i) part[i] = T(0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int ur = (q >> 1) + loc_dr(j), uc = (q & 1) + loc_dc(j);
#pragma unroll
          for (int i = 0; i < D; ++i) {
#pragma unroll
            for (int k = 0; k < D; ++k)
              part[i] += a.ke.v[(loc * D + i) * K + j * D + k] * U[ur][uc][k];
          }
        }
        const T s = a.rho[row * a.sr + e];
#pragma unroll
        for (int i = 0; i < D; ++i) acc[i] += s * part[i];
      } else if constexpr (SRC == kSrcParam) {
        // keff_E rows (loc*D + i) = sum_g rho_g G_g rows: contract each G_g with the
        // element's nodal values, then weight by the child density.
        const T* rho = a.prho + row * a.spr;
        const int pkx = PKX > 0 ? PKX : a.pkx, pky = PKY > 0 ? PKY : a.pky;
#pragma unroll
        for (int gy = 0; gy < pky; ++gy) {
          const long long frow = (long long)(er * pky + gy) * a.fnelx + (long long)ec * pkx;
#pragma unroll
          for (int gx = 0; gx < pkx; ++gx) {
            const T s = rho[frow + gx];
            const T* G = a.ptab + (gy * pkx + gx) * K * K;
            T part[D];
#pragma unroll
            for (int i = 0; i < D; ++i) part[i] = T(0);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int ur = (q >> 1) + loc_dr(j), uc = (q & 1) + loc_dc(j);
#pragma unroll
              for (int i = 0; i < D; ++i) {
#pragma unroll
                for (int k = 0; k < D; ++k) part[i] += G[(loc * D + i) * K + j * D + k] * U[ur][uc][k];
              }
            }
#pragma unroll
            for (int i = 0; i < D; ++i) acc[i] += s * part[i];
          }
        }
      } else {
        const long long nel = (long long)a.nelx * a.nely;
        const T* ke_e = a.keff + row * a.sk + e;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int ur = (q >> 1) + loc_dr(j), uc = (q & 1) + loc_dc(j);
#pragma unroll
          for (int i = 0; i < D; ++i) {
#pragma unroll
            for (int k = 0; k < D; ++k) {
              const int r = loc * D + i, cl = j * D + k;
              const int idx = SRC == kSrcKeffUpper ? upper_index(r, cl, K) : r * K + cl;
              acc[i] += ke_e[(long long)idx * nel] * U[ur][uc][k];
            }
          }
        }
      }
    }

    const long long node = (long long)r * a.nnx + c;
    if (a.rdiag != nullptr) {
      const Vec<T, D> rd = load_node<T, D>(a.rdiag + row * a.srd, node);
#pragma unroll
      for (int i = 0; i < D; ++i) acc[i] += rd.v[i] * U[1][1][i];
    }
    // one vector access per node and array, one mask word per node
    uint32_t word = ~0u;
    if (a.mode != kMatvec && a.mask != nullptr) word = a.mask[(node * D) >> 5];
    Vec<TB, D> bv;
    Vec<T, D> iv;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      bv.v[i] = TB(0);
      iv.v[i] = T(0);
    }
    const bool sweep = a.mode == kSmooth || first;
    if (sweep || a.mode == kResid) bv = load_node<TB, D>(a.b + row * a.sb, node);
    if (sweep) iv = load_node<T, D>(a.invd + row * a.sd, node);
    Vec<TY, D> yo;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const bool fr = (word >> ((node * D + i) & 31)) & 1u;
      T out = T(0);
      if (fr) {
        if (a.mode == kMatvec || a.mode == kOperator) {
          out = acc[i];
        } else if (a.mode == kResid) {
          out = static_cast<T>(bv.v[i]) - acc[i];
        } else {  // kSmooth, kSmoothFirst
          out = U[1][1][i] + a.omega * iv.v[i] * (static_cast<T>(bv.v[i]) - acc[i]);
        }
      }
      yo.v[i] = static_cast<TY>(out);
      if (a.dot == kDotU) dot += (double)U[1][1][i] * (double)yo.v[i];
      if (a.dot == kDotB) dot += (double)bv.v[i] * (double)yo.v[i];
    }
    store_node<TY, D>(a.y + row * a.sy, node, yo);
  }
  if (a.dot != kNoDot) {
    const double s = block_sum(dot, red);
    if (threadIdx.x == 0) a.partial[(long long)row * a.tiles + tt] = s;
  }
}

// Level-0 stencil (density * KE): one 64 x 16 node tile per workgroup.  The tile's
// nodes plus a one-node halo (66 x 18) and the 65 x 17 surrounding element
// densities are staged in LDS with coalesced loads (zeros outside the grid, so the
// compute loop has no boundary branches); each thread then walks 4 consecutive
// node rows with a 3-row register window.
constexpr int kHaloX = kTileX + 2, kHaloY = kTileY + 2;
constexpr int kElemX = kTileX + 1, kElemY = kTileY + 1;

template <typename T, typename TB, typename TY, int D, int SRC, int MODE, int DOT>
__device__ __forceinline__ void stencil_tile_body(const StencilArgs<T, TB, TY>& a, const int bid,
                                                  const int karg_off) {
  constexpr int K = 4 * D;
  constexpr int NR = kRowsPerThread;
  // SRC kSrcDensity: level 0 (density * KE).  kSrcParam: level 1 at 2 x 2 coarsening,
  // keff_E = sum over the 4 level-0 elements g under E of rho_g G_pos(g); the tile
  // stages, per coarse element and level-0 row, the pair of level-0 densities under it.
  constexpr bool kParam = SRC == kSrcParam;
  constexpr int kRows = kParam ? 2 * kElemY : kElemY;  // density rows staged
  __shared__ Vec<T, D> sU[kHaloY][kHaloX];
  __shared__ T sR[kParam ? 1 : kElemY][kParam ? 1 : kElemX];
  __shared__ Vec<T, 2> sRf[kParam ? kRows : 1][kParam ? kElemX : 1];
  __shared__ double red[kWaves];
  // kSmoothFirst: b (in the stencil precision) and D^-1 of the interior nodes, kept from
  // the staging loads for the epilogue
  constexpr bool kFirst = MODE == kSmoothFirst;
  __shared__ Vec<T, D> sBc[kFirst ? kTileY : 1][kFirst ? kTileX : 1];
  __shared__ Vec<T, D> sIc[kFirst ? kTileY : 1][kFirst ? kTileX : 1];
  // free-dof words of the output rows: row r's dofs (gy * nnx + x0) D .. + 64 D - 1 span
  // at most kMW consecutive 32-bit words, the first at (gy * nnx + x0) D / 32
  constexpr int kMW = kTileX * D / 32 + 1;
  constexpr bool kOutMask = MODE != kMatvec;
  __shared__ uint32_t sMW[kOutMask ? kTileY : 1][kOutMask ? kMW : 1];
  constexpr bool kDir = MODE == kOperatorDir;
  constexpr bool kProl = MODE == kSmoothProlong;
  // kSmoothProlong: the coarse correction at the coarse nodes under the staged halo
  // (kPY x kPX from (y0 / 2 - 1, x0 / 2 - 1); x0 and y0 are even)
  constexpr int kPY = kHaloY / 2 + 1, kPX = kHaloX / 2 + 1;
  constexpr int kPE = (kPY * kPX + kBlock - 1) / kBlock;
  __shared__ Vec<T, D> sE[kProl ? kPY : 1][kProl ? kPX : 1];
  const int t = xcd_remap(bid, a.total);
  int row, tt;
  row_tile(t, a.tiles, a.total, row, tt);
  const int by = tt / a.ntx, bx = tt - by * a.ntx;
  const int x0 = bx * kTileX, y0 = by * kTileY;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);  // wave-uniform
  const T* __restrict__ u = a.u + row * a.su;
  if (a.active != nullptr && a.active[row] == 0) {  // whole block: uniform
    if constexpr (kDir) {
      // a row that retired in iteration dk still takes that iteration's x += alpha p
      // (and the held-back one of dk - 1)
      if (a.diters[row] == a.dk) {
        const double al = a.dalpha[row];
        const bool two = a.dpp != nullptr;
        const double alp = two ? a.dalpha_prev[row] : 0.0;
        double* xr = a.dx + row * a.sdx;
        const int gx = x0 + lane;
#pragma unroll
        for (int j = 0; j < NR; ++j) {
          const int gy = y0 + wave * NR + j;
          if (gy < a.nny && gx < a.nnx) {
            const long long node = (long long)gy * a.nnx + gx;
            const Vec<double, D> pv = load_node<double, D>(u, node);
            Vec<double, D> xv = load_node<double, D>(xr, node);
            if (two) {
              const Vec<double, D> pq = load_node<double, D>(a.dpp + row * a.su, node);
#pragma unroll
              for (int q = 0; q < D; ++q) xv.v[q] = fma(alp, pq.v[q], xv.v[q]);
            }
#pragma unroll
            for (int q = 0; q < D; ++q) xv.v[q] = fma(al, pv.v[q], xv.v[q]);
            store_node<double, D>(xr, node, xv);
          }
        }
      }
    }
    return;
  }
  const T* __restrict__ rho = kParam ? a.prho + row * a.spr : a.rho + row * a.sr;
  // kOperatorDir: beta and alpha of this row, and whether it takes x += alpha p_old
  const double dbe = kDir ? a.dbeta[row] : 0.0;
  const double dal = kDir ? a.dalpha[row] : 0.0;
  const bool dxup = kDir && a.diters[row] == a.dk && !a.ddefer;
  const bool dxtwo = dxup && a.dpp != nullptr;
  const double dalp = dxtwo ? a.dalpha_prev[row] : 0.0;
  auto stage_node = [&](long long node, bool owned) -> Vec<T, D> {
    if constexpr (kDir) {
      // p_new = z + beta p_old, the expression of the unfused direction update
      const Vec<double, D> pv = load_node<double, D>(u, node);
      Vec<double, D> zv;
      if (a.dz_f32) {
        const Vec<float, D> zf =
            load_node<float, D>(static_cast<const float*>(a.dz) + row * a.su, node);
#pragma unroll
        for (int q = 0; q < D; ++q) zv.v[q] = (double)zf.v[q];
      } else {
        zv = load_node<double, D>(static_cast<const double*>(a.dz) + row * a.su, node);
      }
      Vec<double, D> pn;
#pragma unroll
      for (int q = 0; q < D; ++q) pn.v[q] = zv.v[q] + dbe * pv.v[q];
      if (owned) {
        store_node<double, D>(a.dp + row * a.su, node, pn);
        if (dxup) {
          double* xr = a.dx + row * a.sdx;
          Vec<double, D> xv = load_node<double, D>(xr, node);
          if (dxtwo) {
            const Vec<double, D> pq = load_node<double, D>(a.dpp + row * a.su, node);
#pragma unroll
            for (int q = 0; q < D; ++q) xv.v[q] = fma(dalp, pq.v[q], xv.v[q]);
          }
#pragma unroll
          for (int q = 0; q < D; ++q) xv.v[q] = fma(dal, pv.v[q], xv.v[q]);
          store_node<double, D>(xr, node, xv);
        }
      }
      return pn;
    } else {
      return load_node<T, D>(u, node);
    }
  };
  // Stage the tile.  Wavefront w loads halo rows w, w + 4, ...; lane l loads column
  // x0 - 1 + l of each, so column validity is per lane and row validity per wavefront.
  // The two trailing node columns and the trailing element column of the halo are
  // loaded by the block's first 2 * kHaloY + kElemY threads.  All loads of a thread are
  // issued before its first LDS write.
  constexpr int kUR = (kHaloY + kWaves - 1) / kWaves;
  constexpr int kRR = (kRows + kWaves - 1) / kWaves;
  constexpr int kTailU = 2 * kHaloY, kTail = kTailU + kRows;
  const int pcy0 = y0 / 2 - 1, pcx0 = x0 / 2 - 1;
  Vec<T, D> pev[kProl ? kPE : 1];
  if constexpr (kProl) {
    const int cnny = a.nely / 2 + 1;
    const T* pe = a.pe + row * a.spe;
#pragma unroll
    for (int m = 0; m < kPE; ++m) {
      const int e = threadIdx.x + m * kBlock, cy = pcy0 + e / kPX, ccx = pcx0 + e % kPX;
#pragma unroll
      for (int c = 0; c < D; ++c) pev[m].v[c] = T(0);
      if (e < kPY * kPX && cy >= 0 && cy < cnny && ccx >= 0 && ccx < a.pcnnx)
        pev[m] = load_node<T, D>(pe, (long long)cy * a.pcnnx + ccx);
    }
  }
  const int cx = x0 - 1 + lane;
  const bool cnode = cx >= 0 && cx < a.nnx;
  const bool celem = cx >= 0 && cx < a.nelx;
  Vec<T, D> wu[kUR];
  T wr[kRR];
  uint32_t wm[kUR];  // free-dof words of the staged nodes (input projection)
  Vec<TB, D> wb[kFirst ? kUR : 1];
  Vec<T, D> wd[kFirst ? kUR : 1];
#pragma unroll
  for (int n = 0; n < kUR; ++n) {
    const int i = wave + n * kWaves, gy = y0 - 1 + i;
#pragma unroll
    for (int c = 0; c < D; ++c) wu[n].v[c] = T(0);
    if constexpr (kFirst) {
#pragma unroll
      for (int c = 0; c < D; ++c) {
        wb[n].v[c] = TB(0);
        wd[n].v[c] = T(0);
      }
    }
    wm[n] = ~0u;
    if (i < kHaloY && gy >= 0 && gy < a.nny && cnode) {
      const long long node = (long long)gy * a.nnx + cx;
      if constexpr (kFirst) {
        wb[n] = load_node<TB, D>(a.b + row * a.sb, node);
        wd[n] = load_node<T, D>(a.invd + row * a.sd, node);
      } else {
        // the tile's own nodes: rows 1 .. kTileY, lanes 1 .. 63 (column x0 + 63 is a
        // trailing column)
        wu[n] = stage_node(node, i >= 1 && i <= kTileY && lane >= 1);
      }
      if (a.proj_in || (kProl && a.mask != nullptr)) wm[n] = a.mask[(node * D) >> 5];
    }
  }
  Vec<T, 2> wr2[kParam ? kRR : 1];
#pragma unroll
  for (int n = 0; n < kRR; ++n) {
    const int i = wave + n * kWaves;
    if constexpr (kParam) {
      // level-0 row 2 ey + (i % 2) under coarse element row ey
      const int ey = y0 - 1 + (i >> 1);
      wr2[n].v[0] = wr2[n].v[1] = T(0);
      if (i < kRows && ey >= 0 && ey < a.nely && celem)
        wr2[n] = load_node<T, 2>(rho, (long long)(2 * ey + (i & 1)) * a.fnelx / 2 + cx);
    } else {
      const int ey = y0 - 1 + i;
      wr[n] = T(0);
      if (i < kElemY && ey >= 0 && ey < a.nely && celem) wr[n] = rho[(long long)ey * a.nelx + cx];
    }
  }
  // trailing columns: thread t < kTailU takes node (t / 2, 64 + t % 2), the next kElemY
  // threads take element (t - kTailU, 64)
  const int tix = threadIdx.x;
  const int ti = tix < kTailU ? (tix >> 1) : tix - kTailU;
  const int tk = tix < kTailU ? kTileX + (tix & 1) : kTileX;
  const int tgy = y0 - 1 + ti, tgx = x0 - 1 + tk;
  Vec<T, D> tu;
  T tr = T(0);
  Vec<TB, D> tb;
  Vec<T, D> td;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    tu.v[c] = T(0);
    tb.v[c] = TB(0);
    td.v[c] = T(0);
  }
  uint32_t tm = ~0u;  // kSmoothProlong: the trailing node's free-dof word
  if (tix < kTailU && tgy >= 0 && tgy < a.nny && tgx < a.nnx) {
    const long long node = (long long)tgy * a.nnx + tgx;
    if constexpr (kFirst) {
      tb = load_node<TB, D>(a.b + row * a.sb, node);
      td = load_node<T, D>(a.invd + row * a.sd, node);
    } else {
      tu = stage_node(node, tk == kTileX && ti >= 1 && ti <= kTileY);
    }
    if (kProl && a.mask != nullptr) tm = a.mask[(node * D) >> 5];
  }
  Vec<T, 2> tr2;
  tr2.v[0] = tr2.v[1] = T(0);
  if constexpr (kParam) {
    const int fr = tix - kTailU, ey = y0 - 1 + (fr >> 1);
    if (tix >= kTailU && tix < kTail && ey >= 0 && ey < a.nely && tgx < a.nelx)
      tr2 = load_node<T, 2>(rho, (long long)(2 * ey + (fr & 1)) * a.fnelx / 2 + tgx);
  } else if (tix >= kTailU && tix < kTail && tgy >= 0 && tgy < a.nely && tgx < a.nelx) {
    tr = rho[(long long)tgy * a.nelx + tgx];
  }
  if constexpr (kFirst) {
    // x0 = w D^-1 b (the first sweep); interior nodes also keep b and D^-1 in LDS and
    // write the copy of b.  The last interior column (x0 + 63) is a trailing-column node.
    auto finish = [&](const Vec<TB, D>& bv, const Vec<T, D>& dv, int li, int lk, int gyy,
                      int gxx) -> Vec<T, D> {
      Vec<T, D> bc, o;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        bc.v[c] = static_cast<T>(bv.v[c]);
        o.v[c] = a.omega * dv.v[c] * bc.v[c];
      }
      if (li >= 1 && li <= kTileY && lk >= 1 && lk <= kTileX && gyy < a.nny && gxx < a.nnx) {
        sBc[li - 1][lk - 1] = bc;
        sIc[li - 1][lk - 1] = dv;
        if (a.bcopy != nullptr) store_node<T, D>(a.bcopy + row * a.sbc, (long long)gyy * a.nnx + gxx, bc);
      }
      return o;
    };
#pragma unroll
    for (int n = 0; n < kUR; ++n) {
      const int i = wave + n * kWaves;
      if (i < kHaloY) wu[n] = finish(wb[n], wd[n], i, lane, y0 - 1 + i, cx);
    }
    if (tix < kTailU) tu = finish(tb, td, ti, tk, tgy, tgx);
  }
  if constexpr (kProl) {
    // xp = x + P e at the free dofs of the staged nodes, the expression of prolong_add
    // (bilinear weights 1, 0 at even and 1/2, 1/2 at odd fine positions)
#pragma unroll
    for (int m = 0; m < kPE; ++m) {
      const int e = threadIdx.x + m * kBlock;
      if (e < kPY * kPX) sE[e / kPX][e % kPX] = pev[m];
    }
    __syncthreads();
    auto prolong = [&](Vec<T, D>& v, uint32_t word, int gy, int gx) {
      const long long node = (long long)gy * a.nnx + gx;
      const uint32_t bits = (word >> ((node * D) & 31)) & ((1u << D) - 1u);
      const int fy = gy & 1, fx = gx & 1;
      const T wy1 = fy ? T(0.5) : T(0), wy0 = T(1) - wy1;
      const T wx1 = fx ? T(0.5) : T(0), wx0 = T(1) - wx1;
      const int ly = (gy >> 1) - pcy0, lx = (gx >> 1) - pcx0;
      const Vec<T, D> c00 = sE[ly][lx];
      Vec<T, D> c01, c10, c11;
#pragma unroll
      for (int k = 0; k < D; ++k) c01.v[k] = c10.v[k] = c11.v[k] = T(0);
      if (fx) c01 = sE[ly][lx + 1];
      if (fy) c10 = sE[ly + 1][lx];
      if (fx && fy) c11 = sE[ly + 1][lx + 1];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const T top = wx0 * c00.v[k] + wx1 * c01.v[k];
        const T bot = wx0 * c10.v[k] + wx1 * c11.v[k];
        if ((bits >> k) & 1u) v.v[k] += wy0 * top + wy1 * bot;
      }
    };
#pragma unroll
    for (int n = 0; n < kUR; ++n) {
      const int i = wave + n * kWaves, gy = y0 - 1 + i;
      if (i < kHaloY && gy >= 0 && gy < a.nny && cnode) prolong(wu[n], wm[n], gy, cx);
    }
    if (tix < kTailU && tgy >= 0 && tgy < a.nny && tgx < a.nnx) prolong(tu, tm, tgy, tgx);
  }
  // Epilogue operands of this thread's output nodes (b, D^-1) are loaded now, before
  // the barrier, so that the epilogue does not wait for them; the output rows' free-dof
  // words are loaded by the first kTileY * kMW threads (one word each) and pass through
  // LDS.
  constexpr bool kSweep = MODE == kSmooth || kProl;  // kSmoothFirst reads them from LDS
  constexpr bool kNeedB = kSweep || MODE == kResid;
  const bool has_mask = kOutMask && a.mask != nullptr;
  uint32_t mw = ~0u;
  if (has_mask && tix < kTileY * kMW) {
    const int r = tix / kMW, gy = y0 + r;
    const long long w = (((long long)gy * a.nnx + x0) * D >> 5) + (tix - r * kMW);
    if (gy < a.nny && w <= ((long long)a.nny * a.nnx * D - 1) >> 5) mw = a.mask[w];
  }
  Vec<TB, D> ob[kNeedB ? NR : 1];
  Vec<T, D> oi[kSweep ? NR : 1];
  {
    const int gxo = x0 + lane, gyo = y0 + wave * NR;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const bool in = gyo + j < a.nny && gxo < a.nnx;
      const long long node = (long long)(gyo + j) * a.nnx + gxo;
      if constexpr (kNeedB) {
#pragma unroll
        for (int c = 0; c < D; ++c) ob[j].v[c] = TB(0);
        if (in) ob[j] = load_node<TB, D>(a.b + row * a.sb, node);
      }
      if constexpr (kSweep) {
#pragma unroll
        for (int c = 0; c < D; ++c) oi[j].v[c] = T(0);
        if (in) oi[j] = load_node<T, D>(a.invd + row * a.sd, node);
      }
    }
  }
  if (a.proj_in) {
#pragma unroll
    for (int n = 0; n < kUR; ++n) {
      const int i = wave + n * kWaves, gy = y0 - 1 + i;
      if (i < kHaloY && gy >= 0 && gy < a.nny && cnode) {
        const long long dof0 = ((long long)gy * a.nnx + cx) * D;
#pragma unroll
        for (int c = 0; c < D; ++c)
          if (!((wm[n] >> ((dof0 + c) & 31)) & 1u)) wu[n].v[c] = T(0);
      }
    }
    if (tix < kTailU && tgy >= 0 && tgy < a.nny && tgx < a.nnx) {
      const long long dof0 = ((long long)tgy * a.nnx + tgx) * D;
      const uint32_t word = a.mask[dof0 >> 5];
#pragma unroll
      for (int c = 0; c < D; ++c)
        if (!((word >> ((dof0 + c) & 31)) & 1u)) tu.v[c] = T(0);
    }
  }
#pragma unroll
  for (int n = 0; n < kUR; ++n) {
    const int i = wave + n * kWaves;
    if (i < kHaloY) sU[i][lane] = wu[n];
  }
#pragma unroll
  for (int n = 0; n < kRR; ++n) {
    const int i = wave + n * kWaves;
    if constexpr (kParam) {
      if (i < kRows) sRf[i][lane] = wr2[n];
    } else if (i < kElemY) {
      sR[i][lane] = wr[n];
    }
  }
  if (tix < kTailU) {
    sU[ti][tk] = tu;
  } else if (tix < kTail) {
    if constexpr (kParam) sRf[ti][tk] = tr2;
    else sR[ti][tk] = tr;
  }
  if constexpr (kOutMask) {
    if (tix < kTileY * kMW) sMW[tix / kMW][tix % kMW] = mw;
  }
  __syncthreads();

  const int tx = lane, ty = wave;
  const int gx = x0 + tx;
  const int i0 = ty * NR;  // LDS row of the window's first (halo) node row
  // fp32: the thread's whole (NR + 2) x 3 node window and (NR + 1) x 2 densities live in
  // registers.  fp64: only the (NR + 1) x 2 sub-window of one element position at a
  // time, read from LDS per position, which keeps the kernel under 96 VGPRs.
  constexpr bool kFullWindow = sizeof(T) == 4;
  Vec<T, D> W[kFullWindow ? NR + 2 : 1][3];
  T R[kFullWindow ? NR + 1 : 1][2];
  if constexpr (kFullWindow) {
#pragma unroll
    for (int r = 0; r < NR + 2; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) W[r][c] = sU[i0 + r][tx + c];
    }
    if constexpr (!kParam) {
#pragma unroll
      for (int r = 0; r < NR + 1; ++r) {
        R[r][0] = sR[i0 + r][tx];
        R[r][1] = sR[i0 + r][tx + 1];
      }
    }
  }
  T acc[NR][D];
#pragma unroll
  for (int j = 0; j < NR; ++j) {
#pragma unroll
    for (int ii = 0; ii < D; ++ii) acc[j][ii] = T(0);
  }
  // element position outer, node row inner: each KE row pair is read once per thread.
  // KE is read from the kernel-argument segment where it is used, so that only one row
  // pair is live in scalar registers (the 64 fp64 entries do not fit at once).
  using Args = StencilArgs<T, TB, TY>;
  using ConstPtr = const __attribute__((address_space(4))) T*;
  const ConstPtr kargs_ke = reinterpret_cast<ConstPtr>(
      (__attribute__((address_space(4))) char*)__builtin_amdgcn_kernarg_segment_ptr() +
      karg_off + offsetof(Args, ke));  // karg_off: where `a` sits in the argument segment
  // level 1: the per-position tables, read the same way (constant address space, one
  // element position's 4 x (D x 4D) rows live at a time)
  const ConstPtr ptab_c =
      reinterpret_cast<ConstPtr>(reinterpret_cast<uintptr_t>(kParam ? a.ptab : nullptr));
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    constexpr int locs[4] = {1, 0, 2, 3};
    const int loc = locs[q];
    if constexpr (kParam) {
      ConstPtr tq = ptab_c + loc * D * K;
#pragma unroll
      for (int j = 0; j < NR; ++j) {
#pragma unroll
        for (int ii = 0; ii < D; ++ii) asm volatile("" : "+v"(acc[j][ii]));
      }
      asm volatile("" : "+s"(tq));
      T tab[4][D][K];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
#pragma unroll
        for (int ii = 0; ii < D; ++ii) {
#pragma unroll
          for (int c = 0; c < K; ++c) tab[g][ii][c] = tq[g * K * K + ii * K + c];
        }
      }
      const int qr = q >> 1, qc = q & 1;
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        // the four level-0 densities under this element: rows 2 er, 2 er + 1 of the pairs
        const Vec<T, 2> r0 = sRf[2 * (i0 + j + qr)][tx + qc];
        const Vec<T, 2> r1 = sRf[2 * (i0 + j + qr) + 1][tx + qc];
        const T rg[4] = {r0.v[0], r0.v[1], r1.v[0], r1.v[1]};
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          T part[D];
#pragma unroll
          for (int ii = 0; ii < D; ++ii) part[ii] = T(0);
#pragma unroll
          for (int jn = 0; jn < 4; ++jn) {
            const Vec<T, D>& un = W[j + qr + loc_dr(jn)][qc + loc_dc(jn)];
#pragma unroll
            for (int ii = 0; ii < D; ++ii) {
#pragma unroll
              for (int kk = 0; kk < D; ++kk) part[ii] += tab[g][ii][jn * D + kk] * un.v[kk];
            }
          }
#pragma unroll
          for (int ii = 0; ii < D; ++ii) acc[j][ii] += rg[g] * part[ii];
        }
      }
      continue;
    }
    // fp64: the 64 KE entries do not fit in scalar registers.  The empty asm makes the
    // row pointer depend on the accumulators of the previous position, so the compiler
    // cannot hoist all four row-pair loads to the top and spill them.
    ConstPtr kq = kargs_ke + loc * D * K;
    if constexpr (sizeof(T) == 8) {
#pragma unroll
      for (int j = 0; j < NR; ++j) {
#pragma unroll
        for (int ii = 0; ii < D; ++ii) asm volatile("" : "+v"(acc[j][ii]));
      }
      asm volatile("" : "+s"(kq));
    }
    T ke[D][K];
#pragma unroll
    for (int ii = 0; ii < D; ++ii) {
#pragma unroll
      for (int c = 0; c < K; ++c) ke[ii][c] = kq[ii * K + c];
    }
    const int qr = q >> 1, qc = q & 1;
    Vec<T, D> S[kFullWindow ? 1 : NR + 1][2];
    T RQ[kFullWindow ? 1 : NR];
    if constexpr (!kFullWindow) {
#pragma unroll
      for (int r = 0; r < NR + 1; ++r) {
#pragma unroll
        for (int c = 0; c < 2; ++c) S[r][c] = sU[i0 + qr + r][tx + qc + c];
      }
#pragma unroll
      for (int j = 0; j < NR; ++j) RQ[j] = sR[i0 + qr + j][tx + qc];
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      T part[D];
#pragma unroll
      for (int ii = 0; ii < D; ++ii) part[ii] = T(0);
#pragma unroll
      for (int jn = 0; jn < 4; ++jn) {
        const Vec<T, D>& un = kFullWindow ? W[j + qr + loc_dr(jn)][qc + loc_dc(jn)]
                                          : S[j + loc_dr(jn)][loc_dc(jn)];
#pragma unroll
        for (int ii = 0; ii < D; ++ii) {
#pragma unroll
          for (int kk = 0; kk < D; ++kk) part[ii] += ke[ii][jn * D + kk] * un.v[kk];
        }
      }
      const T dens = kFullWindow ? R[j + qr][qc] : RQ[j];
#pragma unroll
      for (int ii = 0; ii < D; ++ii) acc[j][ii] += dens * part[ii];
    }
  }

  double dot = 0.0;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const int gy = y0 + i0 + j;
    if (gy < a.nny && gx < a.nnx) {
      const long long node = (long long)gy * a.nnx + gx;
      const Vec<T, D> uc = kFullWindow ? W[j + 1][1] : sU[i0 + 1 + j][tx + 1];
      // one vector access per node and array, one mask word per node (the D dofs of a
      // node share it)
      if (a.rdiag != nullptr) {
        const Vec<T, D> rd = load_node<T, D>(a.rdiag + row * a.srd, node);
#pragma unroll
        for (int ii = 0; ii < D; ++ii) acc[j][ii] += rd.v[ii] * uc.v[ii];
      }
      uint32_t word = ~0u;
      if (has_mask)
        word = sMW[i0 + j][((node * D) >> 5) - ((((long long)gy * a.nnx + x0) * D) >> 5)];
      Vec<TB, D> bv;
      Vec<T, D> iv;
#pragma unroll
      for (int ii = 0; ii < D; ++ii) {
        bv.v[ii] = TB(0);
        iv.v[ii] = T(0);
      }
      if constexpr (kNeedB) bv = ob[j];
      if constexpr (kSweep) iv = oi[j];
      if constexpr (kFirst) {
#pragma unroll
        for (int ii = 0; ii < D; ++ii) bv.v[ii] = static_cast<TB>(sBc[i0 + j][tx].v[ii]);
        iv = sIc[i0 + j][tx];
      }
      Vec<TY, D> yo;
#pragma unroll
      for (int ii = 0; ii < D; ++ii) {
        const bool fr = (word >> ((node * D + ii) & 31)) & 1u;
        T out = T(0);
        if (fr) {
          if constexpr (MODE == kMatvec || MODE == kOperator || MODE == kOperatorDir) {
            out = acc[j][ii];
          } else if constexpr (MODE == kResid) {
            out = static_cast<T>(bv.v[ii]) - acc[j][ii];
          } else {
            out = uc.v[ii] + a.omega * iv.v[ii] * (static_cast<T>(bv.v[ii]) - acc[j][ii]);
          }
        }
        yo.v[ii] = static_cast<TY>(out);
        if constexpr (DOT == kDotU) dot += (double)uc.v[ii] * (double)yo.v[ii];
        if constexpr (DOT == kDotB) dot += (double)bv.v[ii] * (double)yo.v[ii];
      }
      store_node<TY, D>(a.y + row * a.sy, node, yo);
    }
  }
  if constexpr (DOT != kNoDot) {
    const double sum = block_sum(dot, red);
    if (threadIdx.x == 0) a.partial[(long long)row * a.tiles + tt] = sum;
  }
}

template <typename T, typename TB, typename TY, int D, int SRC, int MODE, int DOT>
__global__ __launch_bounds__(kBlock) void stencil_tile_kernel(StencilArgs<T, TB, TY> a) {
  stencil_tile_body<T, TB, TY, D, SRC, MODE, DOT>(a, blockIdx.x, 0);
}

// The same tile pass over the level of several hierarchies in one launch (the uneven
// batch): hierarchy i takes workgroups first[i] .. first[i + 1] - 1.
template <typename T, typename TB, typename TY>
struct StencilGroups {
  int n;
  int first[kMaxBatchGroups + 1];
  StencilArgs<T, TB, TY> a[kMaxBatchGroups];
};

template <typename T, typename TB, typename TY, int D, int SRC, int MODE, int DOT>
__global__ __launch_bounds__(kBlock) void stencil_tile_groups_kernel(StencilGroups<T, TB, TY> g) {
  int i = 0;
  while (i + 1 < g.n && (int)blockIdx.x >= g.first[i + 1]) ++i;
  using G = StencilGroups<T, TB, TY>;
  stencil_tile_body<T, TB, TY, D, SRC, MODE, DOT>(
      g.a[i], blockIdx.x - g.first[i],
      (int)(offsetof(G, a) + i * sizeof(StencilArgs<T, TB, TY>)));
}

template <typename T, typename TB, typename TY, int D, int SRC>
static int launch_tile(const StencilArgs<T, TB, TY>& a, dim3 grid, hipStream_t s) {
  const dim3 block(kBlock);
  switch (a.mode * 3 + a.dot) {
    case kMatvec * 3 + kNoDot:
      hipLaunchKernelGGL((stencil_tile_kernel<T, TB, TY, D, SRC, kMatvec, kNoDot>), grid, block, 0, s, a);
      break;
    case kOperator * 3 + kNoDot:
      hipLaunchKernelGGL((stencil_tile_kernel<T, TB, TY, D, SRC, kOperator, kNoDot>), grid, block, 0, s, a);
      break;
    case kOperator * 3 + kDotU:
      hipLaunchKernelGGL((stencil_tile_kernel<T, TB, TY, D, SRC, kOperator, kDotU>), grid, block, 0, s, a);
      break;
    case kSmooth * 3 + kNoDot:
      hipLaunchKernelGGL((stencil_tile_kernel<T, TB, TY, D, SRC, kSmooth, kNoDot>), grid, block, 0, s, a);
      break;
    case kSmooth * 3 + kDotB:
      hipLaunchKernelGGL((stencil_tile_kernel<T, TB, TY, D, SRC, kSmooth, kDotB>), grid, block, 0, s, a);
      break;
    case kResid * 3 + kNoDot:
      hipLaunchKernelGGL((stencil_tile_kernel<T, TB, TY, D, SRC, kResid, kNoDot>), grid, block, 0, s, a);
      break;
    case kSmoothFirst * 3 + kNoDot:
      hipLaunchKernelGGL((stencil_tile_kernel<T, TB, TY, D, SRC, kSmoothFirst, kNoDot>), grid, block, 0, s, a);
      break;
    default:
      set_error("stencil: unsupported mode / reduction combination");
      return STTOPT_EINVAL;
  }
  return STTOPT_OK;
}

// The kernel arguments of one stencil pass, and its source kind and tile flag.
template <typename T, typename TB, typename TY>
static int stencil_args(const OpDesc& op, int rows, int mode, bool proj_in, int dot,
                        const T* u, long long su, const TB* b, long long sb, const T* invd,
                        long long sd, TY* y, long long sy, double omega, const int* active,
                        double* partial, T* bcopy, long long sbc, StencilArgs<T, TB, TY>& a,
                        int& src_out, bool& param_tile_out) {
  a = StencilArgs<T, TB, TY>{};
  a.nelx = op.nelx;
  a.nely = op.nely;
  a.nnx = op.nnx();
  a.nny = op.nny();
  const int src = op.ptab != nullptr ? kSrcParam
                  : op.keff != nullptr
                      ? (op.keff_node ? kSrcNode : op.keff_upper ? kSrcKeffUpper : kSrcKeff)
                      : kSrcDensity;
  a.ntx = op.ntx();
  // level 0, and in fp32 the parameterized level 1 at 2 x 2, run on 64 x 16 LDS tiles
  const bool p22 = op.pkx == 2 && op.pky == 2;
  const bool param_tile = src == kSrcParam && p22 && sizeof(T) == sizeof(float);
  a.tiles = (src == kSrcDensity || param_tile) ? op.tiles() : op.tiles_small();
  if (src != kSrcDensity && dot != kNoDot) {
    set_error("stencil: reduction partials are level-0 only");
    return STTOPT_EINVAL;
  }
  if (mode == kSmoothFirst && (b == nullptr || invd == nullptr)) {
    set_error("stencil: the fused first sweeps need b and the inverse diagonal");
    return STTOPT_EINVAL;
  }
  if (op.d == 2) {
    // nodes are read and written as one 2-vector: pointers and row strides must keep
    // every node's pair of dofs aligned to the pair
    auto ok = pair_aligned;
    if (!ok(u, su, sizeof(T)) || !ok(y, sy, sizeof(TY)) || !ok(b, sb, sizeof(TB)) ||
        !ok(invd, sd, sizeof(T)) || !ok(op.rdiag, op.srd, sizeof(T)) ||
        !ok(bcopy, sbc, sizeof(T))) {
      set_error("stencil: d = 2 vectors need node-aligned pointers and even row strides");
      return STTOPT_EINVAL;
    }
  }
  const long long total = (long long)a.tiles * rows;
  if (total > 0x7fffffffLL) {
    set_error("stencil: grid too large");
    return STTOPT_EINVAL;
  }
  a.total = (int)total;
  a.mode = mode;
  a.proj_in = (proj_in && op.mask != nullptr) ? 1 : 0;  // projection needs the mask
  a.dot = dot;
  a.u = u;
  a.su = su;
  a.rho = static_cast<const T*>(op.rho);
  a.sr = op.sr;
  a.keff = static_cast<const T*>(op.keff);
  a.sk = op.sk;
  a.rdiag = static_cast<const T*>(op.rdiag);
  a.srd = op.srd;
  a.mask = op.mask;
  a.b = b;
  a.sb = sb;
  a.invd = invd;
  a.sd = sd;
  a.y = y;
  a.sy = sy;
  a.omega = static_cast<T>(omega);
  a.active = active;
  a.partial = partial;
  a.ptab = static_cast<const T*>(op.ptab);
  a.pkx = op.pkx;
  a.pky = op.pky;
  a.fnelx = op.fnelx;
  a.prho = static_cast<const T*>(op.prho);
  a.spr = op.spr;
  a.bcopy = bcopy;
  a.sbc = sbc;
  a.ke = make_ke<T>(op.ke, op.d);
  if (src == kSrcDensity && op.rho == nullptr) {
    set_error("stencil: level has neither density nor keff");
    return STTOPT_EINVAL;
  }
  if (src == kSrcParam && (op.prho == nullptr || op.pkx < 1 || op.pky < 1 ||
                           op.fnelx != op.nelx * op.pkx)) {
    set_error("stencil: parameterized level needs the fine density and a matching block");
    return STTOPT_EINVAL;
  }
  src_out = src;
  param_tile_out = param_tile;
  return STTOPT_OK;
}

template <typename T, typename TB, typename TY>
int stencil(const OpDesc& op, int rows, int mode, bool proj_in, int dot, const T* u,
            long long su, const TB* b, long long sb, const T* invd, long long sd,
            TY* y, long long sy, double omega, const int* active, double* partial,
            hipStream_t s, T* bcopy, long long sbc) {
  if (rows <= 0 || op.nelx <= 0 || op.nely <= 0) return STTOPT_OK;
  StencilArgs<T, TB, TY> a;
  int src = 0;
  bool param_tile = false;
  if (int e = stencil_args<T, TB, TY>(op, rows, mode, proj_in, dot, u, su, b, sb, invd, sd, y,
                                      sy, omega, active, partial, bcopy, sbc, a, src,
                                      param_tile))
    return e;
  const bool p22 = op.pkx == 2 && op.pky == 2;
  dim3 grid((unsigned)a.total), block(kBlock);
#define STT_STENCIL(D, SRC) \
  hipLaunchKernelGGL((stencil_kernel<T, TB, TY, D, SRC>), grid, block, 0, s, a)
#define STT_STENCIL22(D) \
  hipLaunchKernelGGL((stencil_kernel<T, TB, TY, D, kSrcParam, 2, 2>), grid, block, 0, s, a)
#define STT_TILE(D, SRC)                                                 \
  do {                                                                   \
    if (int e_ = launch_tile<T, TB, TY, D, SRC>(a, grid, s)) return e_;  \
  } while (0)
#define STT_PARAM(D)                                 \
  do {                                               \
    if constexpr (sizeof(T) == sizeof(float)) {      \
      if (param_tile) {                              \
        STT_TILE(D, kSrcParam);                      \
        break;                                       \
      }                                              \
    }                                                \
    if (p22) STT_STENCIL22(D);                       \
    else STT_STENCIL(D, kSrcParam);                  \
  } while (0)
  if (op.d == 2) {
    if (src == kSrcParam) STT_PARAM(2);
    else if (src == kSrcKeff) STT_STENCIL(2, kSrcKeff);
    else if (src == kSrcKeffUpper) STT_STENCIL(2, kSrcKeffUpper);
    else if (src == kSrcNode) STT_STENCIL(2, kSrcNode);
    else STT_TILE(2, kSrcDensity);
  } else if (op.d == 1) {
    if (src == kSrcParam) STT_PARAM(1);
    else if (src == kSrcKeff) STT_STENCIL(1, kSrcKeff);
    else if (src == kSrcKeffUpper) STT_STENCIL(1, kSrcKeffUpper);
    else if (src == kSrcNode) STT_STENCIL(1, kSrcNode);
    else STT_TILE(1, kSrcDensity);
#undef STT_STENCIL
#undef STT_STENCIL22
#undef STT_TILE
#undef STT_PARAM
  } else {
    set_error("stencil: d must be 1 or 2");
    return STTOPT_EINVAL;
  }
  STT_CHECK_LAUNCH("stencil_kernel");
  return STTOPT_OK;
}

int smooth_prolong(const OpDesc& op, int rows, const float* x, long long sx, const float* e,
                   long long se, int cnnx, int cnny, const float* b, long long sb,
                   const float* invd, long long sd, float* y, long long sy, double omega,
                   const int* active, hipStream_t s) {
  if (rows <= 0 || op.nelx <= 0 || op.nely <= 0) return STTOPT_OK;
  if (op.ptab == nullptr || op.pkx != 2 || op.pky != 2 || op.nelx % 2 != 0 ||
      op.nely % 2 != 0 || cnnx != op.nelx / 2 + 1 || cnny != op.nely / 2 + 1 ||
      e == nullptr || b == nullptr ||
      invd == nullptr || x == y) {
    set_error("smooth_prolong: needs the parameterized 2 x 2 level, its coarse grid, b, "
              "D^-1 and separate input and output");
    return STTOPT_EINVAL;
  }
  if (op.d == 2 && !pair_aligned(e, se, sizeof(float))) {
    set_error("stencil: d = 2 vectors need node-aligned pointers and even row strides");
    return STTOPT_EINVAL;
  }
  StencilArgs<float, float, float> a;
  int src = 0;
  bool param_tile = false;
  if (int err = stencil_args<float, float, float>(op, rows, kSmooth, false, kNoDot, x, sx, b,
                                                  sb, invd, sd, y, sy, omega, active, nullptr,
                                                  nullptr, 0, a, src, param_tile))
    return err;
  if (!param_tile) {
    set_error("smooth_prolong: the level does not run on LDS tiles");
    return STTOPT_EINVAL;
  }
  a.mode = kSmoothProlong;
  a.pe = e;
  a.spe = se;
  a.pcnnx = cnnx;
  dim3 grid((unsigned)a.total), block(kBlock);
  if (op.d == 2)
    hipLaunchKernelGGL((stencil_tile_kernel<float, float, float, 2, kSrcParam, kSmoothProlong,
                                            kNoDot>),
                       grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((stencil_tile_kernel<float, float, float, 1, kSrcParam, kSmoothProlong,
                                            kNoDot>),
                       grid, block, 0, s, a);
  STT_CHECK_LAUNCH("smooth_prolong");
  return STTOPT_OK;
}

int operator_dir(const OpDesc& op, int rows, const double* p_old, double* p_new,
                 const void* z, int z_f32, double* x, long long xs, const double* alpha,
                 const double* beta, const int* iters, int k, double* Ap,
                 const int* active, double* partial, hipStream_t s, hipEvent_t start,
                 hipEvent_t stop, const double* p_prev, const double* alpha_prev, int defer) {
  if (rows <= 0 || op.nelx <= 0 || op.nely <= 0) return STTOPT_OK;
  const long long n = op.ndof();
  StencilArgs<double, double, double> a;
  int src = 0;
  bool param_tile = false;
  if (int e = stencil_args<double, double, double>(op, rows, kOperator, false, kDotU, p_old, n,
                                                   nullptr, 0, nullptr, 0, Ap, n, 0.0, active,
                                                   partial, nullptr, 0, a, src, param_tile))
    return e;
  if (src != kSrcDensity || active == nullptr || p_new == nullptr || z == nullptr ||
      x == nullptr || alpha == nullptr || beta == nullptr || iters == nullptr ||
      p_new == p_old) {
    set_error("operator_dir: level-0 density operator with separate p buffers, z, x, "
              "alpha, beta, iteration counts and active flags");
    return STTOPT_EINVAL;
  }
  if ((p_prev != nullptr) != (alpha_prev != nullptr) || (defer && p_prev != nullptr)) {
    set_error("operator_dir: the held-back update needs p_prev and alpha_prev, and no deferral");
    return STTOPT_EINVAL;
  }
  if (op.d == 2 && (!pair_aligned(p_new, n, sizeof(double)) ||
                    !pair_aligned(p_prev, n, sizeof(double)) ||
                    !pair_aligned(x, xs, sizeof(double)) ||
                    !pair_aligned(z, n, z_f32 ? sizeof(float) : sizeof(double)))) {
    set_error("operator_dir: d = 2 vectors need node-aligned pointers and even row strides");
    return STTOPT_EINVAL;
  }
  a.mode = kOperatorDir;
  a.dp = p_new;
  a.dz = z;
  a.dz_f32 = z_f32 ? 1 : 0;
  a.dx = x;
  a.sdx = xs;
  a.dalpha = alpha;
  a.dbeta = beta;
  a.diters = iters;
  a.dk = k;
  a.dpp = p_prev;
  a.dalpha_prev = alpha_prev;
  a.ddefer = defer ? 1 : 0;
  const dim3 grid((unsigned)a.total), block(kBlock);
  auto launch = [&](auto kernel) {
    if (start != nullptr || stop != nullptr)
      hipExtLaunchKernelGGL(kernel, grid, block, 0, s, start, stop, 0, a);
    else
      hipLaunchKernelGGL(kernel, grid, block, 0, s, a);
  };
  if (op.d == 2)
    launch(stencil_tile_kernel<double, double, double, 2, kSrcDensity, kOperatorDir, kDotU>);
  else if (op.d == 1)
    launch(stencil_tile_kernel<double, double, double, 1, kSrcDensity, kOperatorDir, kDotU>);
  else {
    set_error("operator_dir: d must be 1 or 2");
    return STTOPT_EINVAL;
  }
  STT_CHECK_LAUNCH("operator_dir");
  return STTOPT_OK;
}

template <typename T>
int stencil_tile_groups(const StencilCall<T>* calls, int n, hipStream_t s) {
  if (n < 1 || n > kMaxBatchGroups) {
    set_error("stencil groups: 1 to 4 hierarchies per launch");
    return STTOPT_EINVAL;
  }
  StencilGroups<T, T, T> g{};
  int total = 0, src = -1, d = 0, mode = -1, dot = -1;
  for (int i = 0; i < n; ++i) {
    const StencilCall<T>& c = calls[i];
    if (c.rows <= 0) continue;
    StencilArgs<T, T, T> a;
    int csrc = 0;
    bool tile = false;
    if (int e = stencil_args<T, T, T>(c.op, c.rows, c.mode, false, c.dot, c.u, c.su, c.b, c.sb,
                                      c.invd, c.sd, c.y, c.sy, c.omega, c.active, c.partial,
                                      nullptr, 0, a, csrc, tile))
      return e;
    if (!(csrc == kSrcDensity || tile) ||
        (src >= 0 && (csrc != src || c.op.d != d || c.mode != mode || c.dot != dot))) {
      set_error("stencil groups: one tile pass (same source, mode and dofs per node) per "
                "launch");
      return STTOPT_EINVAL;
    }
    src = csrc;
    d = c.op.d;
    mode = c.mode;
    dot = c.dot;
    g.a[g.n] = a;
    g.first[g.n] = total;
    total += a.total;
    ++g.n;
  }
  g.first[g.n] = total;
  if (total == 0) return STTOPT_OK;
  const dim3 grid(total), block(kBlock);
#define STT_G(D, SRC, MODE, DOT) \
  hipLaunchKernelGGL((stencil_tile_groups_kernel<T, T, T, D, SRC, MODE, DOT>), grid, block, 0, s, g)
  const int key = (src == kSrcDensity ? 0 : 1) * 1000 + d * 100 + mode * 10 + dot;
  switch (key) {
    case 0 * 1000 + 2 * 100 + kOperator * 10 + kDotU: STT_G(2, kSrcDensity, kOperator, kDotU); break;
    case 0 * 1000 + 1 * 100 + kOperator * 10 + kDotU: STT_G(1, kSrcDensity, kOperator, kDotU); break;
    case 1 * 1000 + 2 * 100 + kSmoothFirst * 10 + kNoDot: STT_G(2, kSrcParam, kSmoothFirst, kNoDot); break;
    case 1 * 1000 + 2 * 100 + kResid * 10 + kNoDot: STT_G(2, kSrcParam, kResid, kNoDot); break;
    case 1 * 1000 + 2 * 100 + kSmooth * 10 + kNoDot: STT_G(2, kSrcParam, kSmooth, kNoDot); break;
    case 1 * 1000 + 1 * 100 + kSmoothFirst * 10 + kNoDot: STT_G(1, kSrcParam, kSmoothFirst, kNoDot); break;
    case 1 * 1000 + 1 * 100 + kResid * 10 + kNoDot: STT_G(1, kSrcParam, kResid, kNoDot); break;
    case 1 * 1000 + 1 * 100 + kSmooth * 10 + kNoDot: STT_G(1, kSrcParam, kSmooth, kNoDot); break;
    default:
      set_error("stencil groups: unsupported pass");
      return STTOPT_EINVAL;
  }
#undef STT_G
  STT_CHECK_LAUNCH("stencil_tile_groups_kernel");
  return STTOPT_OK;
}
template int stencil_tile_groups<double>(const StencilCall<double>*, int, hipStream_t);
template int stencil_tile_groups<float>(const StencilCall<float>*, int, hipStream_t);

#define STT_STENCIL_INST(T, TB, TY)                                                      \
  template int stencil<T, TB, TY>(const OpDesc&, int, int, bool, int, const T*, long long,  \
                                  const TB*, long long, const T*, long long, TY*, long long,  \
                                  double, const int*, double*, hipStream_t, T*, long long)
STT_STENCIL_INST(double, double, double);
STT_STENCIL_INST(float, float, float);
STT_STENCIL_INST(float, double, double);
STT_STENCIL_INST(float, double, float);
#undef STT_STENCIL_INST

// ------------------------------------------------------------------ diagonal --

template <typename TO, int D, bool KEFF>
__global__ __launch_bounds__(256) void diagonal_kernel(int nelx, int nely, int rows,
                                                       const double* rho, long long sr,
                                                       const double* keff, long long sk,
                                                       const double* rdiag, long long srd,
                                                       const uint32_t* mask, KE<double> ke,
                                                       int out_mode, TO* out, long long so) {
  constexpr int K = 4 * D;
  const int nnx = nelx + 1, nny = nely + 1;
  const long long nnodes = (long long)nnx * nny;
  const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  if (node >= nnodes) return;
  const int r = (int)(node / nnx), c = (int)(node - (long long)r * nnx);
  double acc[D];
#pragma unroll
  for (int i = 0; i < D; ++i) acc[i] = 0.0;
  const long long nel = (long long)nelx * nely;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int er = r - 1 + (q >> 1), ec = c - 1 + (q & 1);
    if (er < 0 || er >= nely || ec < 0 || ec >= nelx) continue;
    const long long e = (long long)er * nelx + ec;
    constexpr int locs[4] = {1, 0, 2, 3};
    const int loc = locs[q];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const int a = loc * D + i;
      if constexpr (KEFF)
        acc[i] += keff[row * sk + (long long)(a * K + a) * nel + e];
      else
        acc[i] += rho[row * sr + e] * ke.v[a * K + a];
    }
  }
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const long long dof = node * D + i;
    double v = acc[i];
    if (rdiag != nullptr) v += rdiag[row * srd + dof];
    const bool fr = mask_free(mask, dof);
    if (out_mode == 1 && !fr) v = 1.0;
    if (out_mode == 2) v = fr ? 1.0 / v : 0.0;
    out[row * so + dof] = static_cast<TO>(v);
  }
}

template <typename TO>
int diagonal(const OpDesc& op, int rows, int out_mode, TO* out, long long so,
             hipStream_t s) {
  if (rows <= 0) return STTOPT_OK;
  dim3 grid(ceil_div(op.nnodes(), 256), rows), block(256);
  const KE<double> ke = make_ke<double>(op.ke, op.d);
  const double* rho = static_cast<const double*>(op.rho);
  const double* keff = static_cast<const double*>(op.keff);
  const double* rd = static_cast<const double*>(op.rdiag);
#define STT_DIAG(D, KF)                                                                 \
  hipLaunchKernelGGL((diagonal_kernel<TO, D, KF>), grid, block, 0, s, op.nelx, op.nely, \
                     rows, rho, op.sr, keff, op.sk, rd, op.srd, op.mask, ke, out_mode,  \
                     out, so)
  if (op.d == 2) {
    if (keff) STT_DIAG(2, true); else STT_DIAG(2, false);
  } else {
    if (keff) STT_DIAG(1, true); else STT_DIAG(1, false);
  }
#undef STT_DIAG
  STT_CHECK_LAUNCH("diagonal_kernel");
  return STTOPT_OK;
}
template int diagonal<double>(const OpDesc&, int, int, double*, long long, hipStream_t);
template int diagonal<float>(const OpDesc&, int, int, float*, long long, hipStream_t);

// ---------------------------------------------------------- lumped reaction --

template <typename T, int D>
__global__ __launch_bounds__(256) void lumped_kernel(int nelx, int nely, const T* react,
                                                     long long sr, T* out, long long so) {
  const int nnx = nelx + 1, nny = nely + 1;
  const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  if (node >= (long long)nnx * nny) return;
  const int r = (int)(node / nnx), c = (int)(node - (long long)r * nnx);
  T acc = T(0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int er = r - 1 + (q >> 1), ec = c - 1 + (q & 1);
    if (er < 0 || er >= nely || ec < 0 || ec >= nelx) continue;
    acc += react[row * sr + (long long)er * nelx + ec] / T(4);
  }
#pragma unroll
  for (int i = 0; i < D; ++i) out[row * so + node * D + i] = acc;
}

// ------------------------------------------------------------- elementwise --

// x = w D^-1 b at free dofs (0 at fixed), optionally with b copied into the V-cycle
// precision; V consecutive dofs per thread (V = 2 when every array allows pairs).
template <typename T, typename TB, int V>
__global__ __launch_bounds__(256) void jacobi_first_kernel(long long n, const TB* __restrict__ b,
                                                           long long sb,
                                                           const T* __restrict__ invd,
                                                           long long sd, T omega,
                                                           const uint32_t* __restrict__ mask,
                                                           T* __restrict__ x, long long sx,
                                                           T* __restrict__ bcopy, long long sbc,
                                                           const int* __restrict__ active) {
  const int row = blockIdx.y;
  if (active != nullptr && active[row] == 0) return;
  const long long nv = n / V;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += (long long)gridDim.x * blockDim.x) {
    const uint32_t word = mask != nullptr ? mask[(i * V) >> 5] : ~0u;
    const Vec<TB, V> bv = load_node<TB, V>(b + row * sb, i);
    const Vec<T, V> dv = load_node<T, V>(invd + row * sd, i);
    Vec<T, V> cv, xv;
#pragma unroll
    for (int k = 0; k < V; ++k) {
      cv.v[k] = static_cast<T>(bv.v[k]);
      xv.v[k] = ((word >> ((i * V + k) & 31)) & 1u) ? omega * dv.v[k] * cv.v[k] : T(0);
    }
    if (bcopy != nullptr) store_node<T, V>(bcopy + row * sbc, i, cv);
    store_node<T, V>(x + row * sx, i, xv);
  }
}

static inline dim3 elem_grid(long long n, int rows) {
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  return dim3((unsigned)blocks, rows);
}

template <typename T, typename TB>
int jacobi_first(long long n, int rows, const TB* b, long long sb, const T* invd,
                 long long sd, double omega, const uint32_t* mask, T* x, long long sx,
                 T* bcopy, long long sbc, const int* active, hipStream_t s) {
  if (rows <= 0 || n <= 0) return STTOPT_OK;
  const bool pairs = n % 2 == 0 && pair_aligned(b, sb, sizeof(TB)) &&
                     pair_aligned(invd, sd, sizeof(T)) && pair_aligned(x, sx, sizeof(T)) &&
                     pair_aligned(bcopy, sbc, sizeof(T));
  if (pairs)
    hipLaunchKernelGGL((jacobi_first_kernel<T, TB, 2>), elem_grid(n / 2, rows), dim3(256), 0,
                       s, n, b, sb, invd, sd, static_cast<T>(omega), mask, x, sx, bcopy, sbc,
                       active);
  else
    hipLaunchKernelGGL((jacobi_first_kernel<T, TB, 1>), elem_grid(n, rows), dim3(256), 0, s,
                       n, b, sb, invd, sd, static_cast<T>(omega), mask, x, sx, bcopy, sbc,
                       active);
  STT_CHECK_LAUNCH("jacobi_first_kernel");
  return STTOPT_OK;
}
template int jacobi_first<double, double>(long long, int, const double*, long long,
                                          const double*, long long, double,
                                          const uint32_t*, double*, long long, double*,
                                          long long, const int*, hipStream_t);
template int jacobi_first<float, float>(long long, int, const float*, long long,
                                        const float*, long long, double, const uint32_t*,
                                        float*, long long, float*, long long, const int*,
                                        hipStream_t);
template int jacobi_first<float, double>(long long, int, const double*, long long,
                                         const float*, long long, double, const uint32_t*,
                                         float*, long long, float*, long long, const int*,
                                         hipStream_t);

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void convert_kernel(long long n, const TI* in, long long si,
                                                      TO* out, long long so,
                                                      const int* active) {
  const int row = blockIdx.y;
  if (active != nullptr && active[row] == 0) return;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[row * so + i] = static_cast<TO>(in[row * si + i]);
}

template <typename TI, typename TO>
int convert(long long n, int rows, const TI* in, long long si, TO* out, long long so,
            const int* active, hipStream_t s) {
  if (rows <= 0 || n <= 0) return STTOPT_OK;
  hipLaunchKernelGGL((convert_kernel<TI, TO>), elem_grid(n, rows), dim3(256), 0, s, n, in,
                     si, out, so, active);
  STT_CHECK_LAUNCH("convert_kernel");
  return STTOPT_OK;
}
template int convert<double, double>(long long, int, const double*, long long, double*,
                                     long long, const int*, hipStream_t);
template int convert<float, double>(long long, int, const float*, long long, double*,
                                    long long, const int*, hipStream_t);
template int convert<double, float>(long long, int, const double*, long long, float*,
                                    long long, const int*, hipStream_t);
template int convert<float, float>(long long, int, const float*, long long, float*,
                                   long long, const int*, hipStream_t);

template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(long long n, T value, T* x, long long sx,
                                                   const int* active) {
  const int row = blockIdx.y;
  if (active != nullptr && active[row] == 0) return;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    x[row * sx + i] = value;
}

template <typename T>
int fill(long long n, int rows, T value, T* x, long long sx, const int* active,
         hipStream_t s) {
  if (rows <= 0 || n <= 0) return STTOPT_OK;
  hipLaunchKernelGGL((fill_kernel<T>), elem_grid(n, rows), dim3(256), 0, s, n, value, x, sx,
                     active);
  STT_CHECK_LAUNCH("fill_kernel");
  return STTOPT_OK;
}
template int fill<double>(long long, int, double, double*, long long, const int*,
                          hipStream_t);
template int fill<float>(long long, int, float, float*, long long, const int*, hipStream_t);
template int fill<int>(long long, int, int, int*, long long, const int*, hipStream_t);

template <typename T>
__global__ __launch_bounds__(256) void axpy_kernel(long long n, const T* xin, long long si,
                                                   T* y, long long sy, const int* active) {
  const int row = blockIdx.y;
  if (active != nullptr && active[row] == 0) return;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[row * sy + i] += xin[row * si + i];
}

template <typename T>
int axpy(long long n, int rows, const T* xin, long long si, T* y, long long sy,
         const int* active, hipStream_t s) {
  if (rows <= 0 || n <= 0) return STTOPT_OK;
  hipLaunchKernelGGL((axpy_kernel<T>), elem_grid(n, rows), dim3(256), 0, s, n, xin, si, y,
                     sy, active);
  STT_CHECK_LAUNCH("axpy_kernel");
  return STTOPT_OK;
}
template int axpy<double>(long long, int, const double*, long long, double*, long long,
                          const int*, hipStream_t);
template int axpy<float>(long long, int, const float*, long long, float*, long long,
                         const int*, hipStream_t);

__global__ __launch_bounds__(256) void project_kernel(long long n, const uint32_t* mask,
                                                      double* x, long long sx) {
  const int row = blockIdx.y;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    if (!mask_free(mask, i)) x[row * sx + i] = 0.0;
}

int project_f64(long long n, int rows, const uint32_t* mask, double* x, long long sx,
                hipStream_t s) {
  if (rows <= 0 || n <= 0 || mask == nullptr) return STTOPT_OK;
  hipLaunchKernelGGL(project_kernel, elem_grid(n, rows), dim3(256), 0, s, n, mask, x, sx);
  STT_CHECK_LAUNCH("project_kernel");
  return STTOPT_OK;
}

// ---------------------------------------------------------------- masks ------

__global__ __launch_bounds__(256) void pack_mask_kernel(const uint8_t* m, long long n,
                                                        uint32_t* bits) {
  const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long words = (n + 31) >> 5;
  if (w >= words) return;
  uint32_t v = 0;
  for (int k = 0; k < 32; ++k) {
    const long long i = (w << 5) + k;
    if (i < n && m[i] != 0) v |= (1u << k);
  }
  bits[w] = v;
}

__global__ __launch_bounds__(256) void coarsen_mask_kernel(const uint32_t* fine, int nelx,
                                                           int d, int kx, int ky,
                                                           int cnnx, int cnny,
                                                           uint32_t* coarse) {
  const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)cnnx * cnny * d;
  const long long words = (n + 31) >> 5;
  if (w >= words) return;
  uint32_t v = 0;
  for (int k = 0; k < 32; ++k) {
    const long long i = (w << 5) + k;
    if (i >= n) break;
    const long long node = i / d;
    const int comp = (int)(i - node * d);
    const int R = (int)(node / cnnx), C = (int)(node - (long long)R * cnnx);
    const long long fnode = (long long)(R * ky) * (nelx + 1) + (long long)C * kx;
    if (mask_free(fine, fnode * d + comp)) v |= (1u << k);
  }
  coarse[w] = v;
}

}  // namespace stt

// ===================================================================== C ABI ==

using namespace stt;

extern "C" {

const char* sttopt_last_error(void) { return g_last_error.c_str(); }

int sttopt_version(void) { return 1; }

int sttopt_abi_sizes(int64_t* level_bytes, int64_t* mg_bytes) {
  if (level_bytes) *level_bytes = (int64_t)sizeof(sttopt_level);
  if (mg_bytes) *mg_bytes = (int64_t)sizeof(sttopt_mg);
  return STTOPT_OK;
}

int sttopt_pack_mask(const uint8_t* mask, int64_t ndof, uint32_t* bits, void* stream) {
  if (mask == nullptr || bits == nullptr || ndof < 0) {
    set_error("sttopt_pack_mask: null pointer or negative size");
    return STTOPT_EINVAL;
  }
  const long long words = (ndof + 31) >> 5;
  if (words == 0) return STTOPT_OK;
  hipLaunchKernelGGL(pack_mask_kernel, dim3(ceil_div(words, 256)), dim3(256), 0,
                     (hipStream_t)stream, mask, (long long)ndof, bits);
  STT_CHECK_LAUNCH("pack_mask_kernel");
  return STTOPT_OK;
}

int sttopt_coarsen_mask(const uint32_t* fine_bits, int32_t nelx, int32_t nely, int32_t d,
                        int32_t kx, int32_t ky, uint32_t* coarse_bits, void* stream) {
  if (fine_bits == nullptr || coarse_bits == nullptr || kx < 1 || ky < 1 ||
      nelx % kx != 0 || nely % ky != 0 || (d != 1 && d != 2)) {
    set_error("sttopt_coarsen_mask: invalid arguments");
    return STTOPT_EINVAL;
  }
  const int cnnx = nelx / kx + 1, cnny = nely / ky + 1;
  const long long words = ((long long)cnnx * cnny * d + 31) >> 5;
  hipLaunchKernelGGL(coarsen_mask_kernel, dim3(ceil_div(words, 256)), dim3(256), 0,
                     (hipStream_t)stream, fine_bits, nelx, d, kx, ky, cnnx, cnny,
                     coarse_bits);
  STT_CHECK_LAUNCH("coarsen_mask_kernel");
  return STTOPT_OK;
}

static int check_grid(int32_t d, int32_t nelx, int32_t nely, int32_t rows, const char* who) {
  if ((d != 1 && d != 2) || nelx <= 0 || nely <= 0 || rows < 0) {
    set_error(std::string(who) + ": invalid grid (d, nelx, nely, rows)");
    return STTOPT_EINVAL;
  }
  return STTOPT_OK;
}

int sttopt_apply(int32_t dtype, int32_t mode, int32_t d, int32_t nelx, int32_t nely,
                 int32_t rows, const void* u, int64_t u_stride, const void* density,
                 int64_t density_stride, const double* ke, const uint32_t* mask,
                 const void* rdiag, int64_t rdiag_stride, void* y, int64_t y_stride,
                 void* stream) {
  if (int st = check_grid(d, nelx, nely, rows, "sttopt_apply")) return st;
  if (u == nullptr || density == nullptr || ke == nullptr || y == nullptr ||
      (mode != STTOPT_MODE_MATVEC && mode != STTOPT_MODE_OPERATOR &&
       mode != STTOPT_MODE_OPERATOR_PROJECTED)) {
    set_error("sttopt_apply: null pointer or bad mode");
    return STTOPT_EINVAL;
  }
  const bool opmode = mode != STTOPT_MODE_MATVEC;
  OpDesc op;
  op.d = d;
  op.nelx = nelx;
  op.nely = nely;
  op.rho = density;
  op.sr = density_stride;
  op.rdiag = opmode ? rdiag : nullptr;
  op.srd = rdiag_stride;
  op.mask = opmode ? mask : nullptr;
  for (int i = 0; i < 16 * d * d; ++i) op.ke[i] = ke[i];
  const int m = opmode ? kOperator : kMatvec;
  // the input projection is skipped when the caller guarantees a projected u
  const bool proj = mode == STTOPT_MODE_OPERATOR && mask != nullptr;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == STTOPT_F64)
    return stencil<double, double, double>(op, rows, m, proj, kNoDot,
                                           static_cast<const double*>(u), u_stride, nullptr,
                                           0, nullptr, 0, static_cast<double*>(y), y_stride,
                                           0.0, nullptr, nullptr, s);
  if (dtype == STTOPT_F32)
    return stencil<float, float, float>(op, rows, m, proj, kNoDot,
                                        static_cast<const float*>(u), u_stride, nullptr, 0,
                                        nullptr, 0, static_cast<float*>(y), y_stride, 0.0,
                                        nullptr, nullptr, s);
  set_error("sttopt_apply: dtype must be STTOPT_F64 or STTOPT_F32");
  return STTOPT_EINVAL;
}

int sttopt_apply_keff(int32_t dtype, int32_t d, int32_t nelx, int32_t nely, int32_t rows,
                      const void* u, int64_t u_stride, const void* keff, int64_t keff_stride,
                      const uint32_t* mask, void* y, int64_t y_stride, void* stream) {
  if (int st = check_grid(d, nelx, nely, rows, "sttopt_apply_keff")) return st;
  if (u == nullptr || keff == nullptr || y == nullptr) {
    set_error("sttopt_apply_keff: null pointer");
    return STTOPT_EINVAL;
  }
  OpDesc op;
  op.d = d;
  op.nelx = nelx;
  op.nely = nely;
  op.keff = keff;
  op.sk = keff_stride;
  op.mask = mask;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == STTOPT_F64)
    return stencil<double, double, double>(op, rows, kOperator, mask != nullptr, kNoDot,
                                           static_cast<const double*>(u), u_stride, nullptr,
                                           0, nullptr, 0, static_cast<double*>(y), y_stride,
                                           0.0, nullptr, nullptr, s);
  if (dtype == STTOPT_F32)
    return stencil<float, float, float>(op, rows, kOperator, mask != nullptr, kNoDot,
                                        static_cast<const float*>(u), u_stride, nullptr, 0,
                                        nullptr, 0, static_cast<float*>(y), y_stride, 0.0,
                                        nullptr, nullptr, s);
  set_error("sttopt_apply_keff: dtype must be STTOPT_F64 or STTOPT_F32");
  return STTOPT_EINVAL;
}

int sttopt_diagonal(int32_t dtype, int32_t d, int32_t nelx, int32_t nely, int32_t rows,
                    const void* density, int64_t density_stride, const double* ke,
                    const uint32_t* mask, const void* rdiag, int64_t rdiag_stride, void* out,
                    int64_t out_stride, void* stream) {
  if (int st = check_grid(d, nelx, nely, rows, "sttopt_diagonal")) return st;
  if (density == nullptr || ke == nullptr || out == nullptr || dtype != STTOPT_F64) {
    set_error("sttopt_diagonal: null pointer or dtype other than fp64");
    return STTOPT_EINVAL;
  }
  OpDesc op;
  op.d = d;
  op.nelx = nelx;
  op.nely = nely;
  op.rho = density;
  op.sr = density_stride;
  op.rdiag = rdiag;
  op.srd = rdiag_stride;
  op.mask = mask;
  for (int i = 0; i < 16 * d * d; ++i) op.ke[i] = ke[i];
  return diagonal<double>(op, rows, mask != nullptr ? 1 : 0, static_cast<double*>(out),
                          out_stride, (hipStream_t)stream);
}

int sttopt_lumped_reaction(int32_t dtype, int32_t d, int32_t nelx, int32_t nely,
                           int32_t rows, const void* reaction, int64_t reaction_stride,
                           void* out, int64_t out_stride, void* stream) {
  if (int st = check_grid(d, nelx, nely, rows, "sttopt_lumped_reaction")) return st;
  if (reaction == nullptr || out == nullptr || rows == 0) {
    if (rows == 0) return STTOPT_OK;
    set_error("sttopt_lumped_reaction: null pointer");
    return STTOPT_EINVAL;
  }
  const long long nnodes = (long long)(nelx + 1) * (nely + 1);
  dim3 grid(ceil_div(nnodes, 256), rows), block(256);
  hipStream_t s = (hipStream_t)stream;
  if (dtype == STTOPT_F64) {
    if (d == 2)
      hipLaunchKernelGGL((lumped_kernel<double, 2>), grid, block, 0, s, nelx, nely,
                         static_cast<const double*>(reaction), reaction_stride,
                         static_cast<double*>(out), out_stride);
    else
      hipLaunchKernelGGL((lumped_kernel<double, 1>), grid, block, 0, s, nelx, nely,
                         static_cast<const double*>(reaction), reaction_stride,
                         static_cast<double*>(out), out_stride);
  } else if (dtype == STTOPT_F32) {
    if (d == 2)
      hipLaunchKernelGGL((lumped_kernel<float, 2>), grid, block, 0, s, nelx, nely,
                         static_cast<const float*>(reaction), reaction_stride,
                         static_cast<float*>(out), out_stride);
    else
      hipLaunchKernelGGL((lumped_kernel<float, 1>), grid, block, 0, s, nelx, nely,
                         static_cast<const float*>(reaction), reaction_stride,
                         static_cast<float*>(out), out_stride);
  } else {
    set_error("sttopt_lumped_reaction: bad dtype");
    return STTOPT_EINVAL;
  }
  STT_CHECK_LAUNCH("lumped_kernel");
  return STTOPT_OK;
}

}  // extern "C"
