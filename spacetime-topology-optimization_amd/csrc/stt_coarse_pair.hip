// Coarse levels of the fp32 V-cycle in two passes per level (torch_mg.py:440-471 with
// n_smooth = 2, gamma = 1, 2 x 2 coarsening), on the assembled node stencil
// (STTOPT_KEFF_NODE):
//
//   pre:  x1 = w D^-1 b,  x2 = x1 + w D^-1 (b - A x1),  r = b - A x2,  b_c = P_c R r
//   post: xp = x2 + P_f (P e),  x3 = xp + w D^-1 (b - A xp),  x4 = x3 + w D^-1 (b - A x3)
//
// where the separate path takes six to seven launches per level (first sweeps,
// residual, restriction, prolongation, two sweeps).  A workgroup stages what its tile
// needs in LDS and evaluates the stencil on shrinking regions: the pre pass on the fine
// nodes that feed a 32 x 4 tile of coarse nodes (65 x 9, with x2 on a one-node ring
// around them and x1 on two), the post pass on a 64 x 8 tile of fine nodes (x3 on a
// one-node ring, xp on two).  The intermediate fields never reach HBM; the pre pass
// writes x2 for its own fine nodes and the restricted residual, the post pass writes x4.
// Each thread loads the stencil rows of its nodes once, before anything else, and
// applies them in both phases that need them.  Every value is formed with the operands
// and in the order of the separate kernels
// (stencil_kernel's node-stencil sum, jacobi_first's first sweep, restrict22_kernel,
// prolong_add_kernel), so the results are the same.
#include "stt_kernels.h"

namespace stt {

namespace {

// Tile configurations (the autotuned schedule of a level): the pre pass covers
// CW x CH coarse nodes (staging 2 CW + 5 by 2 CH + 5 fine nodes), the post pass TW x TH
// fine nodes (staging TW + 4 by TH + 4), with NT threads per workgroup.  Every phase
// takes at most two nodes per thread, whose loads are issued together (the phases are
// latency bound: each reads the stencil rows of its nodes from L2).
template <int CW_, int CH_, int TW_, int TH_, int NT_>
struct PairCfg {
  static constexpr int CW = CW_, CH = CH_, TW = TW_, TH = TH_, NT = NT_;
  static constexpr int PW = 2 * CW + 5, PH = 2 * CH + 5;      // pre: staged fine region
  static constexpr int QW = TW + 4, QH = TH + 4;              // post: staged fine region
  static constexpr int EW = TW / 2 + 3, EH = TH / 2 + 3;      // post: staged coarse region
  template <int N>
  static constexpr int per_thread() { return (N + NT - 1) / NT; }
};
using PairWide = PairCfg<32, 4, 64, 8, 512>;    // variant 0
using PairNarrow = PairCfg<16, 4, 32, 8, 256>;  // variant 1: twice the workgroups

// What one node's stencil row needs from HBM: its own record (own block's upper
// triangle, forward blocks E, NW, N, NE), the backward neighbours' forward blocks toward
// it, which of the eight neighbours exist, and the node's free-dof bits.  Loaded once
// per node and kept in registers across the phases that apply the same row.
template <int D>
struct NodeRow {
  static constexpr int NS = node_stencil_stride(D), B = D * D;
  float own[NS];
  float back[4][B];
  bool fwd_in[4], back_in[4];
  uint32_t free_bits;
};

constexpr int kFdr[4] = {0, 1, 1, 1}, kFdc[4] = {1, -1, 0, 1};

template <int D>
__device__ __forceinline__ void load_row(const float* S, const uint32_t* mask, int nnx,
                                         int nny, int r, int c, NodeRow<D>& w) {
  constexpr int NS = NodeRow<D>::NS, ND = 4, B = NodeRow<D>::B;
  const long long node = (long long)r * nnx + c;
  load_record<float, NS>(S + node * NS, w.own);
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const int cc = c + kFdc[f];
    w.fwd_in[f] = r + kFdr[f] < nny && cc >= 0 && cc < nnx;
    const int rr = r - kFdr[f], cb = c - kFdc[f];
    w.back_in[f] = rr >= 0 && cb >= 0 && cb < nnx;
    if (w.back_in[f]) {
      load_record<float, B>(S + ((long long)rr * nnx + cb) * NS + ND + f * B, w.back[f]);
    } else {
#pragma unroll
      for (int q = 0; q < B; ++q) w.back[f][q] = 0.0f;
    }
  }
  const uint32_t word = mask[(node * D) >> 5];
  w.free_bits = (word >> ((node * D) & 31)) & ((1u << D) - 1u);
}

// acc = (A u)_node: the sum of stencil_kernel<..., kSrcNode> in the same order.  U: the
// 3 x 3 neighbourhood, zero outside the grid.
template <int D>
__device__ __forceinline__ void apply_row(const NodeRow<D>& w, const float (&U)[3][3][D],
                                          float (&acc)[D]) {
  constexpr int ND = 4;
#pragma unroll
  for (int i = 0; i < D; ++i) acc[i] = 0.0f;
  int idx = 0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
#pragma unroll
    for (int k = i; k < D; ++k) {
      const float v = w.own[idx++];
      acc[i] += v * U[1][1][k];
      if (k != i) acc[k] += v * U[1][1][i];
    }
  }
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    if (w.fwd_in[f]) {
#pragma unroll
      for (int i = 0; i < D; ++i) {
#pragma unroll
        for (int k = 0; k < D; ++k)
          acc[i] += w.own[ND + (f * D + i) * D + k] * U[1 + kFdr[f]][1 + kFdc[f]][k];
      }
    }
  }
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    if (w.back_in[f]) {
#pragma unroll
      for (int i = 0; i < D; ++i) {
#pragma unroll
        for (int k = 0; k < D; ++k)
          acc[i] += w.back[f][k * D + i] * U[1 - kFdr[f]][1 - kFdc[f]][k];
      }
    }
  }
}

// The 3 x 3 neighbourhood of local node (lr, lc) of a staged field of width W.
template <typename T, int D, int W>
__device__ __forceinline__ void gather3(const T (*f)[D], int lr, int lc, T (&U)[3][3][D]) {
#pragma unroll
  for (int dr = 0; dr < 3; ++dr) {
#pragma unroll
    for (int dc = 0; dc < 3; ++dc) {
#pragma unroll
      for (int k = 0; k < D; ++k) U[dr][dc][k] = f[(lr + dr - 1) * W + lc + dc - 1][k];
    }
  }
}

__device__ __forceinline__ bool dof_free(const uint32_t* mask, long long dof) {
  return (mask[dof >> 5] >> (dof & 31)) & 1u;
}

struct CoarsePairArgs {
  int nnx, nny;          // fine grid nodes of this level
  int cnnx, cnny;        // nodes of the next coarser level (2 x 2)
  int tiles_x, tiles, total;
  const float* S;        // node-stencil records of this level
  long long sS;
  const uint32_t* mask;  // this level's free dofs
  const uint32_t* cmask; // the coarser level's free dofs (pre)
  const float* b;        // this level's right-hand side
  long long sb;
  const float* invd;
  long long sd;
  const float* x2;       // post: the pre-smoothed x
  long long sx2;
  const float* ec;       // post: the coarse correction
  long long sec;
  float* out;            // pre: x2; post: x4
  long long so;
  float* bc;             // pre: the coarser level's right-hand side
  long long sbc;
  float omega;
  const int* active;
};

// The pre pass of tile tt of batch row `row` (the whole workgroup).
template <int D, typename C>
__device__ __forceinline__ void pair_pre_tile(const CoarsePairArgs& a, int row, int tt) {
  __shared__ float sbv[C::PH * C::PW][D];
  __shared__ float sdv[C::PH * C::PW][D];
  __shared__ float sx1[C::PH * C::PW][D];
  __shared__ float sx2[C::PH * C::PW][D];
  __shared__ float srr[C::PH * C::PW][D];
  if (a.active != nullptr && a.active[row] == 0) return;  // whole block
  const int ty = tt / a.tiles_x, tx = tt - ty * a.tiles_x;
  const int I0 = ty * C::CH, J0 = tx * C::CW;     // first coarse node of the tile
  // fine node at local (0, 0): the residual is needed on fine rows 2 I0 - 1 ..
  // 2 (I0 + C::CH) - 1 (local 2 .. C::PH - 3), x2 one node beyond, x1 two
  const int fr0 = 2 * I0 - 3, fc0 = 2 * J0 - 3;
  const float* b = a.b + row * a.sb;
  const float* dinv = a.invd + row * a.sd;
  const float* S = a.S + row * a.sS;
  // The stencil rows of this thread's nodes one node inside the staged region (phase
  // x2; the residual's nodes are a subset, taken by the same threads), loaded before
  // anything else so that every HBM read of the workgroup is in flight at once.
  constexpr int NJ = C::template per_thread<C::PH * C::PW>();
  NodeRow<D> w[NJ];
  bool inner[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int i = threadIdx.x + j * C::NT;
    const int lr = i / C::PW, lc = i % C::PW;
    const int fr = fr0 + lr, fc = fc0 + lc;
    inner[j] = i < C::PH * C::PW && lr >= 1 && lr <= C::PH - 2 && lc >= 1 && lc <= C::PW - 2 &&
               fr >= 0 && fr < a.nny && fc >= 0 && fc < a.nnx;
    if (inner[j]) load_row<D>(S, a.mask, a.nnx, a.nny, fr, fc, w[j]);
  }
  // x1 = w D^-1 b on the whole staged region (jacobi_first's expression; 0 outside)
#pragma unroll
  for (int j = 0; j < C::template per_thread<C::PH * C::PW>(); ++j) {
    const int i = threadIdx.x + j * C::NT;
    if (i >= C::PH * C::PW) break;
    const int fr = fr0 + i / C::PW, fc = fc0 + i % C::PW;
    Vec<float, D> bv, dv;
#pragma unroll
    for (int k = 0; k < D; ++k) bv.v[k] = dv.v[k] = 0.0f;
    if (fr >= 0 && fr < a.nny && fc >= 0 && fc < a.nnx) {
      const long long node = (long long)fr * a.nnx + fc;
      bv = load_node<float, D>(b, node);
      dv = load_node<float, D>(dinv, node);
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      sbv[i][k] = bv.v[k];
      sdv[i][k] = dv.v[k];
      sx1[i][k] = a.omega * dv.v[k] * bv.v[k];
    }
  }
  __syncthreads();
  // x2 = x1 + w D^-1 (b - A x1) one node inside the staged region; x2 of the tile's own
  // fine nodes (those whose coarse node, by integer halving, is in the tile) to HBM
#pragma unroll
  for (int j = 0; j < C::template per_thread<C::PH * C::PW>(); ++j) {
    const int i = threadIdx.x + j * C::NT;
    if (i >= C::PH * C::PW) break;
    const int lr = i / C::PW, lc = i % C::PW;
    if (lr < 1 || lr > C::PH - 2 || lc < 1 || lc > C::PW - 2) continue;
    const int fr = fr0 + lr, fc = fc0 + lc;
    Vec<float, D> xo;
#pragma unroll
    for (int k = 0; k < D; ++k) xo.v[k] = 0.0f;
    if (inner[j]) {
      float U[3][3][D], acc[D];
      gather3<float, D, C::PW>(sx1, lr, lc, U);
      apply_row<D>(w[j], U, acc);
      const long long node = (long long)fr * a.nnx + fc;
#pragma unroll
      for (int k = 0; k < D; ++k)
        xo.v[k] = ((w[j].free_bits >> k) & 1u)
                      ? U[1][1][k] + a.omega * sdv[i][k] * (sbv[i][k] - acc[k])
                      : 0.0f;
      if (fr >= 2 * I0 && fr < 2 * I0 + 2 * C::CH && fc >= 2 * J0 && fc < 2 * J0 + 2 * C::CW)
        store_node<float, D>(a.out + row * a.so, node, xo);
    }
#pragma unroll
    for (int k = 0; k < D; ++k) sx2[i][k] = xo.v[k];
  }
  __syncthreads();
  // r = b - A x2 on the fine nodes that feed the tile's coarse nodes
#pragma unroll
  for (int j = 0; j < C::template per_thread<C::PH * C::PW>(); ++j) {
    const int i = threadIdx.x + j * C::NT;
    if (i >= C::PH * C::PW) break;
    const int lr = i / C::PW, lc = i % C::PW;
    if (lr < 2 || lr > C::PH - 3 || lc < 2 || lc > C::PW - 3) continue;
    const int fr = fr0 + lr, fc = fc0 + lc;
    float ro[D];
#pragma unroll
    for (int k = 0; k < D; ++k) ro[k] = 0.0f;
    if (inner[j]) {
      float U[3][3][D], acc[D];
      gather3<float, D, C::PW>(sx2, lr, lc, U);
      apply_row<D>(w[j], U, acc);
#pragma unroll
      for (int k = 0; k < D; ++k) ro[k] = ((w[j].free_bits >> k) & 1u) ? sbv[i][k] - acc[k] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) srr[i][k] = ro[k];
  }
  __syncthreads();
  // full weighting to the tile's coarse nodes, in restrict22_kernel's order
  for (int i = threadIdx.x; i < C::CW * C::CH; i += C::NT) {
    const int I = I0 + i / C::CW, J = J0 + i % C::CW;
    if (I >= a.cnny || J >= a.cnnx) continue;
    float acc[D];
#pragma unroll
    for (int k = 0; k < D; ++k) acc[k] = 0.0f;
    const float wh = 1.0f / 2.0f, w1 = 2.0f / 2.0f;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      const int R = 2 * I + dy;
      if (R < 0 || R >= a.nny) continue;
      const float wy = float(2 - (dy < 0 ? -dy : dy)) / 2.0f;
      const int lr = R - fr0, lc = 2 * J - fc0;
      float racc[D];
#pragma unroll
      for (int k = 0; k < D; ++k) racc[k] = 0.0f;
      if (J > 0) {
#pragma unroll
        for (int k = 0; k < D; ++k) racc[k] += wh * srr[lr * C::PW + lc - 1][k];
      }
#pragma unroll
      for (int k = 0; k < D; ++k) racc[k] += w1 * srr[lr * C::PW + lc][k];
      if (2 * J + 1 < a.nnx) {
#pragma unroll
        for (int k = 0; k < D; ++k) racc[k] += wh * srr[lr * C::PW + lc + 1][k];
      }
#pragma unroll
      for (int k = 0; k < D; ++k) acc[k] += wy * racc[k];
    }
    const long long cnode = (long long)I * a.cnnx + J;
    Vec<float, D> o;
#pragma unroll
    for (int k = 0; k < D; ++k) o.v[k] = dof_free(a.cmask, cnode * D + k) ? acc[k] : 0.0f;
    store_node<float, D>(a.bc + row * a.sbc, cnode, o);
  }
}

// The post pass of tile tt of batch row `row` (the whole workgroup).
template <int D, typename C>
__device__ __forceinline__ void pair_post_tile(const CoarsePairArgs& a, int row, int tt) {
  __shared__ float sbv[C::QH * C::QW][D];
  __shared__ float sdv[C::QH * C::QW][D];
  __shared__ float sxp[C::QH * C::QW][D];
  __shared__ float sx3[C::QH * C::QW][D];
  __shared__ float sec[C::EH * C::EW][D];
  if (a.active != nullptr && a.active[row] == 0) return;  // whole block
  const int ty = tt / a.tiles_x, tx = tt - ty * a.tiles_x;
  const int r0 = ty * C::TH, c0 = tx * C::TW;          // first output node of the tile
  const int fr0 = r0 - 2, fc0 = c0 - 2;            // fine node at local (0, 0)
  const int er0 = r0 / 2 - 1, ec0 = c0 / 2 - 1;    // coarse node at local (0, 0)
  const float* b = a.b + row * a.sb;
  const float* dinv = a.invd + row * a.sd;
  const float* x2 = a.x2 + row * a.sx2;
  const float* e = a.ec + row * a.sec;
  const float* S = a.S + row * a.sS;
  // the stencil rows of this thread's nodes one node around the tile (phase x3; the
  // tile's own nodes, phase x4, are a subset taken by the same threads)
  constexpr int NJ = C::template per_thread<C::QH * C::QW>();
  NodeRow<D> w[NJ];
  bool inner[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int i = threadIdx.x + j * C::NT;
    const int lr = i / C::QW, lc = i % C::QW;
    const int fr = fr0 + lr, fc = fc0 + lc;
    inner[j] = i < C::QH * C::QW && lr >= 1 && lr <= C::QH - 2 && lc >= 1 && lc <= C::QW - 2 &&
               fr >= 0 && fr < a.nny && fc >= 0 && fc < a.nnx;
    if (inner[j]) load_row<D>(S, a.mask, a.nnx, a.nny, fr, fc, w[j]);
  }
  for (int i = threadIdx.x; i < C::EH * C::EW; i += C::NT) {
    const int cr = er0 + i / C::EW, cc = ec0 + i % C::EW;
    Vec<float, D> v;
#pragma unroll
    for (int k = 0; k < D; ++k) v.v[k] = 0.0f;
    if (cr >= 0 && cr < a.cnny && cc >= 0 && cc < a.cnnx)
      v = load_node<float, D>(e, (long long)cr * a.cnnx + cc);
#pragma unroll
    for (int k = 0; k < D; ++k) sec[i][k] = v.v[k];
  }
  Vec<float, D> xv[C::template per_thread<C::QH * C::QW>()];
#pragma unroll
  for (int j = 0; j < C::template per_thread<C::QH * C::QW>(); ++j) {
    const int i = threadIdx.x + j * C::NT;
    if (i >= C::QH * C::QW) break;
    const int fr = fr0 + i / C::QW, fc = fc0 + i % C::QW;
    Vec<float, D> bv, dv;
#pragma unroll
    for (int k = 0; k < D; ++k) bv.v[k] = dv.v[k] = xv[j].v[k] = 0.0f;
    if (fr >= 0 && fr < a.nny && fc >= 0 && fc < a.nnx) {
      const long long node = (long long)fr * a.nnx + fc;
      bv = load_node<float, D>(b, node);
      dv = load_node<float, D>(dinv, node);
      xv[j] = load_node<float, D>(x2, node);
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      sbv[i][k] = bv.v[k];
      sdv[i][k] = dv.v[k];
    }
  }
  __syncthreads();
  // xp = x2 + P e at the free dofs (prolong_add_kernel's expression)
#pragma unroll
  for (int j = 0; j < C::template per_thread<C::QH * C::QW>(); ++j) {
    const int i = threadIdx.x + j * C::NT;
    if (i >= C::QH * C::QW) break;
    const int fr = fr0 + i / C::QW, fc = fc0 + i % C::QW;
    Vec<float, D> v = xv[j];
    if (fr >= 0 && fr < a.nny && fc >= 0 && fc < a.nnx) {
      const long long node = (long long)fr * a.nnx + fc;
      const uint32_t word = a.mask[(node * D) >> 5];
      const uint32_t bits = (word >> ((node * D) & 31)) & ((1u << D) - 1u);
      if (bits != 0) {
        const int I0 = fr / 2, fy = fr - 2 * I0;
        const int J0 = fc / 2, fx = fc - 2 * J0;
        const float wy1 = float(fy) / 2.0f, wy0 = 1.0f - wy1;
        const float wx1 = float(fx) / 2.0f, wx0 = 1.0f - wx1;
        const int l00 = (I0 - er0) * C::EW + (J0 - ec0);
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const float a00 = sec[l00][k];
          const float a01 = fx > 0 ? sec[l00 + 1][k] : 0.0f;
          const float a10 = fy > 0 ? sec[l00 + C::EW][k] : 0.0f;
          const float a11 = (fx > 0 && fy > 0) ? sec[l00 + C::EW + 1][k] : 0.0f;
          const float top = wx0 * a00 + wx1 * a01;
          const float bot = wx0 * a10 + wx1 * a11;
          if ((bits >> k) & 1u) v.v[k] += wy0 * top + wy1 * bot;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < D; ++k) sxp[i][k] = v.v[k];
  }
  __syncthreads();
  // x3 one node around the tile, then x4 on it
#pragma unroll
  for (int j = 0; j < C::template per_thread<C::QH * C::QW>(); ++j) {
    const int i = threadIdx.x + j * C::NT;
    if (i >= C::QH * C::QW) break;
    const int lr = i / C::QW, lc = i % C::QW;
    if (lr < 1 || lr > C::QH - 2 || lc < 1 || lc > C::QW - 2) continue;
    const int fr = fr0 + lr, fc = fc0 + lc;
    float xo[D];
#pragma unroll
    for (int k = 0; k < D; ++k) xo[k] = 0.0f;
    if (inner[j]) {
      float U[3][3][D], acc[D];
      gather3<float, D, C::QW>(sxp, lr, lc, U);
      apply_row<D>(w[j], U, acc);
#pragma unroll
      for (int k = 0; k < D; ++k)
        xo[k] = ((w[j].free_bits >> k) & 1u)
                    ? U[1][1][k] + a.omega * sdv[i][k] * (sbv[i][k] - acc[k])
                    : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) sx3[i][k] = xo[k];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < C::template per_thread<C::QH * C::QW>(); ++j) {
    const int i = threadIdx.x + j * C::NT;
    if (i >= C::QH * C::QW) break;
    const int lr = i / C::QW, lc = i % C::QW;
    if (lr < 2 || lr > C::QH - 3 || lc < 2 || lc > C::QW - 3) continue;
    const int fr = fr0 + lr, fc = fc0 + lc;
    if (!inner[j]) continue;  // outside the grid
    float U[3][3][D], acc[D];
    gather3<float, D, C::QW>(sx3, lr, lc, U);
    apply_row<D>(w[j], U, acc);
    const long long node = (long long)fr * a.nnx + fc;
    Vec<float, D> xo;
#pragma unroll
    for (int k = 0; k < D; ++k)
      xo.v[k] = ((w[j].free_bits >> k) & 1u)
                    ? U[1][1][k] + a.omega * sdv[i][k] * (sbv[i][k] - acc[k])
                    : 0.0f;
    store_node<float, D>(a.out + row * a.so, node, xo);
  }
}

// One launch per level (the fine grid of one hierarchy), or one launch over the same
// level of several hierarchies: group i takes workgroups first[i] .. first[i + 1] - 1.
// Launch index bid of a pass over rows x tiles: its row and tile (XCD-aware order).
template <int D, typename C>
__device__ __forceinline__ void pair_pre_body(const CoarsePairArgs& a, int bid) {
  const int t = xcd_remap(bid, a.total);
  int row, tt;
  row_tile(t, a.tiles, a.total, row, tt);
  pair_pre_tile<D, C>(a, row, tt);
}

template <int D, typename C>
__device__ __forceinline__ void pair_post_body(const CoarsePairArgs& a, int bid) {
  const int t = xcd_remap(bid, a.total);
  int row, tt;
  row_tile(t, a.tiles, a.total, row, tt);
  pair_post_tile<D, C>(a, row, tt);
}

constexpr int kMaxPairGroups = 8;
struct PairGroups {
  int n;
  int first[kMaxPairGroups + 1];
  CoarsePairArgs a[kMaxPairGroups];
};

__device__ __forceinline__ int group_of(const PairGroups& g, int b) {
  int i = 0;
  while (i + 1 < g.n && b >= g.first[i + 1]) ++i;
  return i;
}

template <int D, typename C>
__global__ __launch_bounds__(C::NT) void coarse_pair_pre_kernel(PairGroups g) {
  const int i = group_of(g, blockIdx.x);
  pair_pre_body<D, C>(g.a[i], blockIdx.x - g.first[i]);
}

template <int D, typename C>
__global__ __launch_bounds__(C::NT) void coarse_pair_post_kernel(PairGroups g) {
  const int i = group_of(g, blockIdx.x);
  pair_post_body<D, C>(g.a[i], blockIdx.x - g.first[i]);
}

int check_pair_level(const OpDesc& op, const char* what) {
  if (op.keff == nullptr || !op.keff_node || op.mask == nullptr || (op.d != 1 && op.d != 2) ||
      op.nelx % 2 != 0 || op.nely % 2 != 0) {
    set_error(std::string(what) +
              ": needs a node-stencil level with a free-dof mask, coarsened 2 x 2");
    return STTOPT_EINVAL;
  }
  return STTOPT_OK;
}

template <typename C>
CoarsePairArgs pair_args(const PairLevel& L, bool pre) {
  const OpDesc& op = L.op;
  CoarsePairArgs a{};
  a.nnx = op.nnx();
  a.nny = op.nny();
  a.cnnx = op.nelx / 2 + 1;
  a.cnny = op.nely / 2 + 1;
  if (pre) {
    a.tiles_x = ceil_div(a.cnnx, C::CW);
    a.tiles = a.tiles_x * ceil_div(a.cnny, C::CH);
  } else {
    a.tiles_x = ceil_div(a.nnx, C::TW);
    a.tiles = a.tiles_x * ceil_div(a.nny, C::TH);
  }
  a.total = a.tiles * L.rows;
  a.S = static_cast<const float*>(op.keff);
  a.sS = op.sk;
  a.mask = op.mask;
  a.cmask = L.cmask;
  a.b = L.b;
  a.sb = L.sb;
  a.invd = L.invd;
  a.sd = L.sd;
  a.x2 = L.x2;
  a.sx2 = L.sx2;
  a.ec = L.ec;
  a.sec = L.sec;
  a.out = pre ? const_cast<float*>(L.x2) : L.x4;
  a.so = pre ? L.sx2 : L.sx4;
  a.bc = L.bc;
  a.sbc = L.sbc;
  a.omega = static_cast<float>(L.omega);
  a.active = L.active;
  return a;
}

template <typename C>
int launch_pair(const PairLevel* lv, int n, bool pre, hipStream_t s) {
  PairGroups g{};
  int total = 0, d = 0;
  for (int i = 0; i < n; ++i) {
    if (int e = check_pair_level(lv[i].op, pre ? "coarse_pair_pre" : "coarse_pair_post"))
      return e;
    if (i > 0 && lv[i].op.d != d) {
      set_error("coarse pair pass: levels of one launch need the same dofs per node");
      return STTOPT_EINVAL;
    }
    d = lv[i].op.d;
    if (lv[i].rows <= 0) continue;
    g.a[g.n] = pair_args<C>(lv[i], pre);
    g.first[g.n] = total;
    total += g.a[g.n].total;
    ++g.n;
  }
  g.first[g.n] = total;
  if (total == 0) return STTOPT_OK;
  if (pre) {
    if (d == 2)
      hipLaunchKernelGGL((coarse_pair_pre_kernel<2, C>), dim3(total), dim3(C::NT), 0, s, g);
    else
      hipLaunchKernelGGL((coarse_pair_pre_kernel<1, C>), dim3(total), dim3(C::NT), 0, s, g);
  } else {
    if (d == 2)
      hipLaunchKernelGGL((coarse_pair_post_kernel<2, C>), dim3(total), dim3(C::NT), 0, s, g);
    else
      hipLaunchKernelGGL((coarse_pair_post_kernel<1, C>), dim3(total), dim3(C::NT), 0, s, g);
  }
  STT_CHECK_LAUNCH(pre ? "coarse_pair_pre_kernel" : "coarse_pair_post_kernel");
  return STTOPT_OK;
}

// ---- the bottom of the cycle in one persistent launch (vcycle_bottom) ----
constexpr int kBottomPhases = 2 * kMaxBottomLevels + 1;
// a wait gives up after this many ticks of the 100 MHz wall clock (0.2 s; a wait of a
// working launch lasts microseconds)
constexpr unsigned long long kBottomWaitTicks = 20000000ull;

struct BottomArgs {
  int nlev, rows, nphase;
  int per_row[kBottomPhases];           // items per row of each phase
  unsigned first[kBottomPhases + 1];    // first item of each phase
  CoarsePairArgs pre[kMaxBottomLevels], post[kMaxBottomLevels];
  long long n;
  const double* cinv;
  long long scinv;
  const float* cb;
  long long scb;
  const uint32_t* cmask;
  float* cx;
  long long scx;
  const int* active;
  unsigned* queue;  // [0] next item, [1] workgroups that have left the loop
  unsigned* done;   // [phase][row] items done
  int* err;
};

// Wave 0 of a workgroup (all its lanes, so that no branch of the work loop diverges
// within a wave): poll *c until it reaches `need` (giving up after kBottomWaitTicks,
// with *err set, and at once when an earlier wait has given up), then acquire: the
// stores its producers made before their release are visible to the loads that follow.
__device__ __forceinline__ void bottom_wait(const unsigned* c, unsigned need, int* err) {
  const long long t0 = wall_clock64();
  while (__builtin_amdgcn_readfirstlane(
             __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < need) {
    if (__builtin_amdgcn_readfirstlane(
            __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0)
      break;
    if ((unsigned long long)(wall_clock64() - t0) > kBottomWaitTicks) {
      atomicExch(err, 1);
      break;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // stale L2 lines dropped
}

// Wave 0 (all lanes): add one to *c (lane 0's add; the others add zero) and return the
// value before the add, the same in every lane.
__device__ __forceinline__ unsigned bottom_take(unsigned* c) {
  const unsigned v = atomicAdd(c, (threadIdx.x & (kWave - 1)) == 0 ? 1u : 0u);
  return __builtin_amdgcn_readfirstlane(v);
}

// Items are taken in list order, and an item depends only on items of the previous
// phase of its row, all earlier in the list and held by workgroups that are running:
// the earliest unfinished item always waits on nothing, so every wait ends.  Every
// branch of the loop is uniform per wavefront (the item and the phase are read as
// uniform, and the counter updates are made by all of wave 0): with a branch on thread 0
// before the loop's back edge the compiler ran the next item's barriers for the other
// lanes before thread 0 had counted the item and taken the next one, so the workgroup
// repeated its item without end.
template <int D>
__global__ __launch_bounds__(PairNarrow::NT) void vcycle_bottom_kernel(BottomArgs a) {
  __shared__ unsigned s_item;
  __shared__ unsigned s_last;
  const unsigned total = a.first[a.nphase];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  for (;;) {
    if (wave == 0) s_item = bottom_take(a.queue);
    __syncthreads();
    const unsigned item = __builtin_amdgcn_readfirstlane(s_item);
    if (item >= total) break;
    int p = 0;
    while (item >= a.first[p + 1]) ++p;
    const int local = (int)(item - a.first[p]);
    const int row = local / a.per_row[p], tile = local - row * a.per_row[p];
    if (p > 0) {
      if (wave == 0)
        bottom_wait(a.done + (p - 1) * a.rows + row, (unsigned)a.per_row[p - 1], a.err);
      __syncthreads();
    }
    if (p < a.nlev) {
      pair_pre_tile<D, PairNarrow>(a.pre[p], row, tile);
    } else if (p == a.nlev) {
      coarse_solve_body<float, true>(a.n, a.cinv, a.scinv, a.cb, a.scb, a.cmask, a.cx, a.scx,
                                     a.active, tile, row);
    } else {
      pair_post_tile<D, PairNarrow>(a.post[a.nphase - 1 - p], row, tile);
    }
    // every store of the item has completed (and its LDS is free for the next item);
    // then release them and count the item
    __syncthreads();
    if (wave == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // the L2 written back
      (void)bottom_take(a.done + p * a.rows + row);
    }
  }
  // the last workgroup to leave the loop zeroes the counters for the next launch (the
  // fence completes this workgroup's counter updates before its exit is counted, so no
  // update lands after the reset)
  if (wave == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    s_last = bottom_take(a.queue + 1) == gridDim.x - 1 ? 1u : 0u;
  }
  __syncthreads();
  if (s_last != 0) {
    const int nd = a.nphase * a.rows;
    for (int i = threadIdx.x; i < nd; i += blockDim.x)
      __hip_atomic_store(a.done + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x < 2)
      __hip_atomic_store(a.queue + threadIdx.x, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Workgroups of the persistent launch: as many as fit on the device at once.
int bottom_grid(int d) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  static thread_local int cached_dev = -1, cached[2] = {0, 0};
  if (cached_dev != dev) {
    cached_dev = dev;
    cached[0] = cached[1] = 0;
  }
  int& g = cached[d - 1];
  if (g == 0) {
    int cus = 0, per_cu = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    hipError_t e = d == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                &per_cu, vcycle_bottom_kernel<2>, PairNarrow::NT, 0)
                          : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                &per_cu, vcycle_bottom_kernel<1>, PairNarrow::NT, 0);
    if (e != hipSuccess || cus <= 0 || per_cu <= 0) return 256;
    g = cus * per_cu;
  }
  return g;
}

}  // namespace

long long bottom_sync_words(int rows) { return 2 + (long long)kBottomPhases * rows; }

int vcycle_bottom(const BottomCall& c, unsigned* sync, int* err, hipStream_t s) {
  if (c.nlev < 1 || c.nlev > kMaxBottomLevels || sync == nullptr || err == nullptr ||
      c.cinv == nullptr || c.cb == nullptr || c.cx == nullptr || c.n <= 0 || c.n % 2 != 0 ||
      c.scinv % 2 != 0 || (reinterpret_cast<uintptr_t>(c.cinv) & 15) != 0) {
    set_error("vcycle_bottom: 1 to 3 pair levels and an even-sized, aligned coarsest inverse");
    return STTOPT_EINVAL;
  }
  const int d = c.lv[0].op.d, rows = c.lv[0].rows;
  BottomArgs a{};
  a.nlev = c.nlev;
  a.rows = rows;
  a.nphase = 2 * c.nlev + 1;
  for (int l = 0; l < c.nlev; ++l) {
    if (int e = check_pair_level(c.lv[l].op, "vcycle_bottom")) return e;
    if (c.lv[l].op.d != d || c.lv[l].rows != rows) {
      set_error("vcycle_bottom: the levels need the same dofs per node and rows");
      return STTOPT_EINVAL;
    }
    a.pre[l] = pair_args<PairNarrow>(c.lv[l], true);
    a.post[l] = pair_args<PairNarrow>(c.lv[l], false);
    a.per_row[l] = a.pre[l].tiles;
    a.per_row[a.nphase - 1 - l] = a.post[l].tiles;
  }
  if (rows <= 0) return STTOPT_OK;
  a.per_row[c.nlev] = (int)ceil_div(c.n, (long long)(PairNarrow::NT / kWave));
  long long total = 0;
  for (int p = 0; p < a.nphase; ++p) {
    a.first[p] = (unsigned)total;
    total += (long long)a.per_row[p] * rows;
  }
  if (total >= 0x7fffffffLL) {
    set_error("vcycle_bottom: too many items");
    return STTOPT_EINVAL;
  }
  a.first[a.nphase] = (unsigned)total;
  a.n = c.n;
  a.cinv = c.cinv;
  a.scinv = c.scinv;
  a.cb = c.cb;
  a.scb = c.scb;
  a.cmask = c.cmask;
  a.cx = c.cx;
  a.scx = c.scx;
  a.active = c.lv[0].active;
  a.queue = sync;
  a.done = sync + 2;
  a.err = err;
  const int grid = (int)(total < bottom_grid(d) ? total : bottom_grid(d));
  if (d == 2)
    hipLaunchKernelGGL(vcycle_bottom_kernel<2>, dim3(grid), dim3(PairNarrow::NT), 0, s, a);
  else
    hipLaunchKernelGGL(vcycle_bottom_kernel<1>, dim3(grid), dim3(PairNarrow::NT), 0, s, a);
  STT_CHECK_LAUNCH("vcycle_bottom_kernel");
  return STTOPT_OK;
}

int coarse_pair_groups(const PairLevel* lv, int n, bool pre, int variant, hipStream_t s) {
  if (n < 1 || n > kMaxPairGroups) {
    set_error("coarse pair pass: 1 to 8 levels per launch");
    return STTOPT_EINVAL;
  }
  if (variant == 1) return launch_pair<PairNarrow>(lv, n, pre, s);
  return launch_pair<PairWide>(lv, n, pre, s);
}

int coarse_pair_pre(const OpDesc& op, int rows, const float* b, long long sb,
                    const float* invd, long long sd, float* x2, long long sx2,
                    const uint32_t* cmask, float* bc, long long sbc, double omega,
                    const int* active, hipStream_t s, int variant) {
  PairLevel L{};
  L.op = op;
  L.rows = rows;
  L.b = b;
  L.sb = sb;
  L.invd = invd;
  L.sd = sd;
  L.x2 = x2;
  L.sx2 = sx2;
  L.cmask = cmask;
  L.bc = bc;
  L.sbc = sbc;
  L.omega = omega;
  L.active = active;
  return coarse_pair_groups(&L, 1, true, variant, s);
}

int coarse_pair_post(const OpDesc& op, int rows, const float* x2, long long sx2,
                     const float* ec, long long sec, const float* b, long long sb,
                     const float* invd, long long sd, float* x4, long long sx4, double omega,
                     const int* active, hipStream_t s, int variant) {
  PairLevel L{};
  L.op = op;
  L.rows = rows;
  L.b = b;
  L.sb = sb;
  L.invd = invd;
  L.sd = sd;
  L.x2 = x2;
  L.sx2 = sx2;
  L.ec = ec;
  L.sec = sec;
  L.x4 = x4;
  L.sx4 = sx4;
  L.omega = omega;
  L.active = active;
  return coarse_pair_groups(&L, 1, false, variant, s);
}

}  // namespace stt
