// Geometric multigrid on the structured Q4 grid: on-device Galerkin coarse
// operators, full-weighting restriction / bilinear prolongation, the dense
// coarsest solve, and the V-cycle driver (torch_mg.py:293-476).
#include "stt_kernels.h"

namespace stt {

// ------------------------------------------------------------- Galerkin -----
// Explicit form (used with a reaction term, where no per-position table exists):
//   keff_E[a][b] = sum_m sum_{p,q} S_m[p][a] keff_{child m}[p][q] S_m[q][b]
// One wavefront per (64 coarse elements, one entry (a, b), one row): the entry is
// uniform across the wave, so the interpolation weights are scalar loads, and the 64
// lanes read 64 neighbouring fine elements.
template <int D>
__global__ __launch_bounds__(64) void galerkin_kernel(int nelx, int nely, int kx, int ky,
                                                      const double* fk, long long sfk,
                                                      const double* S, double* ck,
                                                      long long sck) {
  constexpr int K = 4 * D, KK = K * K;
  const int cnelx = nelx / kx, cnely = nely / ky;
  const long long cnel = (long long)cnelx * cnely;
  const long long fnel = (long long)nelx * nely;
  const long long E = (long long)blockIdx.x * 64 + threadIdx.x;
  const int entry = blockIdx.y;
  const int row = blockIdx.z;
  if (E >= cnel) return;
  const int A = entry / K, B = entry - (entry / K) * K;
  const int cr = (int)(E / cnelx), cc = (int)(E - (long long)cr * cnelx);
  double acc = 0.0;
  for (int my = 0; my < ky; ++my) {
    for (int mx = 0; mx < kx; ++mx) {
      const int m = my * kx + mx;
      const long long f = (long long)(cr * ky + my) * nelx + (cc * kx + mx);
      const double* Sm = S + m * KK;
      for (int p = 0; p < K; ++p) {
        const double spa = Sm[p * K + A];
        if (spa == 0.0) continue;
        double inner = 0.0;
        for (int q = 0; q < K; ++q) {
          const double sqb = Sm[q * K + B];
          if (sqb == 0.0) continue;
          inner += fk[row * sfk + (long long)(p * K + q) * fnel + f] * sqb;
        }
        acc += spa * inner;
      }
    }
  }
  ck[row * sck + (long long)entry * cnel + E] = acc;
}

// Density form, any block size: keff_E[a][b] = sum over the pkx x pky fine elements g
// under E of rho_g G_pos(g)[a][b] (+ react_g GM_pos(g)[a][b]).  By linearity this equals
// the level-by-level Galerkin product when G holds the per-position products of KE
// through every intermediate level.  One thread per coarse element and EG consecutive
// entries; the table rows are uniform across the wavefront (scalar loads).
template <int D, int EG>
__global__ __launch_bounds__(256) void galerkin_param_kernel(
    int fnelx, int pkx, int pky, int cnelx, long long cnel, const double* __restrict__ rho,
    long long sr, const double* __restrict__ react, long long sre,
    const double* __restrict__ g, const double* __restrict__ gm, double* __restrict__ ck,
    long long sck) {
  constexpr int KK = 16 * D * D;
  const long long E = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int e0 = blockIdx.y * EG;
  const int row = blockIdx.z;
  if (E >= cnel) return;
  const int cr = (int)(E / cnelx), cc = (int)(E - (long long)cr * cnelx);
  const double* rr = rho + row * sr;
  const double* re = react != nullptr ? react + row * sre : nullptr;
  double acc[EG];
#pragma unroll
  for (int q = 0; q < EG; ++q) acc[q] = 0.0;
  for (int gy = 0; gy < pky; ++gy) {
    const long long frow = (long long)(cr * pky + gy) * fnelx + (long long)cc * pkx;
#pragma unroll 4
    for (int gx = 0; gx < pkx; ++gx) {
      const int pos = gy * pkx + gx;
      const double r = rr[frow + gx];
      const double* G = g + (long long)pos * KK + e0;
#pragma unroll
      for (int q = 0; q < EG; ++q) acc[q] += r * G[q];
      if (re != nullptr) {
        const double t = re[frow + gx];
        const double* M = gm + (long long)pos * KK + e0;
#pragma unroll
        for (int q = 0; q < EG; ++q) acc[q] += t * M[q];
      }
    }
  }
  double* out = ck + row * sck + E;
#pragma unroll
  for (int q = 0; q < EG; ++q) out[(long long)(e0 + q) * cnel] = acc[q];
}

// Packed upper triangle of symmetric entry-major element matrices, (A + A^T) / 2.
template <typename TO, int K>
__global__ __launch_bounds__(256) void pack_keff_kernel(long long nel,
                                                        const double* __restrict__ keff,
                                                        long long sk, TO* __restrict__ out,
                                                        long long so) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  if (e >= nel) return;
  const double* in = keff + row * sk + e;
  TO* o = out + row * so + e;
  int idx = 0;
#pragma unroll
  for (int r = 0; r < K; ++r) {
#pragma unroll
    for (int c = r; c < K; ++c) {
      o[(long long)idx * nel] =
          static_cast<TO>(0.5 * (in[(long long)(r * K + c) * nel] + in[(long long)(c * K + r) * nel]));
      ++idx;
    }
  }
}

// Assembled node stencil of symmetric entry-major element matrices (STTOPT_KEFF_NODE):
// one record of node_stencil_stride(d) values per node: the upper triangle of its own
// d x d block (zero padded to 4 values), then the d x d couplings to its E, NW, N, NE
// neighbours (node row + 0/1, column + 1/-1/0/1), each the sum over the elements
// holding both nodes of (A + A^T) / 2.  The other four neighbours' couplings are those
// neighbours' forward blocks, transposed.  Sums in fp64, one rounding to TO.
template <typename TO, int D>
__global__ __launch_bounds__(256) void node_stencil_kernel(int nelx, int nely,
                                                           const double* __restrict__ keff,
                                                           long long sk, TO* __restrict__ out,
                                                           long long so) {
  constexpr int K = 4 * D, ND = 4, NS = node_stencil_stride(D);
  const int nnx = nelx + 1;
  const long long nn = (long long)nnx * (nely + 1), nel = (long long)nelx * nely;
  const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  if (node >= nn) return;
  const int r = (int)(node / nnx), c = (int)(node - (long long)r * nnx);
  const double* ke = keff + row * sk;
  constexpr int fdr[4] = {0, 1, 1, 1}, fdc[4] = {1, -1, 0, 1};  // E, NW, N, NE
  double acc[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) acc[k] = 0.0;
  // local index of node (dr, dc) of an element (offsets from its first node row / column)
  auto local = [](int dr, int dc) { return dr == 1 ? (dc == 0 ? 0 : 1) : (dc == 1 ? 2 : 3); };
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int er = r - 1 + (q >> 1), ec = c - 1 + (q & 1);
    if (er < 0 || er >= nely || ec < 0 || ec >= nelx) continue;
    const long long e = (long long)er * nelx + ec;
    const int ln = local(r - er, c - ec);
    auto sym = [&](int a, int b) {
      return 0.5 * (ke[(long long)(a * K + b) * nel + e] + ke[(long long)(b * K + a) * nel + e]);
    };
    int idx = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
#pragma unroll
      for (int k = i; k < D; ++k) acc[idx++] += sym(ln * D + i, ln * D + k);
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const int mr = r + fdr[f] - er, mc = c + fdc[f] - ec;
      if (mr < 0 || mr > 1 || mc < 0 || mc > 1) continue;
      const int lm = local(mr, mc);
#pragma unroll
      for (int i = 0; i < D; ++i) {
#pragma unroll
        for (int k = 0; k < D; ++k) acc[ND + (f * D + i) * D + k] += sym(ln * D + i, lm * D + k);
      }
    }
  }
  TO* o = out + row * so + node * NS;
#pragma unroll
  for (int k = 0; k < NS; ++k) o[k] = static_cast<TO>(acc[k]);
}

// ------------------------------------------------------------ transfers -----

// Both transfers work on whole nodes (the D dofs as one vector).  KX = KY = 0 takes
// the coarsening factors at run time; the common 2 x 2 case is compiled with them.
template <typename T, int D, int KX, int KY>
__global__ __launch_bounds__(256) void restrict_kernel(int cnelx, int cnely, int kx_, int ky_,
                                                       const T* __restrict__ fine,
                                                       long long sf,
                                                       const uint32_t* __restrict__ cmask,
                                                       T* __restrict__ coarse, long long sc,
                                                       const int* __restrict__ active) {
  const int row = blockIdx.y;
  if (active != nullptr && active[row] == 0) return;
  const int kx = KX > 0 ? KX : kx_, ky = KY > 0 ? KY : ky_;
  const int cnnx = cnelx + 1, cnny = cnely + 1;
  const int fnnx = cnelx * kx + 1, fnny = cnely * ky + 1;
  const long long cnode = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cnode >= (long long)cnnx * cnny) return;
  const int I = (int)(cnode / cnnx), J = (int)(cnode - (long long)I * cnnx);
  const T* f = fine + row * sf;
  T acc[D];
#pragma unroll
  for (int k = 0; k < D; ++k) acc[k] = T(0);
#pragma unroll
  for (int dy = -(ky - 1); dy <= ky - 1; ++dy) {
    const int R = I * ky + dy;
    if (R < 0 || R >= fnny) continue;
    const T wy = T(ky - (dy < 0 ? -dy : dy)) / T(ky);
    T racc[D];
#pragma unroll
    for (int k = 0; k < D; ++k) racc[k] = T(0);
#pragma unroll
    for (int dx = -(kx - 1); dx <= kx - 1; ++dx) {
      const int C = J * kx + dx;
      if (C < 0 || C >= fnnx) continue;
      const T wx = T(kx - (dx < 0 ? -dx : dx)) / T(kx);
      const Vec<T, D> v = load_node<T, D>(f, (long long)R * fnnx + C);
#pragma unroll
      for (int k = 0; k < D; ++k) racc[k] += wx * v.v[k];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) acc[k] += wy * racc[k];
  }
  const uint32_t word = cmask != nullptr ? cmask[(cnode * D) >> 5] : ~0u;
  Vec<T, D> out;
#pragma unroll
  for (int k = 0; k < D; ++k) out.v[k] = ((word >> ((cnode * D + k) & 31)) & 1u) ? acc[k] : T(0);
  store_node<T, D>(coarse + row * sc, cnode, out);
}

// 2 x 2 restriction with one load per fine node and row: each thread loads the fine
// nodes 2J and 2J + 1 of the three fine rows around its coarse node J; node 2J - 1 is
// the left lane's 2J' + 1 (a wave-wide DPP shift), loaded directly only by lane 0.  The
// sums run in the order of restrict_kernel (dx = -1, 0, 1), so results are identical.
template <typename T>
__device__ __forceinline__ T lane_from_left(T v) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, false));
  } else {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
  }
}

template <typename T, int D>
__device__ __forceinline__ void restrict22_body(int cnelx, int cnely, const T* __restrict__ fine,
                                                long long sf,
                                                const uint32_t* __restrict__ cmask,
                                                T* __restrict__ coarse, long long sc,
                                                const int* __restrict__ active, const int bx,
                                                const int row) {
  if (active != nullptr && active[row] == 0) return;  // whole block
  const int cnnx = cnelx + 1, cnny = cnely + 1;
  const int fnnx = 2 * cnelx + 1, fnny = 2 * cnely + 1;
  const long long cn = (long long)cnnx * cnny;
  const long long cnode = (long long)bx * blockDim.x + threadIdx.x;
  const bool valid = cnode < cn;  // every lane takes part in the shifts
  const long long cnc = valid ? cnode : cn - 1;
  const int I = (int)(cnc / cnnx), J = (int)(cnc - (long long)I * cnnx);
  const int lane = threadIdx.x & (kWave - 1);
  const bool own_left = lane == 0;  // no left lane in this wavefront
  const T* f = fine + row * sf;
  T acc[D];
#pragma unroll
  for (int k = 0; k < D; ++k) acc[k] = T(0);
#pragma unroll
  for (int dy = -1; dy <= 1; ++dy) {
    const int R = 2 * I + dy;
    const bool rin = R >= 0 && R < fnny;
    const long long base = (long long)(rin ? R : 0) * fnnx;
    Vec<T, D> a, b, l;
#pragma unroll
    for (int k = 0; k < D; ++k) a.v[k] = b.v[k] = l.v[k] = T(0);
    if (rin) a = load_node<T, D>(f, base + 2 * J);
    if (rin && 2 * J + 1 < fnnx) b = load_node<T, D>(f, base + 2 * J + 1);
#pragma unroll
    for (int k = 0; k < D; ++k) l.v[k] = lane_from_left(b.v[k]);
    if (own_left && rin && J > 0) l = load_node<T, D>(f, base + 2 * J - 1);
    if (J == 0) {
#pragma unroll
      for (int k = 0; k < D; ++k) l.v[k] = T(0);
    }
    if (!rin) continue;
    const T wy = T(2 - (dy < 0 ? -dy : dy)) / T(2);
    T racc[D];
#pragma unroll
    for (int k = 0; k < D; ++k) racc[k] = T(0);
    const T wh = T(1) / T(2), w1 = T(2) / T(2);
    if (J > 0) {
#pragma unroll
      for (int k = 0; k < D; ++k) racc[k] += wh * l.v[k];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) racc[k] += w1 * a.v[k];
    if (2 * J + 1 < fnnx) {
#pragma unroll
      for (int k = 0; k < D; ++k) racc[k] += wh * b.v[k];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) acc[k] += wy * racc[k];
  }
  if (!valid) return;
  const uint32_t word = cmask != nullptr ? cmask[(cnode * D) >> 5] : ~0u;
  Vec<T, D> out;
#pragma unroll
  for (int k = 0; k < D; ++k) out.v[k] = ((word >> ((cnode * D + k) & 31)) & 1u) ? acc[k] : T(0);
  store_node<T, D>(coarse + row * sc, cnode, out);
}

template <typename T, int D, int KX, int KY>
__device__ __forceinline__ void prolong_add_body(int cnelx, int cnely, int kx_, int ky_,
                                                 const T* __restrict__ coarse, long long sc,
                                                 const uint32_t* __restrict__ fmask,
                                                 T* __restrict__ fine, long long sf,
                                                 const int* __restrict__ active, const int bx,
                                                 const int row) {
  if (active != nullptr && active[row] == 0) return;
  const int kx = KX > 0 ? KX : kx_, ky = KY > 0 ? KY : ky_;
  const int cnnx = cnelx + 1;
  const int fnnx = cnelx * kx + 1, fnny = cnely * ky + 1;
  const long long fnode = (long long)bx * blockDim.x + threadIdx.x;
  if (fnode >= (long long)fnnx * fnny) return;
  const uint32_t word = fmask != nullptr ? fmask[(fnode * D) >> 5] : ~0u;
  const uint32_t bits = (word >> ((fnode * D) & 31)) & ((1u << D) - 1u);
  if (bits == 0) return;  // every dof of the node fixed: nothing to add
  const int R = (int)(fnode / fnnx), C = (int)(fnode - (long long)R * fnnx);
  const int I0 = R / ky, fy = R - I0 * ky;
  const int J0 = C / kx, fx = C - J0 * kx;
  const T wy1 = T(fy) / T(ky), wy0 = T(1) - wy1;
  const T wx1 = T(fx) / T(kx), wx0 = T(1) - wx1;
  const T* c = coarse + row * sc;
  const long long c00 = (long long)I0 * cnnx + J0;
  const Vec<T, D> a00 = load_node<T, D>(c, c00);
  Vec<T, D> a01, a10, a11;
#pragma unroll
  for (int k = 0; k < D; ++k) a01.v[k] = a10.v[k] = a11.v[k] = T(0);
  if (fx > 0) a01 = load_node<T, D>(c, c00 + 1);
  if (fy > 0) a10 = load_node<T, D>(c, c00 + cnnx);
  if (fx > 0 && fy > 0) a11 = load_node<T, D>(c, c00 + cnnx + 1);
  T* fr = fine + row * sf;
  Vec<T, D> v = load_node<T, D>(fr, fnode);
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const T top = wx0 * a00.v[k] + wx1 * a01.v[k];
    const T bot = wx0 * a10.v[k] + wx1 * a11.v[k];
    if ((bits >> k) & 1u) v.v[k] += wy0 * top + wy1 * bot;
  }
  store_node<T, D>(fr, fnode, v);
}

template <typename T, int D>
__global__ __launch_bounds__(256) void restrict22_kernel(int cnelx, int cnely,
                                                         const T* __restrict__ fine,
                                                         long long sf,
                                                         const uint32_t* __restrict__ cmask,
                                                         T* __restrict__ coarse, long long sc,
                                                         const int* __restrict__ active) {
  restrict22_body<T, D>(cnelx, cnely, fine, sf, cmask, coarse, sc, active, blockIdx.x,
                        blockIdx.y);
}

template <typename T, int D, int KX, int KY>
__global__ __launch_bounds__(256) void prolong_add_kernel(int cnelx, int cnely, int kx_,
                                                          int ky_, const T* __restrict__ coarse,
                                                          long long sc,
                                                          const uint32_t* __restrict__ fmask,
                                                          T* __restrict__ fine, long long sf,
                                                          const int* __restrict__ active) {
  prolong_add_body<T, D, KX, KY>(cnelx, cnely, kx_, ky_, coarse, sc, fmask, fine, sf, active,
                                 blockIdx.x, blockIdx.y);
}

// 2 x 2 transfers of several hierarchies in one launch: hierarchy i takes workgroups
// first[i] .. first[i + 1] - 1, bx_count per row.
struct TransferGroups {
  int n;
  int first[kMaxBatchGroups + 1];
  TransferCall a[kMaxBatchGroups];
  int bx_count[kMaxBatchGroups];
};

__device__ __forceinline__ int transfer_group(const int* first, int n, int b) {
  int i = 0;
  while (i + 1 < n && b >= first[i + 1]) ++i;
  return i;
}

template <int D>
__global__ __launch_bounds__(256) void restrict22_groups_kernel(TransferGroups g) {
  const int i = transfer_group(g.first, g.n, blockIdx.x);
  const int b = blockIdx.x - g.first[i], bx = b % g.bx_count[i], row = b / g.bx_count[i];
  const TransferCall& c = g.a[i];
  restrict22_body<float, D>(c.cnelx, c.cnely, c.src, c.ss, c.mask, c.dst, c.sd, c.active, bx,
                            row);
}

template <int D>
__global__ __launch_bounds__(256) void prolong22_groups_kernel(TransferGroups g) {
  const int i = transfer_group(g.first, g.n, blockIdx.x);
  const int b = blockIdx.x - g.first[i], bx = b % g.bx_count[i], row = b / g.bx_count[i];
  const TransferCall& c = g.a[i];
  prolong_add_body<float, D, 2, 2>(c.cnelx, c.cnely, 2, 2, c.src, c.ss, c.mask, c.dst, c.sd,
                                   c.active, bx, row);
}

int transfer22_groups(const TransferCall* calls, int n, bool restrict_, hipStream_t s) {
  if (n < 1 || n > kMaxBatchGroups) {
    set_error("transfer groups: 1 to 4 hierarchies per launch");
    return STTOPT_EINVAL;
  }
  TransferGroups g{};
  int total = 0, d = 0;
  for (int i = 0; i < n; ++i) {
    const TransferCall& c = calls[i];
    if (c.rows <= 0) continue;
    if ((c.d != 1 && c.d != 2) || (g.n > 0 && c.d != d) ||
        (c.d == 2 && !(pair_aligned(c.src, c.ss, sizeof(float)) &&
                       pair_aligned(c.dst, c.sd, sizeof(float))))) {
      set_error("transfer groups: one dofs-per-node per launch, node-aligned vectors");
      return STTOPT_EINVAL;
    }
    d = c.d;
    // restriction: one thread per coarse node; prolongation: per fine node
    const long long nodes = restrict_ ? (long long)(c.cnelx + 1) * (c.cnely + 1)
                                      : (long long)(2 * c.cnelx + 1) * (2 * c.cnely + 1);
    g.a[g.n] = c;
    g.bx_count[g.n] = ceil_div(nodes, 256);
    g.first[g.n] = total;
    total += g.bx_count[g.n] * c.rows;
    ++g.n;
  }
  g.first[g.n] = total;
  if (total == 0) return STTOPT_OK;
  if (restrict_) {
    if (d == 2) hipLaunchKernelGGL(restrict22_groups_kernel<2>, dim3(total), dim3(256), 0, s, g);
    else hipLaunchKernelGGL(restrict22_groups_kernel<1>, dim3(total), dim3(256), 0, s, g);
  } else {
    if (d == 2) hipLaunchKernelGGL(prolong22_groups_kernel<2>, dim3(total), dim3(256), 0, s, g);
    else hipLaunchKernelGGL(prolong22_groups_kernel<1>, dim3(total), dim3(256), 0, s, g);
  }
  STT_CHECK_LAUNCH("transfer22_groups_kernel");
  return STTOPT_OK;
}

template <typename T>
int restrict_to(int d, int cnelx, int cnely, int kx, int ky, int rows, const T* fine,
                long long sf, const uint32_t* cmask, T* coarse, long long sc,
                const int* active, hipStream_t s) {
  if (rows <= 0) return STTOPT_OK;
  if (d == 2 && !(pair_aligned(fine, sf, sizeof(T)) && pair_aligned(coarse, sc, sizeof(T)))) {
    set_error("restrict: d = 2 vectors need node-aligned pointers and even row strides");
    return STTOPT_EINVAL;
  }
  const long long n = (long long)(cnelx + 1) * (cnely + 1);
  const dim3 grid(ceil_div(n, 256), rows), block(256);
  const bool k22 = kx == 2 && ky == 2;
  if (d == 2) {
    if (k22)
      hipLaunchKernelGGL((restrict22_kernel<T, 2>), grid, block, 0, s, cnelx, cnely, fine, sf,
                         cmask, coarse, sc, active);
    else
      hipLaunchKernelGGL((restrict_kernel<T, 2, 0, 0>), grid, block, 0, s, cnelx, cnely, kx,
                         ky, fine, sf, cmask, coarse, sc, active);
  } else {
    if (k22)
      hipLaunchKernelGGL((restrict22_kernel<T, 1>), grid, block, 0, s, cnelx, cnely, fine, sf,
                         cmask, coarse, sc, active);
    else
      hipLaunchKernelGGL((restrict_kernel<T, 1, 0, 0>), grid, block, 0, s, cnelx, cnely, kx,
                         ky, fine, sf, cmask, coarse, sc, active);
  }
  STT_CHECK_LAUNCH("restrict_kernel");
  return STTOPT_OK;
}

template <typename T>
int prolong_add(int d, int cnelx, int cnely, int kx, int ky, int rows, const T* coarse,
                long long sc, const uint32_t* fmask, T* fine, long long sf,
                const int* active, hipStream_t s) {
  if (rows <= 0) return STTOPT_OK;
  if (d == 2 && !(pair_aligned(fine, sf, sizeof(T)) && pair_aligned(coarse, sc, sizeof(T)))) {
    set_error("prolong: d = 2 vectors need node-aligned pointers and even row strides");
    return STTOPT_EINVAL;
  }
  const long long n = (long long)(cnelx * kx + 1) * (cnely * ky + 1);
  const dim3 grid(ceil_div(n, 256), rows), block(256);
  const bool k22 = kx == 2 && ky == 2;
  if (d == 2) {
    if (k22)
      hipLaunchKernelGGL((prolong_add_kernel<T, 2, 2, 2>), grid, block, 0, s, cnelx, cnely,
                         kx, ky, coarse, sc, fmask, fine, sf, active);
    else
      hipLaunchKernelGGL((prolong_add_kernel<T, 2, 0, 0>), grid, block, 0, s, cnelx, cnely,
                         kx, ky, coarse, sc, fmask, fine, sf, active);
  } else {
    if (k22)
      hipLaunchKernelGGL((prolong_add_kernel<T, 1, 2, 2>), grid, block, 0, s, cnelx, cnely,
                         kx, ky, coarse, sc, fmask, fine, sf, active);
    else
      hipLaunchKernelGGL((prolong_add_kernel<T, 1, 0, 0>), grid, block, 0, s, cnelx, cnely,
                         kx, ky, coarse, sc, fmask, fine, sf, active);
  }
  STT_CHECK_LAUNCH("prolong_add_kernel");
  return STTOPT_OK;
}

template int restrict_to<double>(int, int, int, int, int, int, const double*, long long,
                                 const uint32_t*, double*, long long, const int*,
                                 hipStream_t);
template int restrict_to<float>(int, int, int, int, int, int, const float*, long long,
                                const uint32_t*, float*, long long, const int*, hipStream_t);
template int prolong_add<double>(int, int, int, int, int, int, const double*, long long,
                                 const uint32_t*, double*, long long, const int*,
                                 hipStream_t);
template int prolong_add<float>(int, int, int, int, int, int, const float*, long long,
                                const uint32_t*, float*, long long, const int*, hipStream_t);

// ---------------------------------------------------------- coarse solve ----
// (coarse_solve_body: stt_kernels.h)

template <typename T, bool PAIRS>
__global__ __launch_bounds__(256) void coarse_solve_kernel(long long n, const double* cinv,
                                                           long long scinv, const T* b,
                                                           long long sb, const uint32_t* mask,
                                                           T* x, long long sx,
                                                           const int* active) {
  coarse_solve_body<T, PAIRS>(n, cinv, scinv, b, sb, mask, x, sx, active, blockIdx.x,
                              blockIdx.y);
}

// The coarsest solves of several hierarchies in one launch (each at its own coarsest
// level): hierarchy i takes workgroups first[i] .. first[i + 1] - 1, bx_count per row.
struct CoarseSolveGroups {
  int n;
  int first[kMaxBatchGroups + 1];
  CoarseSolveCall a[kMaxBatchGroups];
  int bx_count[kMaxBatchGroups];
};

__global__ __launch_bounds__(256) void coarse_solve_groups_kernel(CoarseSolveGroups g) {
  int i = 0;
  while (i + 1 < g.n && (int)blockIdx.x >= g.first[i + 1]) ++i;
  const int b = blockIdx.x - g.first[i], bx = b % g.bx_count[i], row = b / g.bx_count[i];
  const CoarseSolveCall& c = g.a[i];
  coarse_solve_body<float, true>(c.n, c.cinv, c.scinv, c.b, c.sb, c.mask, c.x, c.sx, c.active,
                                 bx, row);
}

int coarse_solve_groups(const CoarseSolveCall* calls, int n, hipStream_t s) {
  if (n < 1 || n > kMaxBatchGroups) {
    set_error("coarse solve groups: 1 to 4 hierarchies per launch");
    return STTOPT_EINVAL;
  }
  CoarseSolveGroups g{};
  int total = 0;
  for (int i = 0; i < n; ++i) {
    const CoarseSolveCall& c = calls[i];
    if (c.rows <= 0 || c.n <= 0) continue;
    if (c.cinv == nullptr || c.n % 2 != 0 || c.scinv % 2 != 0 ||
        (reinterpret_cast<uintptr_t>(c.cinv) & 15) != 0) {
      set_error("coarse solve groups: even-sized inverses with 16-byte aligned rows");
      return STTOPT_EINVAL;
    }
    g.a[g.n] = c;
    g.bx_count[g.n] = ceil_div(c.n, 4);
    g.first[g.n] = total;
    total += g.bx_count[g.n] * c.rows;
    ++g.n;
  }
  g.first[g.n] = total;
  if (total == 0) return STTOPT_OK;
  hipLaunchKernelGGL(coarse_solve_groups_kernel, dim3(total), dim3(256), 0, s, g);
  STT_CHECK_LAUNCH("coarse_solve_groups_kernel");
  return STTOPT_OK;
}

template <typename T>
int coarse_solve(long long n, int rows, const double* cinv, long long scinv, const T* b,
                 long long sb, const uint32_t* mask, T* x, long long sx, const int* active,
                 hipStream_t s) {
  if (rows <= 0 || n <= 0) return STTOPT_OK;
  if (cinv == nullptr) {
    set_error("coarse_solve: coarsest level has no dense inverse");
    return STTOPT_EINVAL;
  }
  const bool pairs = n % 2 == 0 && scinv % 2 == 0 &&
                     (reinterpret_cast<uintptr_t>(cinv) & 15) == 0;
  if (pairs)
    hipLaunchKernelGGL((coarse_solve_kernel<T, true>), dim3(ceil_div(n, 4), rows), dim3(256),
                       0, s, n, cinv, scinv, b, sb, mask, x, sx, active);
  else
    hipLaunchKernelGGL((coarse_solve_kernel<T, false>), dim3(ceil_div(n, 4), rows), dim3(256),
                       0, s, n, cinv, scinv, b, sb, mask, x, sx, active);
  STT_CHECK_LAUNCH("coarse_solve_kernel");
  return STTOPT_OK;
}
template int coarse_solve<double>(long long, int, const double*, long long, const double*,
                                  long long, const uint32_t*, double*, long long,
                                  const int*, hipStream_t);
template int coarse_solve<float>(long long, int, const double*, long long, const float*,
                                 long long, const uint32_t*, float*, long long, const int*,
                                 hipStream_t);

// ------------------------------------------------------- dense assembly -----
// Entry (i, j) of the coarsest operator is gathered from the (up to four)
// elements holding both nodes, then symmetrized as (K_ij + K_ji) / 2; fixed dofs
// get identity rows and columns (torch_mg._dense_cholesky, torch_mg.py:269-289).

template <int D>
__device__ double dense_entry(int nelx, int nely, int r, int c, int ci, int r2, int c2,
                              int cj, const double* keff, const double* rho,
                              const double* ke) {
  constexpr int K = 4 * D;
  const long long nel = (long long)nelx * nely;
  double v = 0.0;
  for (int q = 0; q < 4; ++q) {
    const int er = r - 1 + (q >> 1), ec = c - 1 + (q & 1);
    if (er < 0 || er >= nely || ec < 0 || ec >= nelx) continue;
    const int pr = r2 - er, pc = c2 - ec;  // second node relative to element top-left
    if (pr < 0 || pr > 1 || pc < 0 || pc > 1) continue;
    const int jl = pr == 1 ? (pc == 0 ? 0 : 1) : (pc == 1 ? 2 : 3);
    const int il = q == 0 ? 1 : q == 1 ? 0 : q == 2 ? 2 : 3;
    const int a = il * D + ci, b = jl * D + cj;
    const long long e = (long long)er * nelx + ec;
    if (keff != nullptr)
      v += keff[(long long)(a * K + b) * nel + e];
    else
      v += rho[e] * ke[a * K + b];
  }
  return v;
}

template <int D>
__global__ __launch_bounds__(256) void dense_assemble_kernel(int nelx, int nely,
                                                             const double* keff, long long sk,
                                                             const double* rho, long long sr,
                                                             const double* rdiag, long long srd,
                                                             const double* ke,
                                                             const uint32_t* mask,
                                                             double* out, long long so) {
  const int row = blockIdx.y;
  const int nnx = nelx + 1, nny = nely + 1;
  const long long n = (long long)nnx * nny * D;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  // one thread per (dof i, neighbour slot): 9 neighbour nodes x D components
  const long long i = t / (9 * D);
  if (i >= n) return;
  const int slot = (int)(t - i * 9 * D);
  const int nb = slot / D, cj = slot - (slot / D) * D;
  const long long node = i / D;
  const int ci = (int)(i - node * D);
  const int r = (int)(node / nnx), c = (int)(node - (long long)r * nnx);
  const int r2 = r + nb / 3 - 1, c2 = c + nb % 3 - 1;
  if (r2 < 0 || r2 >= nny || c2 < 0 || c2 >= nnx) return;
  const long long j = ((long long)r2 * nnx + c2) * D + cj;
  const double* kf = keff != nullptr ? keff + row * sk : nullptr;
  const double* rh = rho != nullptr ? rho + row * sr : nullptr;
  double v = 0.5 * (dense_entry<D>(nelx, nely, r, c, ci, r2, c2, cj, kf, rh, ke) +
                    dense_entry<D>(nelx, nely, r2, c2, cj, r, c, ci, kf, rh, ke));
  if (i == j && rdiag != nullptr) v += rdiag[row * srd + i];
  const bool fi = mask_free(mask, i), fj = mask_free(mask, j);
  if (!(fi && fj)) v = (i == j && !fi) ? 1.0 : 0.0;
  out[row * so + i * n + j] = v;
}

}  // namespace stt

using namespace stt;

extern "C" {

int sttopt_galerkin(int32_t d, int32_t nelx, int32_t nely, int32_t kx, int32_t ky,
                    int32_t rows, const double* density, int64_t density_stride,
                    const double* reaction, int64_t reaction_stride, const double* g,
                    const double* gm, const double* fine_keff, int64_t fine_keff_stride,
                    const double* s, double* coarse_keff, int64_t coarse_keff_stride,
                    void* stream) {
  const bool from_density = fine_keff == nullptr;
  const int kmax = from_density ? 4096 : 3;  // block (density form) or coarsening factor
  if ((d != 1 && d != 2) || nelx <= 0 || nely <= 0 || kx < 1 || ky < 1 || kx > kmax ||
      ky > kmax || kx * ky > 4096 || nelx % kx || nely % ky || coarse_keff == nullptr ||
      rows < 0) {
    set_error("sttopt_galerkin: invalid arguments");
    return STTOPT_EINVAL;
  }
  if (rows == 0) return STTOPT_OK;
  if (from_density && (density == nullptr || g == nullptr || (reaction && !gm))) {
    set_error("sttopt_galerkin: level-1 build needs density and G (and GM with reaction)");
    return STTOPT_EINVAL;
  }
  if (!from_density && s == nullptr) {
    set_error("sttopt_galerkin: coarse build needs the interpolation matrices S");
    return STTOPT_EINVAL;
  }
  const long long cnel = (long long)(nelx / kx) * (nely / ky);
  const int K = 4 * d;
  hipStream_t st = (hipStream_t)stream;
  if (from_density) {
    // few coarse elements: narrower entry groups for more parallelism
    const bool wide = cnel * rows >= 65536;
    const int eg = d == 1 ? 16 : (wide ? 32 : 8);
    const dim3 pgrid(ceil_div(cnel, 256), (K * K) / eg, rows), pblock(256);
#define STT_GALP(D, EG)                                                                 \
  hipLaunchKernelGGL((galerkin_param_kernel<D, EG>), pgrid, pblock, 0, st, nelx, kx, ky, \
                     nelx / kx, cnel, density, density_stride, reaction, reaction_stride, g, gm, \
                     coarse_keff, coarse_keff_stride)
    if (d == 1) STT_GALP(1, 16);
    else if (wide) STT_GALP(2, 32);
    else STT_GALP(2, 8);
#undef STT_GALP
    STT_CHECK_LAUNCH("galerkin_param_kernel");
    return STTOPT_OK;
  }
  dim3 grid(ceil_div(cnel, 64), K * K, rows), block(64);
#define STT_GAL(D)                                                                \
  hipLaunchKernelGGL((galerkin_kernel<D>), grid, block, 0, st, nelx, nely, kx, ky,  \
                     fine_keff, fine_keff_stride, s, coarse_keff, coarse_keff_stride)
  if (d == 2) STT_GAL(2);
  else STT_GAL(1);
#undef STT_GAL
  STT_CHECK_LAUNCH("galerkin_kernel");
  return STTOPT_OK;
}

int sttopt_pack_keff(int32_t dtype, int32_t d, int64_t nel, int32_t rows, const double* keff,
                     int64_t keff_stride, void* out, int64_t out_stride, void* stream) {
  if ((d != 1 && d != 2) || nel <= 0 || rows < 0 || keff == nullptr || out == nullptr ||
      (dtype != STTOPT_F64 && dtype != STTOPT_F32)) {
    set_error("sttopt_pack_keff: invalid arguments");
    return STTOPT_EINVAL;
  }
  if (rows == 0) return STTOPT_OK;
  const dim3 grid(ceil_div(nel, 256), rows), block(256);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == STTOPT_F32) {
    float* o = static_cast<float*>(out);
    if (d == 2)
      hipLaunchKernelGGL((pack_keff_kernel<float, 8>), grid, block, 0, st, nel, keff, keff_stride, o, out_stride);
    else
      hipLaunchKernelGGL((pack_keff_kernel<float, 4>), grid, block, 0, st, nel, keff, keff_stride, o, out_stride);
  } else {
    double* o = static_cast<double*>(out);
    if (d == 2)
      hipLaunchKernelGGL((pack_keff_kernel<double, 8>), grid, block, 0, st, nel, keff, keff_stride, o, out_stride);
    else
      hipLaunchKernelGGL((pack_keff_kernel<double, 4>), grid, block, 0, st, nel, keff, keff_stride, o, out_stride);
  }
  STT_CHECK_LAUNCH("pack_keff_kernel");
  return STTOPT_OK;
}

int sttopt_node_stencil(int32_t dtype, int32_t d, int32_t nelx, int32_t nely, int32_t rows,
                        const double* keff, int64_t keff_stride, void* out, int64_t out_stride,
                        void* stream) {
  if ((d != 1 && d != 2) || nelx <= 0 || nely <= 0 || rows < 0 || keff == nullptr ||
      out == nullptr || (dtype != STTOPT_F64 && dtype != STTOPT_F32)) {
    set_error("sttopt_node_stencil: invalid arguments");
    return STTOPT_EINVAL;
  }
  if (rows == 0) return STTOPT_OK;
  const long long nn = (long long)(nelx + 1) * (nely + 1);
  const dim3 grid(ceil_div(nn, 256), rows), block(256);
  hipStream_t st = (hipStream_t)stream;
#define STT_NODE(TO, DD)                                                                \
  hipLaunchKernelGGL((node_stencil_kernel<TO, DD>), grid, block, 0, st, nelx, nely, keff, \
                     keff_stride, static_cast<TO*>(out), out_stride)
  if (dtype == STTOPT_F32) {
    if (d == 2) STT_NODE(float, 2);
    else STT_NODE(float, 1);
  } else {
    if (d == 2) STT_NODE(double, 2);
    else STT_NODE(double, 1);
  }
#undef STT_NODE
  STT_CHECK_LAUNCH("node_stencil_kernel");
  return STTOPT_OK;
}

int sttopt_keff_diagonal(int32_t dtype_out, int32_t d, int32_t nelx, int32_t nely,
                         int32_t rows, const double* keff, int64_t keff_stride,
                         const uint32_t* mask, int32_t inv, void* out, int64_t out_stride,
                         void* stream) {
  if ((d != 1 && d != 2) || nelx <= 0 || nely <= 0 || keff == nullptr || out == nullptr) {
    set_error("sttopt_keff_diagonal: invalid arguments");
    return STTOPT_EINVAL;
  }
  OpDesc op;
  op.d = d;
  op.nelx = nelx;
  op.nely = nely;
  op.keff = keff;
  op.sk = keff_stride;
  op.mask = mask;
  const int mode = inv ? 2 : 1;
  if (dtype_out == STTOPT_F64)
    return diagonal<double>(op, rows, mode, static_cast<double*>(out), out_stride,
                            (hipStream_t)stream);
  if (dtype_out == STTOPT_F32)
    return diagonal<float>(op, rows, mode, static_cast<float*>(out), out_stride,
                           (hipStream_t)stream);
  set_error("sttopt_keff_diagonal: bad dtype");
  return STTOPT_EINVAL;
}

}  // extern "C"

namespace stt {

// Per-position diagonals of the level-0 element matrices under one element of a level
// applied from the level-0 densities: v[(gy kx + gx) 4d + a], at most 4 positions.
struct ParamDiag {
  double v[4 * 8];
};

// Jacobi diagonal of the level (nelx x nely elements, kx x ky level-0 elements under
// each) whose element matrices are sum_g rho_g G_pos(g): one thread per node, the four
// elements around it in a fixed order, fp64; 1 at fixed dofs.
template <int D>
__global__ __launch_bounds__(256) void param_diagonal_kernel(int nelx, int nely, int kx, int ky,
                                                             const double* rho, long long sr,
                                                             ParamDiag pd, const uint32_t* mask,
                                                             double* out, long long so) {
  constexpr int K = 4 * D;
  const int nnx = nelx + 1, nny = nely + 1;
  const long long nnodes = (long long)nnx * nny;
  const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  if (node >= nnodes) return;
  const int r = (int)(node / nnx), c = (int)(node - (long long)r * nnx);
  const long long fnelx = (long long)nelx * kx;
  const double* rr = rho + row * sr;
  double acc[D];
#pragma unroll
  for (int i = 0; i < D; ++i) acc[i] = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int er = r - 1 + (q >> 1), ec = c - 1 + (q & 1);
    if (er < 0 || er >= nely || ec < 0 || ec >= nelx) continue;
    constexpr int locs[4] = {1, 0, 2, 3};  // the node's corner of that element
    const int loc = locs[q];
    for (int gy = 0; gy < ky; ++gy) {
      for (int gx = 0; gx < kx; ++gx) {
        const double rg = rr[(long long)(ky * er + gy) * fnelx + (long long)kx * ec + gx];
        const double* t = pd.v + (gy * kx + gx) * K + loc * D;
#pragma unroll
        for (int i = 0; i < D; ++i) acc[i] += rg * t[i];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const long long dof = node * D + i;
    out[row * so + dof] = mask_free(mask, dof) ? acc[i] : 1.0;
  }
}

}  // namespace stt

extern "C" {

int sttopt_param_diagonal(int32_t d, int32_t nelx, int32_t nely, int32_t kx, int32_t ky,
                          int32_t rows, const double* density, int64_t density_stride,
                          const double* table_diag, const uint32_t* mask, double* out,
                          int64_t out_stride, void* stream) {
  if ((d != 1 && d != 2) || nelx <= 0 || nely <= 0 || kx < 1 || ky < 1 || kx * ky > 4 ||
      rows < 0 || density == nullptr || table_diag == nullptr || mask == nullptr ||
      out == nullptr) {
    set_error("sttopt_param_diagonal: invalid arguments (d 1 or 2, at most 4 positions)");
    return STTOPT_EINVAL;
  }
  if (rows == 0) return STTOPT_OK;
  ParamDiag pd{};
  for (int i = 0; i < kx * ky * 4 * d; ++i) pd.v[i] = table_diag[i];
  const long long nnodes = (long long)(nelx + 1) * (nely + 1);
  const dim3 grid((unsigned)ceil_div(nnodes, 256), (unsigned)rows), block(256);
  hipStream_t s = (hipStream_t)stream;
  if (d == 2)
    hipLaunchKernelGGL(param_diagonal_kernel<2>, grid, block, 0, s, nelx, nely, kx, ky, density,
                       (long long)density_stride, pd, mask, out, (long long)out_stride);
  else
    hipLaunchKernelGGL(param_diagonal_kernel<1>, grid, block, 0, s, nelx, nely, kx, ky, density,
                       (long long)density_stride, pd, mask, out, (long long)out_stride);
  STT_CHECK_LAUNCH("param_diagonal_kernel");
  return STTOPT_OK;
}

}  // extern "C"

namespace stt {
template <typename TO>
__global__ __launch_bounds__(256) void invert_diag_kernel(long long n, const double* diag,
                                                          long long sd, const uint32_t* mask,
                                                          TO* out, long long so) {
  const int row = blockIdx.y;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[row * so + i] =
        mask_free(mask, i) ? static_cast<TO>(1.0 / diag[row * sd + i]) : TO(0);
}
}  // namespace stt

extern "C" {

int sttopt_invert_diagonal(int32_t dtype_out, int64_t n, int32_t rows, const double* diag,
                           int64_t diag_stride, const uint32_t* mask, void* out,
                           int64_t out_stride, void* stream) {
  if (diag == nullptr || out == nullptr || n < 0 || rows < 0) {
    set_error("sttopt_invert_diagonal: invalid arguments");
    return STTOPT_EINVAL;
  }
  if (n == 0 || rows == 0) return STTOPT_OK;
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  dim3 grid((unsigned)blocks, rows);
  if (dtype_out == STTOPT_F64)
    hipLaunchKernelGGL((invert_diag_kernel<double>), grid, dim3(256), 0, (hipStream_t)stream,
                       (long long)n, diag, diag_stride, mask, static_cast<double*>(out),
                       out_stride);
  else if (dtype_out == STTOPT_F32)
    hipLaunchKernelGGL((invert_diag_kernel<float>), grid, dim3(256), 0, (hipStream_t)stream,
                       (long long)n, diag, diag_stride, mask, static_cast<float*>(out),
                       out_stride);
  else {
    set_error("sttopt_invert_diagonal: bad dtype");
    return STTOPT_EINVAL;
  }
  STT_CHECK_LAUNCH("invert_diag_kernel");
  return STTOPT_OK;
}

int sttopt_dense_assemble(int32_t d, int32_t nelx, int32_t nely, int32_t rows,
                          const double* keff, int64_t keff_stride, const double* density,
                          int64_t density_stride, const double* rdiag, int64_t rdiag_stride,
                          const double* ke, const uint32_t* mask, double* out,
                          int64_t out_stride, void* stream) {
  if ((d != 1 && d != 2) || nelx <= 0 || nely <= 0 || out == nullptr ||
      (keff == nullptr && (density == nullptr || ke == nullptr))) {
    set_error("sttopt_dense_assemble: invalid arguments");
    return STTOPT_EINVAL;
  }
  if (rows == 0) return STTOPT_OK;
  const long long n = (long long)(nelx + 1) * (nely + 1) * d;
  hipStream_t st = (hipStream_t)stream;
  // ke is a device array here (the kernel reads it per element)
  if (int e = fill<double>(n * n, rows, 0.0, out, out_stride, nullptr, st)) return e;
  const long long threads = n * 9 * d;
  dim3 grid(ceil_div(threads, 256), rows), block(256);
  if (d == 2)
    hipLaunchKernelGGL(dense_assemble_kernel<2>, grid, block, 0, st, nelx, nely, keff,
                       keff_stride, density, density_stride, rdiag, rdiag_stride, ke, mask,
                       out, out_stride);
  else
    hipLaunchKernelGGL(dense_assemble_kernel<1>, grid, block, 0, st, nelx, nely, keff,
                       keff_stride, density, density_stride, rdiag, rdiag_stride, ke, mask,
                       out, out_stride);
  STT_CHECK_LAUNCH("dense_assemble_kernel");
  return STTOPT_OK;
}

int sttopt_restrict(int32_t dtype, int32_t d, int32_t cnelx, int32_t cnely, int32_t kx,
                    int32_t ky, int32_t rows, const void* fine, int64_t fine_stride,
                    const uint32_t* cmask, void* coarse, int64_t coarse_stride,
                    void* stream) {
  if ((d != 1 && d != 2) || cnelx <= 0 || cnely <= 0 || kx < 1 || ky < 1 || !fine ||
      !coarse) {
    set_error("sttopt_restrict: invalid arguments");
    return STTOPT_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  if (dtype == STTOPT_F64)
    return restrict_to<double>(d, cnelx, cnely, kx, ky, rows,
                               static_cast<const double*>(fine), fine_stride, cmask,
                               static_cast<double*>(coarse), coarse_stride, nullptr, s);
  if (dtype == STTOPT_F32)
    return restrict_to<float>(d, cnelx, cnely, kx, ky, rows, static_cast<const float*>(fine),
                              fine_stride, cmask, static_cast<float*>(coarse), coarse_stride,
                              nullptr, s);
  set_error("sttopt_restrict: bad dtype");
  return STTOPT_EINVAL;
}

int sttopt_prolong_add(int32_t dtype, int32_t d, int32_t cnelx, int32_t cnely, int32_t kx,
                       int32_t ky, int32_t rows, const void* coarse, int64_t coarse_stride,
                       const uint32_t* fmask, void* fine, int64_t fine_stride, void* stream) {
  if ((d != 1 && d != 2) || cnelx <= 0 || cnely <= 0 || kx < 1 || ky < 1 || !fine ||
      !coarse) {
    set_error("sttopt_prolong_add: invalid arguments");
    return STTOPT_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  if (dtype == STTOPT_F64)
    return prolong_add<double>(d, cnelx, cnely, kx, ky, rows,
                               static_cast<const double*>(coarse), coarse_stride, fmask,
                               static_cast<double*>(fine), fine_stride, nullptr, s);
  if (dtype == STTOPT_F32)
    return prolong_add<float>(d, cnelx, cnely, kx, ky, rows,
                              static_cast<const float*>(coarse), coarse_stride, fmask,
                              static_cast<float*>(fine), fine_stride, nullptr, s);
  set_error("sttopt_prolong_add: bad dtype");
  return STTOPT_EINVAL;
}

}  // extern "C"
