// Batched in-place inverse of dense SPD matrices (the coarsest multigrid operator).
//
// The reference factors the coarsest operator with torch.linalg.cholesky
// (torch_mg.py:269-290) and applies it with two triangular solves per V-cycle
// (torch_mg.py:448-451).  Here the coarse solve is a dense matvec with the inverse, and
// the inverse is formed by the block symmetric sweep (Gauss-Jordan elimination without
// pivoting, stable for SPD input): for pivot block K with D = A_KK^-1,
//
//   A_IJ <- A_IJ - A_IK D A_KJ   (I, J != K)      A_KJ <- D A_KJ,  A_IK <- A_IK D
//   A_KK <- -D
//
// After every block has been swept the matrix holds -A^-1; the last step writes with
// the sign flipped.  The matrix stays symmetric throughout, so only its lower block
// triangle is kept up to date: the update computes and writes the lower tiles, and the
// last step also writes each tile's mirror, leaving the whole inverse.
//
// A banded matrix (entries zero beyond a half-bandwidth b) keeps every tile more than
// c = (b + 63) / 64 blocks from the diagonal at zero until the sweep reaches it: at step K
// the tiles (I, J) with max(I, J) > K + c see A_IK = 0 or A_KJ = 0, so neither their
// update nor the pivot row and column change them.  Step K therefore works on the
// leading (K + c + 1) blocks only (the coarsest operator of a Q4 grid with nnx nodes
// per row has b = d (nnx + 1) + d - 1, two blocks at 33 x 17 nodes: about half the
// tiles of the full sweep).
//
// Per pivot block (64 wide): sweep_panel_kernel forms P = D R from the row panel
// R = A_K: (its tiles right of the diagonal read as the transposes of the lower tiles
// below it), and sweep_update_kernel writes every lower tile from A, R, P and D -- it
// reads no tile that another workgroup of the same launch writes.  D of the first block comes from
// sweep_pivot_kernel; the update launch of step K inverts the updated diagonal block
// K + 1 in the workgroup that writes it (dispatched first), so the next pivot block is
// ready when step K ends.
#include "stt_kernels.h"

namespace stt {

constexpr int kSB = 64;  // sweep block

struct SweepWork {
  double* D[2];  // rows x kSB x kSB each: step K uses D[K & 1]
  double* R;     // rows x kSB x n
  double* P;     // rows x kSB x n
  int* bad;      // one flag
};

static SweepWork carve_sweep(double* work, int n, int rows) {
  SweepWork w;
  w.D[0] = work;
  w.D[1] = w.D[0] + (long long)rows * kSB * kSB;
  w.R = w.D[1] + (long long)rows * kSB * kSB;
  w.P = w.R + (long long)rows * kSB * n;
  w.bad = reinterpret_cast<int*>(w.P + (long long)rows * kSB * n);
  return w;
}

constexpr int kPB = 4;  // pivots swept per barrier in the register sweep

// Sweep of a bk x bk block held in registers by a 256-thread block: thread t holds
// column j = t % 64 of rows t / 64 + 4 q (q < 16), padded with the identity past bk.
// On return m holds -(block)^-1; the result is false if a pivot was not positive (or
// not finite).  Every thread sees the same pivots.
// The pivots are swept kPB at a time (the block form of the sweep, as between the
// 64-wide blocks): the lanes of pivot set P publish their columns through LDS (col:
// two buffers of 64 x kPB, one barrier per set), every thread sweeps the kPB x kPB
// block M = A_PP in registers (M <- -A_PP^-1, pivot by pivot), and then, with
// u = -M v where v = A_Pj off P and v = -e_j on it,
//   A_ij <- keep A_ij - A_iP u   (i not in P; keep = 0 in the columns of P)
//   A_ij <- u_(i - p0)           (i in P)
// The barrier, the LDS round trip and the divisions are on the critical path of every
// sweep launch (one wave per SIMD, nothing to hide their latency), hence the sets.
// Sweeping the identity padding past bk leaves the real entries unchanged.
__device__ bool sweep_block_registers(double (&m)[16], int bk, double (*col)[kSB][kPB]) {
  const int j = threadIdx.x & (kSB - 1);
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x / kSB);
  bool bad_pivot = false;
  for (int p0 = 0; p0 < bk; p0 += kPB) {
    double (*c)[kPB] = col[(p0 / kPB) & 1];
    const int jj = j - p0;
    const bool colp = jj >= 0 && jj < kPB;
    if (colp) {
#pragma unroll
      for (int q = 0; q < 16; ++q) c[g + 4 * q][jj] = m[q];
    }
    __syncthreads();
    double M[kPB][kPB], v[kPB];
#pragma unroll
    for (int a = 0; a < kPB; ++a) {
#pragma unroll
      for (int b = 0; b < kPB; ++b) M[a][b] = c[p0 + a][b];
      v[a] = colp ? (a == jj ? -1.0 : 0.0) : c[j][a];  // A_Pj = A_jP (symmetric)
    }
#pragma unroll
    for (int a = 0; a < kPB; ++a) {
      const double piv = M[a][a];
      const bool ok = piv > 0.0 && isfinite(piv);
      bad_pivot |= !ok;
      const double d = ok ? 1.0 / piv : 0.0;
#pragma unroll
      for (int r = 0; r < kPB; ++r) {
        if (r == a) continue;
        const double dra = d * M[r][a];
#pragma unroll
        for (int t = 0; t < kPB; ++t) {
          if (t != a) M[r][t] = fma(-dra, M[a][t], M[r][t]);
        }
      }
#pragma unroll
      for (int r = 0; r < kPB; ++r) {
        if (r != a) {
          M[r][a] *= d;
          M[a][r] = M[r][a];
        }
      }
      M[a][a] = -d;
    }
    double u[kPB];
#pragma unroll
    for (int a = 0; a < kPB; ++a) {
      double acc = 0.0;
#pragma unroll
      for (int b = 0; b < kPB; ++b) acc = fma(-M[a][b], v[b], acc);
      u[a] = acc;
    }
    const double keep = colp ? 0.0 : 1.0;
    // The rows' LDS reads in two halves of eight rows: their offset (zero) passes
    // through an empty asm that takes u, then the first half's last result, so the
    // compiler cannot issue all 64 reads ahead of the block's sweep (that would take the
    // look-ahead workgroup of the update launch past four waves per SIMD).
    int hoff = 0;
    asm volatile("" : "+v"(hoff) : "v"(u[kPB - 1]));
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h == 1) asm volatile("" : "+v"(hoff) : "v"(m[7]));
#pragma unroll
      for (int qq = 0; qq < 8; ++qq) {
        const int q = 8 * h + qq;
        const double* ci = c[g + 4 * q + hoff];
        double t = m[q] * keep;
#pragma unroll
        for (int a = 0; a < kPB; ++a) t = fma(-ci[a], u[a], t);
        m[q] = t;
      }
    }
    // rows of P: row p0 + a sits in wave (p0 + a) % 4 at q = (p0 + a) / 4
#pragma unroll
    for (int a = 0; a < kPB; ++a) {
      const int r = p0 + a;
      if ((r & 3) == g) {
        const int qp = r >> 2;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          if (q == qp) m[q] = u[a];
        }
      }
    }
  }
  return !bad_pivot;
}

// D = A_KK^-1 of the first pivot block (later ones come from the update launch).
__global__ __launch_bounds__(256) void sweep_pivot_kernel(int n, int k0, int bk,
                                                          const double* __restrict__ a,
                                                          long long sa, double* __restrict__ D,
                                                          int* __restrict__ bad) {
  __shared__ double col[2][kSB][kPB];
  const int row = blockIdx.x;
  const double* A = a + row * sa;
  const int j = threadIdx.x & (kSB - 1), g = threadIdx.x / kSB;
  double m[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int i = g + 4 * q;
    m[q] = (i < bk && j < bk) ? A[(long long)(k0 + i) * n + (k0 + j)] : (i == j ? 1.0 : 0.0);
  }
  if (!sweep_block_registers(m, bk, col) && threadIdx.x == 0) *bad = 1;
  double* Dr = D + (long long)row * kSB * kSB;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int i = g + 4 * q;
    Dr[i * kSB + j] = (i < bk && j < bk) ? -m[q] : 0.0;
  }
}

constexpr int kKC = 16;  // k rows per LDS chunk

// The 64 x 64 tile of a 256-thread block is computed on the fp64 matrix cores: wave w
// owns the 32 x 32 quadrant (w / 2, w % 2) as 2 x 2 products of 16 x 16 blocks
// (v_mfma_f64_16x16x4f64: lane l supplies A(l % 16, l / 16) and B(l / 16, l % 16) of a
// 16 x 4 by 4 x 16 step and holds D(l / 16 + 4 r, l % 16), r < 4).  Entry q of a
// thread's 16 results (q = 8 ib + 4 jb + r) is tile element (tile_row(q), tile_col(q)).
typedef double d4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int tile_row(int q) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  return 32 * (w >> 1) + 16 * (q >> 3) + (l >> 4) + 4 * (q & 3);
}
__device__ __forceinline__ int tile_col(int q) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  return 32 * (w & 1) + 16 * ((q >> 2) & 1) + (l & 15);
}

// out[q] += sum_{k < bk} X(k, tile_row(q)) Y(k, tile_col(q)) for the 64 x 64 tile.  X
// and Y are k-major (row k at X + k ldx) with xcols / ycols valid columns; they pass
// through LDS 16 rows at a time (zero past bk and past the valid columns), the next
// chunk's loads in flight while the current one is multiplied.  Returns after a
// barrier-free last chunk: callers that reuse sX / sY synchronize first.
__device__ __forceinline__ void gemm_tile_kmajor(const double* __restrict__ X, long long ldx,
                                                 int xcols, const double* __restrict__ Y,
                                                 long long ldy, int ycols, int bk,
                                                 double (&out)[16], double* sX, double* sY) {
  constexpr int kPer = kKC * kSB / 256;  // elements of each chunk per thread
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int ra = 32 * (w >> 1) + (l & 15), cb = 32 * (w & 1) + (l & 15), kl = l >> 4;
  d4_t acc[2][2];
#pragma unroll
  for (int ib = 0; ib < 2; ++ib) {
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[ib][jb][r] = out[8 * ib + 4 * jb + r];
    }
  }
  double nx[kPer], ny[kPer];
  auto fetch = [&](int kc) {
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int e = threadIdx.x + r * 256;
      const int k = e / kSB, c = e & (kSB - 1);
      const bool kin = kc + k < bk;
      nx[r] = (kin && c < xcols) ? X[(long long)(kc + k) * ldx + c] : 0.0;
      ny[r] = (kin && c < ycols) ? Y[(long long)(kc + k) * ldy + c] : 0.0;
    }
  };
  fetch(0);
  for (int kc = 0; kc < bk; kc += kKC) {
    __syncthreads();  // the previous chunk is consumed
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      sX[threadIdx.x + r * 256] = nx[r];
      sY[threadIdx.x + r * 256] = ny[r];
    }
    __syncthreads();
    if (kc + kKC < bk) fetch(kc + kKC);
#pragma unroll
    for (int kk = 0; kk < kKC; kk += 4) {
      const double* xr = sX + (kk + kl) * kSB;
      const double* yr = sY + (kk + kl) * kSB;
      const double a0 = xr[ra], a1 = xr[ra + 16];
      const double b0 = yr[cb], b1 = yr[cb + 16];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
#pragma unroll
  for (int ib = 0; ib < 2; ++ib) {
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
#pragma unroll
      for (int r = 0; r < 4; ++r) out[8 * ib + 4 * jb + r] = acc[ib][jb][r];
    }
  }
}

// R = A[K rows, column block J] (copy) and P = D R for column block J = blockIdx.x.
// Only the lower block triangle of A is current: a block J > K of the panel is the
// transpose of the lower tile (J, K), transposed through LDS (coalesced on both sides).
__global__ __launch_bounds__(256) void sweep_panel_kernel(int n, int k0, int bk,
                                                          const double* __restrict__ a,
                                                          long long sa,
                                                          const double* __restrict__ D,
                                                          double* __restrict__ R,
                                                          double* __restrict__ P) {
  __shared__ __align__(16) double sX[kKC * kSB];
  __shared__ __align__(16) double sY[kKC * kSB];
  __shared__ double sT[kSB][kSB + 1];
  const int row = blockIdx.y;
  const int j0 = blockIdx.x * kSB;
  const int bj = min(kSB, n - j0);
  const double* A = a + row * sa;
  double* Rr = R + (long long)row * kSB * n + j0;
  double* Pr = P + (long long)row * kSB * n + j0;
  // the panel block as a k-major operand (row k at Y + k n)
  const double* Y;
  if (j0 > k0) {
    // R(k, c) = A(j0 + c, k0 + k): rows of the lower tile into LDS, then rows of R out
    for (int e = threadIdx.x; e < kSB * kSB; e += 256) {
      const int c = e / kSB, k = e & (kSB - 1);
      sT[c][k] = (k < bk && c < bj) ? A[(long long)(j0 + c) * n + (k0 + k)] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kSB * kSB; e += 256) {
      const int k = e / kSB, c = e & (kSB - 1);
      if (k < bk && c < bj) Rr[(long long)k * n + c] = sT[c][k];
    }
    __syncthreads();  // the block's writes of R are visible to its own reads below
    Y = Rr;
  } else {
    const double* Ak = A + (long long)k0 * n + j0;  // row k of the panel: Ak + k n
    for (int e = threadIdx.x; e < kSB * kSB; e += 256) {
      const int k = e / kSB, c = e & (kSB - 1);
      if (k < bk && c < bj) Rr[(long long)k * n + c] = Ak[(long long)k * n + c];
    }
    Y = Ak;
  }
  double out[16] = {};
  // P(i, j) = sum_k D(i, k) R(k, j); D is symmetric, so its rows serve as X(k, i)
  gemm_tile_kmajor(D + (long long)row * kSB * kSB, kSB, bk, Y, n, bj, bk, out, sX, sY);
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int i = tile_row(q), j = tile_col(q);
    if (i < bk && j < bj) Pr[(long long)i * n + j] = out[q];
  }
}

// Lower-triangle tile t = (I, J), I >= J, of matrix blockIdx.x: every lower tile is
// written once per sweep step, and in the last step also, off the diagonal, its mirror
// (through an LDS transpose).  A workgroup reads only its own tile of A, and no other
// workgroup writes that tile or reads the mirror.
// Four waves per SIMD (the LDS limit; without the hint the compiler settles for three).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void
sweep_update_kernel(int n, int kb, int k0, int bk, double* __restrict__ a, long long sa,
                    const double* __restrict__ D, const double* __restrict__ R,
                    const double* __restrict__ P, double sign, int ntile,
                    double* __restrict__ Dnext, int* __restrict__ bad) {
  // the k-chunks of the product and, afterwards, the transpose of the mirror tile
  __shared__ __align__(16) double lds[kSB * (kSB + 1)];
  double* sX = lds;
  double* sY = lds + kKC * kSB;
  double (*sT)[kSB + 1] = reinterpret_cast<double (*)[kSB + 1]>(lds);
  __shared__ double col[2][kSB][kPB];
  // grid (rows, tiles): the matrices are the fast dimension, so tile slot 0 of every
  // matrix is dispatched first; that slot is the next pivot block's diagonal tile,
  // whose workgroup also inverts the block
  const int row = blockIdx.x;
  const int kn = kb + 1;
  const int tnext = kn * (kn + 1) / 2 + kn;
  const bool ahead = Dnext != nullptr;
  const int t = ahead ? (int)((blockIdx.y + tnext) % ntile) : (int)blockIdx.y;
  int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) ++I;
  while (I * (I + 1) / 2 > t) --I;
  const int J = t - I * (I + 1) / 2;
  const int i0 = I * kSB, j0 = J * kSB;
  const int bi = min(kSB, n - i0), bj = min(kSB, n - j0);
  double* A = a + row * sa;
  const double* Rr = R + (long long)row * kSB * n;
  const double* Pr = P + (long long)row * kSB * n;
  const double* Dr = D + (long long)row * kSB * kSB;
  double out[16] = {};
  if (I != kb && J != kb) {
    // A_IJ - R_KI^T P_KJ  (R_KI^T = A_IK before this step)
    gemm_tile_kmajor(Rr + i0, n, bi, Pr + j0, n, bj, bk, out, sX, sY);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int i = tile_row(q), j = tile_col(q);
      out[q] = (i < bi && j < bj) ? A[(long long)(i0 + i) * n + (j0 + j)] - out[q] : 0.0;
    }
  } else if (I == kb && J == kb) {
#pragma unroll
    for (int q = 0; q < 16; ++q) out[q] = -Dr[tile_row(q) * kSB + tile_col(q)];
  } else if (I == kb) {
    // (K, J) = P_KJ
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int i = tile_row(q), j = tile_col(q);
      out[q] = (i < bi && j < bj) ? Pr[(long long)i * n + j0 + j] : 0.0;
    }
  } else {
    // (I, K) = A_IK D = P_KI^T
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int i = tile_row(q), j = tile_col(q);
      out[q] = (i < bi && j < bj) ? Pr[(long long)j * n + i0 + i] : 0.0;
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int i = tile_row(q), j = tile_col(q);
    out[q] *= sign;
    if (i < bi && j < bj) A[(long long)(i0 + i) * n + (j0 + j)] = out[q];
  }
  if (ahead && I == kn && J == kn) {
    // look-ahead: D_{K+1} = (this updated tile)^-1, re-laid out through LDS into the
    // pivot layout (column t % 64, rows t / 64 + 4 q)
    __syncthreads();  // the product's last chunk is consumed before sT overwrites it
#pragma unroll
    for (int q = 0; q < 16; ++q) sT[tile_row(q)][tile_col(q)] = out[q];
    __syncthreads();
    const int jj = threadIdx.x & (kSB - 1), g = threadIdx.x / kSB;
    double m[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int i = g + 4 * q;
      m[q] = (i < bi && jj < bj) ? sT[i][jj] : (i == jj ? 1.0 : 0.0);
    }
    if (!sweep_block_registers(m, bi, col) && threadIdx.x == 0) *bad = 1;
    double* Dr2 = Dnext + (long long)row * kSB * kSB;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int i = g + 4 * q;
      Dr2[i * kSB + jj] = (i < bi && jj < bj) ? -m[q] : 0.0;
    }
    return;  // a diagonal tile: no mirror to write
  }
  if (I != J && sign < 0.0) {
    // the last step: mirror tile (J, I), element (r, c) = out(c, r), written row by row
    __syncthreads();  // the product's last chunk is consumed before sT overwrites it
#pragma unroll
    for (int q = 0; q < 16; ++q) sT[tile_row(q)][tile_col(q)] = out[q];
    __syncthreads();
    const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int r = 4 * ti + x;
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        const int c = 4 * tj + y;
        if (r < bj && c < bi) A[(long long)(j0 + r) * n + (i0 + c)] = sT[c][r];
      }
    }
  }
}

}  // namespace stt

using namespace stt;

extern "C" {

int64_t sttopt_spd_inverse_work_count(int32_t n, int32_t rows) {
  if (n <= 0 || rows <= 0) return 0;
  // two D, R, P in doubles plus one double slot for the flag
  return 2LL * rows * kSB * kSB + 2LL * rows * kSB * n + 1;
}

int sttopt_spd_inverse_band_async(int32_t n, int32_t band, int32_t rows, double* a,
                                  int64_t a_stride, double* work, int32_t* not_spd,
                                  void* stream) {
  if (n <= 0 || rows < 0 || band < 0 || a == nullptr || work == nullptr ||
      not_spd == nullptr || (rows > 1 && a_stride < (int64_t)n * n)) {
    set_error("sttopt_spd_inverse: invalid arguments");
    return STTOPT_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  if (int e = fill<int>(1, 1, 0, not_spd, 1, nullptr, s)) return e;
  if (rows == 0) return STTOPT_OK;
  SweepWork w = carve_sweep(work, n, rows);
  w.bad = not_spd;
  const int nb = (n + kSB - 1) / kSB;
  // blocks off the diagonal that the band reaches; at least one, so that step K always
  // updates diagonal block K + 1, whose workgroup inverts it for the next step
  const int reach = max(1, (int)((min(band, n - 1) + (int64_t)kSB - 1) / kSB));
  hipLaunchKernelGGL(sweep_pivot_kernel, dim3(rows), dim3(256), 0, s, n, 0,
                     n < kSB ? n : kSB, a, a_stride, w.D[0], w.bad);
  for (int kb = 0; kb < nb; ++kb) {
    const int k0 = kb * kSB, bk = n - k0 < kSB ? n - k0 : kSB;
    const int lead = min(nb, kb + reach + 1);  // blocks this step changes
    const int ntile = lead * (lead + 1) / 2;
    double* Dk = w.D[kb & 1];
    double* Dn = kb + 1 < nb ? w.D[(kb + 1) & 1] : nullptr;
    hipLaunchKernelGGL(sweep_panel_kernel, dim3(lead, rows), dim3(256), 0, s, n, k0, bk, a,
                       a_stride, Dk, w.R, w.P);
    hipLaunchKernelGGL(sweep_update_kernel, dim3(rows, ntile), dim3(256), 0, s, n, kb, k0,
                       bk, a, a_stride, Dk, w.R, w.P, kb == nb - 1 ? -1.0 : 1.0, ntile, Dn,
                       w.bad);
    STT_CHECK_LAUNCH("sweep");
  }
  return STTOPT_OK;
}

int sttopt_spd_inverse_band(int32_t n, int32_t band, int32_t rows, double* a,
                            int64_t a_stride, double* work, void* stream) {
  if (n <= 0 || rows < 0 || work == nullptr) {
    set_error("sttopt_spd_inverse: invalid arguments");
    return STTOPT_EINVAL;
  }
  if (rows == 0) return STTOPT_OK;
  int* bad = carve_sweep(work, n, rows).bad;
  if (int e = sttopt_spd_inverse_band_async(n, band, rows, a, a_stride, work, bad, stream))
    return e;
  int v = 0;
  if (int e = read_device_int(bad, (hipStream_t)stream, &v)) return e;
  if (v) {
    set_error("sttopt_spd_inverse: a matrix is not positive definite");
    return STTOPT_ENOTSPD;
  }
  return STTOPT_OK;
}

int sttopt_spd_inverse(int32_t n, int32_t rows, double* a, int64_t a_stride, double* work,
                       void* stream) {
  return sttopt_spd_inverse_band(n, n > 0 ? n - 1 : 0, rows, a, a_stride, work, stream);
}

}  // extern "C"
