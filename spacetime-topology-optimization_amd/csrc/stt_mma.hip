// The MMA subproblem (Svanberg's mmasub / subsolv, sttopt/mma.py:150-557) on the
// device, for m constraints over n design variables with m << n.
//
// The subproblem data P, Q are [m][n] row-major.  Every n-sized quantity of the
// primal-dual Newton method lives on the device and is produced by column kernels (one
// thread per variable j, looping over the m constraint rows where a column of P or Q is
// needed: coalesced across j).  The m-sized reductions over n are row GEMVs
// (sum_j P_ij A_j + Q_ij B_j, one workgroup per row and column chunk), and the m x m
// matrix GG diag(1/diagx) GG^T of the m < n elimination is a split-n Gram product with
// 4 x 4 register blocks over the lower triangle.  The bordered (m+1) x (m+1) Newton
// system is factored by one workgroup in LDS (Gaussian elimination with partial
// pivoting, as torch.linalg.solve), and the m-sized updates run in that same
// workgroup.  The host drives the barrier and Newton loops exactly as the reference
// does and reads back two numbers per line-search candidate (the residual's norm and
// maximum), which is what decides the loop.
#include <cfloat>
#include <cmath>

#include "stt_kernels.h"

namespace stt {

namespace {

constexpr int kMmaMaxM = 126;  // bordered system (m + 1) <= 127 rows in LDS
constexpr int kCol = 256;      // column-kernel block
constexpr int kColBlocks = 1024;

// Device workspace of one subsolv call (doubles), carved by the host binding in this
// order.  Column partials: [kColBlocks][kColSlots].
enum ColSlot { kSumSq = 0, kMaxAbs = 1, kMaxAlfa = 2, kMaxBeta = 3, kMaxXX = 4, kColSlots = 5 };

struct MmaDev {
  int m, n;
  // problem
  const double *low, *upp, *alfa, *beta, *p0, *q0, *P, *Q, *a, *b, *c, *d;
  double a0;
  // primal-dual state: n-vectors, m-vectors, scalars
  double *x, *xsi, *eta;
  double *y, *lam, *mu, *s;
  double* zz;  // [z, zet, dz, dzet, steg, norm, max, epsi-scaled scratch]
  // Newton directions and column data
  double *dx, *dxsi, *deta, *dy, *dlam, *dmu, *ds;
  double *delx, *diagx, *uxinv1, *xlinv1, *uxinv2, *xlinv2, *PTl, *QTl, *PTd, *QTd, *wA, *wB;
  double *uxc, *xlc;  // 1 / (upp - x') and 1 / (x' - low) at a line-search candidate
  // reductions
  double* colpart;   // [kColBlocks][kColSlots]
  double* rowpart;   // [2][m][nchunk]
  double* grampart;  // [gchunk][m][m]
  double* alam;      // [m][m] lower triangle
  double* gvec;      // [m]
  double* ggd;       // [m]
  double* out;       // [2]: norm, max of the last residual
  int nchunk, gchunk, colblk;
};

enum { kZ = 0, kZet = 1, kDz = 2, kDzet = 3, kSteg = 4 };

__device__ __forceinline__ double block_max(double v, double* lds) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, kWave));
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double r = -INFINITY;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i) r = fmax(r, lds[i]);
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------- columns ----

enum ColMode { kColInit = 0, kColPass1 = 1, kColStep = 2, kColResid = 3, kColApply = 4,
               kColStepGivenDx = 5 };

// t for a line-search candidate: steg / 2^jj (steg from the device); jj < 0: 0, the
// current point
__device__ __forceinline__ double cand_t(const MmaDev& D, int jj) {
  return jj < 0 ? 0.0 : D.zz[kSteg] * ldexp(1.0, -jj);
}

template <int MODE>
__global__ __launch_bounds__(kCol) void mma_col_kernel(MmaDev D, double epsi, int jj) {
  __shared__ double sv[kMmaMaxM + 2];
  __shared__ double red[kCol / kWave];
  const int m = D.m, n = D.n;
  if (MODE == kColPass1 || MODE == kColStep || MODE == kColStepGivenDx) {
    const double* v = MODE == kColPass1 ? D.lam : D.dlam;
    for (int i = threadIdx.x; i < m; i += blockDim.x) sv[i] = v[i];
    __syncthreads();
  }
  const double t = (MODE == kColResid || MODE == kColApply) ? cand_t(D, jj) : 0.0;
  double sumsq = 0.0, mabs = 0.0, malfa = -INFINITY, mbeta = -INFINITY, mxx = -INFINITY;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (long long)gridDim.x * blockDim.x) {
    const double lo = D.low[j], up = D.upp[j], al = D.alfa[j], be = D.beta[j];
    if constexpr (MODE == kColInit) {
      const double x = 0.5 * (al + be);
      D.x[j] = x;
      D.xsi[j] = fmax(1.0 / (x - al), 1.0);
      D.eta[j] = fmax(1.0 / (be - x), 1.0);
      D.dx[j] = D.dxsi[j] = D.deta[j] = 0.0;
      D.PTd[j] = D.QTd[j] = 0.0;
      // P^T lam and Q^T lam at the initial lam = 1
      double pl = 0.0, ql = 0.0;
      for (int i = 0; i < m; ++i) {
        pl += D.P[(long long)i * n + j];
        ql += D.Q[(long long)i * n + j];
      }
      D.PTl[j] = pl;
      D.QTl[j] = ql;
    } else if constexpr (MODE == kColPass1) {
      const double x = D.x[j], xsi = D.xsi[j], eta = D.eta[j];
      double pl = 0.0, ql = 0.0;
#pragma unroll 8
      for (int i = 0; i < m; ++i) {
        pl += D.P[(long long)i * n + j] * sv[i];
        ql += D.Q[(long long)i * n + j] * sv[i];
      }
      const double ux1 = up - x, xl1 = x - lo;
      const double ux2 = ux1 * ux1, xl2 = xl1 * xl1;
      const double ux3 = ux1 * ux2, xl3 = xl1 * xl2;
      const double plam = D.p0[j] + pl, qlam = D.q0[j] + ql;
      const double dpsidx = plam / ux2 - qlam / xl2;
      const double delx = dpsidx - epsi / (x - al) + epsi / (be - x);
      double diagx = plam / ux3 + qlam / xl3;
      diagx = 2 * diagx + xsi / (x - al) + eta / (be - x);
      D.PTl[j] = pl;
      D.QTl[j] = ql;
      D.delx[j] = delx;
      D.diagx[j] = diagx;
      D.uxinv1[j] = 1.0 / ux1;
      D.xlinv1[j] = 1.0 / xl1;
      const double ui2 = 1.0 / ux2, xi2 = 1.0 / xl2;
      D.uxinv2[j] = ui2;
      D.xlinv2[j] = xi2;
      const double q = delx / diagx;
      D.wA[j] = ui2 * q;
      D.wB[j] = -xi2 * q;
    } else if constexpr (MODE == kColStep || MODE == kColStepGivenDx) {
      const double x = D.x[j], xsi = D.xsi[j], eta = D.eta[j];
      double pd = 0.0, qd = 0.0;
#pragma unroll 8
      for (int i = 0; i < m; ++i) {
        pd += D.P[(long long)i * n + j] * sv[i];
        qd += D.Q[(long long)i * n + j] * sv[i];
      }
      D.PTd[j] = pd;
      D.QTd[j] = qd;
      double dx;
      if constexpr (MODE == kColStep) {
        const double dg = D.diagx[j];
        dx = -D.delx[j] / dg - (D.uxinv2[j] * pd - D.xlinv2[j] * qd) / dg;
        D.dx[j] = dx;
      } else {
        dx = D.dx[j];  // from the m >= n solve
      }
      const double dxsi = -xsi + epsi / (x - al) - (xsi * dx) / (x - al);
      const double deta = -eta + epsi / (be - x) + (eta * dx) / (be - x);
      D.dxsi[j] = dxsi;
      D.deta[j] = deta;
      malfa = fmax(malfa, -1.01 * dx / (x - al));
      mbeta = fmax(mbeta, 1.01 * dx / (be - x));
      mxx = fmax(mxx, fmax(-1.01 * dxsi / xsi, -1.01 * deta / eta));
    } else if constexpr (MODE == kColResid) {
      const double x = D.x[j] + t * D.dx[j];
      const double xsi = D.xsi[j] + t * D.dxsi[j];
      const double eta = D.eta[j] + t * D.deta[j];
      const double ux1 = up - x, xl1 = x - lo;
      const double ux2 = ux1 * ux1, xl2 = xl1 * xl1;
      const double plam = D.p0[j] + (D.PTl[j] + t * D.PTd[j]);
      const double qlam = D.q0[j] + (D.QTl[j] + t * D.QTd[j]);
      const double dpsidx = plam / ux2 - qlam / xl2;
      const double rex = dpsidx - xsi + eta;
      const double rexsi = xsi * (x - al) - epsi;
      const double reeta = eta * (be - x) - epsi;
      sumsq += rex * rex + rexsi * rexsi + reeta * reeta;
      // fmax ignores a NaN operand: propagate it as the reference's max does
      const double am = fmax(fabs(rex), fmax(fabs(rexsi), fabs(reeta)));
      mabs = (isnan(rex) || isnan(rexsi) || isnan(reeta) || isnan(mabs)) ? NAN : fmax(mabs, am);
      D.uxc[j] = 1.0 / ux1;
      D.xlc[j] = 1.0 / xl1;
    } else if constexpr (MODE == kColApply) {
      D.x[j] += t * D.dx[j];
      D.xsi[j] += t * D.dxsi[j];
      D.eta[j] += t * D.deta[j];
      // P^T lam follows lam += t dlam (the next Newton pass recomputes it)
      D.PTl[j] += t * D.PTd[j];
      D.QTl[j] += t * D.QTd[j];
    }
  }
  if constexpr (MODE == kColResid) {
    const double s1 = block_sum(sumsq, red);
    __syncthreads();
    const double s2 = block_max(mabs, red);
    if (threadIdx.x == 0) {
      D.colpart[blockIdx.x * kColSlots + kSumSq] = s1;
      D.colpart[blockIdx.x * kColSlots + kMaxAbs] = s2;
    }
  } else if constexpr (MODE == kColStep || MODE == kColStepGivenDx) {
    const double v1 = block_max(malfa, red);
    const double v2 = block_max(mbeta, red);
    const double v3 = block_max(mxx, red);
    if (threadIdx.x == 0) {
      D.colpart[blockIdx.x * kColSlots + kMaxAlfa] = v1;
      D.colpart[blockIdx.x * kColSlots + kMaxBeta] = v2;
      D.colpart[blockIdx.x * kColSlots + kMaxXX] = v3;
    }
  }
}

// ------------------------------------------------------------------- rows ----

// part[v][i][chunk] = sum over the chunk's columns j of P_ij A_v[j] + Q_ij B_v[j], for
// the kGemvRows rows i = blockIdx.y kGemvRows + q: the column values are read once for
// all of them; each row's sum runs in the same order as with one row per block.
constexpr int kGemvRows = 4;
__global__ __launch_bounds__(256) void mma_gemv_kernel(int m, int n, const double* __restrict__ P,
                                                       const double* __restrict__ Q,
                                                       const double* A0, const double* B0,
                                                       const double* A1, const double* B1,
                                                       int nv, double* __restrict__ part,
                                                       int nchunk) {
  __shared__ double red[256 / kWave];
  const int i0 = blockIdx.y * kGemvRows, ch = blockIdx.x;
  const long long per = (n + nchunk - 1) / nchunk;
  const long long j0 = ch * per, j1 = j0 + per < n ? j0 + per : n;
  double s0[kGemvRows], s1[kGemvRows];
#pragma unroll
  for (int q = 0; q < kGemvRows; ++q) s0[q] = s1[q] = 0.0;
  for (long long j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    const double a0 = A0[j], b0 = B0[j];
    const double a1 = nv > 1 ? A1[j] : 0.0, b1 = nv > 1 ? B1[j] : 0.0;
#pragma unroll
    for (int q = 0; q < kGemvRows; ++q) {
      const int i = i0 + q;
      if (i < m) {
        const double p = P[(long long)i * n + j], qv = Q[(long long)i * n + j];
        s0[q] += p * a0 + qv * b0;
        if (nv > 1) s1[q] += p * a1 + qv * b1;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kGemvRows; ++q) {
    const int i = i0 + q;
    if (i >= m) break;  // uniform
    if (q > 0) __syncthreads();
    const double t0 = block_sum(s0[q], red);
    if (threadIdx.x == 0) part[(long long)i * nchunk + ch] = t0;
    if (nv > 1) {
      __syncthreads();
      const double t1 = block_sum(s1[q], red);
      if (threadIdx.x == 0) part[((long long)m + i) * nchunk + ch] = t1;
    }
  }
}

// out_v[i] = sum over the chunks of part[v][i][*] (one workgroup per (i, v))
__global__ __launch_bounds__(256) void mma_rows_reduce_kernel(int m, int nchunk,
                                                              const double* __restrict__ part,
                                                              double* __restrict__ out0,
                                                              double* __restrict__ out1) {
  __shared__ double red[256 / kWave];
  const int i = blockIdx.x, v = blockIdx.y;
  double s = 0.0;
  for (int ch = threadIdx.x; ch < nchunk; ch += blockDim.x)
    s += part[((long long)v * m + i) * nchunk + ch];
  s = block_sum(s, red);
  if (threadIdx.x == 0) (v == 0 ? out0 : out1)[i] = s;
}

// ------------------------------------------------------------------- Gram ----

// One pass over P and Q for the Newton system: per column chunk ch,
//   part[ch][i][k] (i >= k, i, k < m) = sum_j GG_ij GG_kj / diagx_j,
//   part[ch][m][k] = sum_j delx_j GG_kj / diagx_j          (GG delx/diagx, k < m),
//   part[ch][m + 1][k] = sum_j P_kj uxinv1_j + Q_kj xlinv1_j  (the row sums gvec),
// GG = P diag(uxinv2) - Q diag(xlinv2), with row stride m + 1.  The first two are one
// product on the fp64 matrix cores over m + 1 rows (the extra row is delx); the rows are
// padded to nt 16-row tiles and wave w owns the lower tile pairs (I, K) = w, w + 4, ...
// in the order (0,0), (1,0), (1,1), (2,0), ..., accumulating each over the chunk's
// columns (v_mfma_f64_16x16x4f64: lane l supplies A(l % 16, l / 16) and B(l / 16, l % 16)
// of a 16 x 4 by 4 x 16 step and holds D(l / 16 + 4 r, l % 16), r < 4).  The rows and
// the rows / diagx pass through LDS SW columns at a time, k-major ([column][row], rows
// padded to 16 nt + 1); thread t forms column t % SW for rows t / SW + (256 / SW) q and
// also sums its P and Q values into gvec, and the next stage's loads are in flight while
// the current one is multiplied.  PW: tile pairs per wave at most.
// R: rows per thread and stage, m + 1 <= 256 / SW * R (R = 8 covers every m of the
// tile count; R = 5 keeps m + 1 <= 20 at four waves per SIMD).
typedef double gram_d4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline int gram_tiles(int m) { return (m + 16) / 16; }  // m + 1 rows
__host__ __device__ inline int gram_stage_cols(int m) {
  const int nt = gram_tiles(m);
  return nt <= 2 ? 64 : (nt <= 4 ? 32 : 16);
}
__host__ __device__ inline int gram_lds_doubles(int m) {
  return 2 * gram_stage_cols(m) * (16 * gram_tiles(m) + 1);
}
__host__ __device__ inline long long gram_part_doubles(int m) {
  return (long long)(m + 1) * (m + 1) + m;
}

template <int SW, int PW, int R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(R == 5 ? 4 : 1)))
void mma_gram_kernel(MmaDev D) {
  extern __shared__ double gsm[];
  const int m = D.m, n = D.n, me = m + 1;
  const int nt = gram_tiles(m), ld = 16 * nt + 1;
  double* sA = gsm;            // [SW][ld]: GG and delx
  double* sB = gsm + SW * ld;  // [SW][ld]: the same / diagx
  const int ch = blockIdx.x;
  const long long per = (n + D.gchunk - 1) / D.gchunk;
  const long long j0 = ch * per, j1 = j0 + per < n ? j0 + per : n;
  constexpr int kStep = 256 / SW;
  const int jj = threadIdx.x % SW, r0 = threadIdx.x / SW;
  // rows m + 1 .. 16 nt - 1 stay zero in every stage
  for (int e = threadIdx.x; e < SW * ld; e += 256) {
    const int i = e % ld;
    if (i >= me) sA[e] = sB[e] = 0.0;
  }
  double nP[R], nQ[R], nux = 0.0, nxl = 0.0, ndg = 1.0, ndl = 0.0;
  double nu1 = 0.0, nx1 = 0.0;
  auto fetch = [&](long long jb) {
    const long long j = jb + jj;
    const bool in = j < j1;
    nux = in ? D.uxinv2[j] : 0.0;
    nxl = in ? D.xlinv2[j] : 0.0;
    ndg = in ? D.diagx[j] : 1.0;
    ndl = in ? D.delx[j] : 0.0;
    nu1 = in ? D.uxinv1[j] : 0.0;
    nx1 = in ? D.xlinv1[j] : 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = r0 + kStep * q;
      nP[q] = (in && i < m) ? D.P[(long long)i * n + j] : 0.0;
      nQ[q] = (in && i < m) ? D.Q[(long long)i * n + j] : 0.0;
    }
  };
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int kl = l >> 4, lr = l & 15;
  const int npair = nt * (nt + 1) / 2;
  int pI[PW], pK[PW];
#pragma unroll
  for (int q = 0; q < PW; ++q) {
    const int p = w + 4 * q;
    int I = 0;
    while ((I + 1) * (I + 2) / 2 <= p) ++I;
    pI[q] = I;
    pK[q] = p - I * (I + 1) / 2;
  }
  gram_d4 acc[PW];
#pragma unroll
  for (int q = 0; q < PW; ++q) acc[q] = gram_d4{0.0, 0.0, 0.0, 0.0};
  double gv[R];
#pragma unroll
  for (int q = 0; q < R; ++q) gv[q] = 0.0;
  fetch(j0);
  for (long long jb = j0; jb < j1; jb += SW) {
    __syncthreads();  // the previous stage is consumed (and the zero rows written)
    const double dinv = 1.0 / ndg;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = r0 + kStep * q;
      if (i < m) {
        const double g = nP[q] * nux - nQ[q] * nxl;
        sA[jj * ld + i] = g;
        sB[jj * ld + i] = g * dinv;
        gv[q] += nP[q] * nu1 + nQ[q] * nx1;
      } else if (i == m) {
        sA[jj * ld + i] = ndl;
        sB[jj * ld + i] = ndl * dinv;
      }
    }
    __syncthreads();
    if (jb + SW < j1) fetch(jb + SW);
#pragma unroll
    for (int q = 0; q < PW; ++q) {
      if (w + 4 * q >= npair) continue;  // uniform across the wavefront
      const double* a = sA + kl * ld + 16 * pI[q] + lr;
      const double* b = sB + kl * ld + 16 * pK[q] + lr;
      // four steps at a time: their LDS operands only are live (occupancy)
#pragma unroll 4
      for (int kk = 0; kk < SW; kk += 4)
        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk * ld], b[kk * ld], acc[q], 0, 0, 0);
    }
  }
  double* out = D.grampart + (long long)ch * gram_part_doubles(m);
#pragma unroll
  for (int q = 0; q < PW; ++q) {
    if (w + 4 * q >= npair) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * pI[q] + kl + 4 * r, k = 16 * pK[q] + lr;
      if (i < me && k <= i) out[i * me + k] = acc[q][r];
    }
  }
  // gvec: the SW column threads of each row group, summed in column order through LDS
  __syncthreads();  // the last stage's operands are consumed
  double* red = gsm;  // [16 nt][SW]
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = r0 + kStep * q;
    if (i < m) red[i * SW + jj] = gv[q];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < m; i += 256) {
    double t = 0.0;
    for (int c = 0; c < SW; ++c) t += red[i * SW + c];
    out[me * me + i] = t;
  }
}

// Sums over the chunks of the Gram pass, one workgroup per output: alam[i][k] (i >= k;
// every such entry is written by the Gram kernel's blocks), then ggd[k], then gvec[k].
__global__ __launch_bounds__(256) void mma_gram_reduce_kernel(MmaDev D) {
  __shared__ double red[256 / kWave];
  const int m = D.m, me = m + 1;
  const int e = blockIdx.x;
  long long src;
  double* dst;
  if (e < m * m) {
    const int i = e / m, k = e - i * m;
    if (i < k) return;  // uniform per block
    src = (long long)i * me + k;
    dst = D.alam + e;
  } else if (e < m * m + m) {
    const int k = e - m * m;
    src = (long long)m * me + k;
    dst = D.ggd + k;
  } else {
    const int k = e - m * m - m;
    src = (long long)me * me + k;
    dst = D.gvec + k;
  }
  const long long stride = gram_part_doubles(m);
  double s = 0.0;
  for (int ch = threadIdx.x; ch < D.gchunk; ch += blockDim.x) s += D.grampart[ch * stride + src];
  s = block_sum(s, red);
  if (threadIdx.x == 0) *dst = s;
}

// ----------------------------------------------------------- small (1 WG) ----

enum SmallMode { kSmInit = 0, kSmResid = 1, kSmSolve = 2, kSmSteg = 3, kSmApply = 4 };

// In-place Gaussian elimination with partial pivoting of the N x N system in LDS
// (row stride ld), right-hand side rhs; the solution replaces rhs.
__device__ void lu_solve(double* A, int ld, double* rhs, int N, double* red, int* piv) {
  for (int k = 0; k < N; ++k) {
    if (threadIdx.x == 0) {
      int p = k;
      double best = fabs(A[k * ld + k]);
      for (int i = k + 1; i < N; ++i) {
        const double v = fabs(A[i * ld + k]);
        if (v > best) {
          best = v;
          p = i;
        }
      }
      *piv = p;
    }
    __syncthreads();
    const int p = *piv;
    if (p != k) {
      for (int c = threadIdx.x; c < N; c += blockDim.x) {
        const double tmp = A[k * ld + c];
        A[k * ld + c] = A[p * ld + c];
        A[p * ld + c] = tmp;
      }
      if (threadIdx.x == 0) {
        const double tmp = rhs[k];
        rhs[k] = rhs[p];
        rhs[p] = tmp;
      }
    }
    __syncthreads();
    const double pv = A[k * ld + k];
    // rows below k: each thread a set of (row, column) entries
    const int rows = N - k - 1, cols = N - k;
    for (int e = threadIdx.x; e < rows * cols; e += blockDim.x) {
      const int i = k + 1 + e / cols, c = k + e % cols;
      const double f = A[i * ld + k] / pv;
      if (c > k) A[i * ld + c] -= f * A[k * ld + c];
    }
    __syncthreads();
    for (int i = k + 1 + threadIdx.x; i < N; i += blockDim.x) {
      const double f = A[i * ld + k] / pv;
      rhs[i] -= f * rhs[k];
      A[i * ld + k] = 0.0;
    }
    __syncthreads();
  }
  // back substitution
  for (int k = N - 1; k >= 0; --k) {
    if (threadIdx.x == 0) rhs[k] /= A[k * ld + k];
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += blockDim.x) rhs[i] -= A[i * ld + k] * rhs[k];
    __syncthreads();
  }
  (void)red;
}

constexpr int kLd = kMmaMaxM + 2;  // 128

template <int MODE>
__global__ __launch_bounds__(256) void mma_small_kernel(MmaDev D, double epsi, int jj,
                                                        int branch_mlt) {
  __shared__ double sAA[kLd * kLd];
  __shared__ double sb[kLd];
  __shared__ double red[256 / kWave];
  __shared__ int piv;
  const int m = D.m;
  double* zz = D.zz;
  if constexpr (MODE == kSmInit) {
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      D.y[i] = 1.0;
      D.lam[i] = 1.0;
      D.mu[i] = fmax(1.0, 0.5 * D.c[i]);
      D.s[i] = 1.0;
      D.dy[i] = D.dlam[i] = D.dmu[i] = D.ds[i] = 0.0;
    }
    if (threadIdx.x == 0) {
      zz[kZ] = 1.0;
      zz[kZet] = 1.0;
      zz[kDz] = zz[kDzet] = 0.0;
      zz[kSteg] = 0.0;
    }
  } else if constexpr (MODE == kSmResid) {
    // candidate t = steg / 2^jj (jj < 0: the current point)
    const double t = jj < 0 ? 0.0 : zz[kSteg] * ldexp(1.0, -jj);
    double sumsq = 0.0, mabs = 0.0;
    bool nan_seen = false;
    for (int b = threadIdx.x; b < D.colblk; b += blockDim.x) {
      sumsq += D.colpart[b * kColSlots + kSumSq];
      const double v = D.colpart[b * kColSlots + kMaxAbs];
      nan_seen |= isnan(v);
      mabs = fmax(mabs, v);
    }
    const double z = zz[kZ] + t * zz[kDz], zet = zz[kZet] + t * zz[kDzet];
    double alam = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      const double y = D.y[i] + t * D.dy[i], lam = D.lam[i] + t * D.dlam[i];
      const double mu = D.mu[i] + t * D.dmu[i], s = D.s[i] + t * D.ds[i];
      const double g = D.gvec[i];
      const double rey = D.c[i] + D.d[i] * y - mu - lam;
      const double relam = g - D.a[i] * z - y + s - D.b[i];
      const double remu = mu * y - epsi;
      const double res = lam * s - epsi;
      sumsq += rey * rey + relam * relam + remu * remu + res * res;
      nan_seen |= isnan(rey) || isnan(relam) || isnan(remu) || isnan(res);
      mabs = fmax(mabs, fmax(fmax(fabs(rey), fabs(relam)), fmax(fabs(remu), fabs(res))));
      alam += lam * D.a[i];
    }
    sumsq = block_sum(sumsq, red);
    __syncthreads();
    const double asum = block_sum(alam, red);
    __syncthreads();
    mabs = block_max(mabs, red);
    __shared__ int anynan;
    if (threadIdx.x == 0) anynan = 0;
    __syncthreads();
    if (nan_seen) anynan = 1;
    __syncthreads();
    if (threadIdx.x == 0) {
      const double rez = D.a0 - zet - asum;
      const double rezet = zet * z - epsi;
      sumsq += rez * rez + rezet * rezet;
      mabs = fmax(mabs, fmax(fabs(rez), fabs(rezet)));
      const bool bad = anynan || isnan(rez) || isnan(rezet);
      D.out[0] = sqrt(sumsq);
      D.out[1] = bad ? NAN : mabs;
    }
  } else if constexpr (MODE == kSmSolve) {
    // Newton system at the current point; gvec from the row GEMV (v = 0), GG delx/diagx
    // (v = 1, m < n branch)
    const double z = zz[kZ], zet = zz[kZet];
    __shared__ double sdely[kLd], sdiagy[kLd], sdiaglamyi[kLd], sdellam[kLd];
    __shared__ double sdelz;
    double alam = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      const double y = D.y[i], lam = D.lam[i];
      const double g = D.gvec[i];
      const double dely = D.c[i] + D.d[i] * y - lam - epsi / y;
      const double dellam = g - D.a[i] * z - y - D.b[i] + epsi / lam;
      const double diagy = D.d[i] + D.mu[i] / y;
      sdely[i] = dely;
      sdiagy[i] = diagy;
      sdiaglamyi[i] = D.s[i] / lam + 1.0 / diagy;
      sdellam[i] = dellam;
      alam += D.a[i] * lam;
    }
    alam = block_sum(alam, red);
    if (threadIdx.x == 0) sdelz = D.a0 - alam - epsi / z;
    __syncthreads();
    const double delz = sdelz;
    int N;
    if (branch_mlt) {
      N = m + 1;
      for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
        const int i = e / m, k = e - i * m;
        const double v = i >= k ? D.alam[i * m + k] : D.alam[k * m + i];
        sAA[i * kLd + k] = v + (i == k ? sdiaglamyi[i] : 0.0);
      }
      for (int i = threadIdx.x; i < m; i += blockDim.x) {
        sAA[i * kLd + m] = D.a[i];
        sAA[m * kLd + i] = D.a[i];
        sb[i] = sdellam[i] + sdely[i] / sdiagy[i] - D.ggd[i];
      }
      if (threadIdx.x == 0) {
        sAA[m * kLd + m] = -zet / z;
        sb[m] = delz;
      }
      __syncthreads();
      lu_solve(sAA, kLd, sb, N, red, &piv);
      for (int i = threadIdx.x; i < m; i += blockDim.x) D.dlam[i] = sb[i];
      if (threadIdx.x == 0) zz[kDz] = sb[m];
      __syncthreads();
    } else {
      // m >= n: the (n + 1) system in x and z; GG = P diag(uxinv2) - Q diag(xlinv2)
      const int n = D.n;
      N = n + 1;
      for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int r = e / n, c = e - r * n;
        double v = r == c ? D.diagx[r] : 0.0;
        for (int i = 0; i < m; ++i) {
          const double gr = D.P[(long long)i * n + r] * D.uxinv2[r] - D.Q[(long long)i * n + r] * D.xlinv2[r];
          const double gc = D.P[(long long)i * n + c] * D.uxinv2[c] - D.Q[(long long)i * n + c] * D.xlinv2[c];
          const double dli = 1.0 / (D.s[i] / D.lam[i] + 1.0 / (D.d[i] + D.mu[i] / D.y[i]));
          v += gr * dli * gc;
        }
        sAA[r * kLd + c] = v;
      }
      for (int r = threadIdx.x; r <= n; r += blockDim.x) {
        double axz = 0.0, bx = 0.0;
        for (int i = 0; i < m; ++i) {
          const double y = D.y[i], lam = D.lam[i];
          const double diagy = D.d[i] + D.mu[i] / y;
          const double diaglamyi = D.s[i] / lam + 1.0 / diagy;
          const double dely = D.c[i] + D.d[i] * y - lam - epsi / y;
          const double dellam = D.gvec[i] - D.a[i] * z - y - D.b[i] + epsi / lam;
          const double dellamyi = dellam + dely / diagy;
          if (r < n) {
            const double g = D.P[(long long)i * n + r] * D.uxinv2[r] - D.Q[(long long)i * n + r] * D.xlinv2[r];
            axz += -g * (D.a[i] / diaglamyi);
            bx += g * (dellamyi / diaglamyi);
          } else {
            axz += D.a[i] * (D.a[i] / diaglamyi);
            bx += D.a[i] * (dellamyi / diaglamyi);
          }
        }
        if (r < n) {
          sAA[r * kLd + n] = axz;
          sAA[n * kLd + r] = axz;
          sb[r] = -(D.delx[r] + bx);
        } else {
          sAA[n * kLd + n] = zet / z + axz;
          sb[n] = -(delz - bx);
        }
      }
      __syncthreads();
      lu_solve(sAA, kLd, sb, N, red, &piv);
      const double dz = sb[n];
      for (int j = threadIdx.x; j < n; j += blockDim.x) D.dx[j] = sb[j];
      for (int i = threadIdx.x; i < m; i += blockDim.x) {
        const double y = D.y[i], lam = D.lam[i];
        const double diagy = D.d[i] + D.mu[i] / y;
        const double diaglamyi = D.s[i] / lam + 1.0 / diagy;
        const double dely = D.c[i] + D.d[i] * y - lam - epsi / y;
        const double dellam = D.gvec[i] - D.a[i] * z - y - D.b[i] + epsi / lam;
        const double dellamyi = dellam + dely / diagy;
        double ggdx = 0.0;
        for (int j = 0; j < n; ++j) {
          const double g = D.P[(long long)i * n + j] * D.uxinv2[j] - D.Q[(long long)i * n + j] * D.xlinv2[j];
          ggdx += g * sb[j];
        }
        D.dlam[i] = ggdx / diaglamyi - dz * (D.a[i] / diaglamyi) + dellamyi / diaglamyi;
      }
      if (threadIdx.x == 0) zz[kDz] = dz;
      __syncthreads();
    }
    // remaining m-sized directions and the m part of the step bound
    const double dz = zz[kDz];
    double mxx = -INFINITY;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      const double y = D.y[i], lam = D.lam[i], mu = D.mu[i], s = D.s[i];
      const double dlam = D.dlam[i];
      const double dy = -sdely[i] / sdiagy[i] + dlam / sdiagy[i];
      const double dmu = -mu + epsi / y - (mu * dy) / y;
      const double ds = -s + epsi / lam - (s * dlam) / lam;
      D.dy[i] = dy;
      D.dmu[i] = dmu;
      D.ds[i] = ds;
      mxx = fmax(mxx, fmax(fmax(-1.01 * dy / y, -1.01 * dlam / lam),
                           fmax(-1.01 * dmu / mu, -1.01 * ds / s)));
    }
    mxx = block_max(mxx, red);
    if (threadIdx.x == 0) {
      const double dzet = -zet + epsi / z - zet * dz / z;
      zz[kDzet] = dzet;
      mxx = fmax(mxx, fmax(-1.01 * dz / z, -1.01 * dzet / zet));
      zz[kSteg] = mxx;  // the m part; kSmSteg folds in the column maxima
    }
  } else if constexpr (MODE == kSmSteg) {
    double ma = -INFINITY, mb = -INFINITY, mx = -INFINITY;
    for (int b = threadIdx.x; b < D.colblk; b += blockDim.x) {
      ma = fmax(ma, D.colpart[b * kColSlots + kMaxAlfa]);
      mb = fmax(mb, D.colpart[b * kColSlots + kMaxBeta]);
      mx = fmax(mx, D.colpart[b * kColSlots + kMaxXX]);
    }
    ma = block_max(ma, red);
    mb = block_max(mb, red);
    mx = block_max(mx, red);
    if (threadIdx.x == 0) {
      const double stmxx = fmax(mx, zz[kSteg]);
      zz[kSteg] = 1.0 / fmax(fmax(ma, mb), fmax(stmxx, 1.0));
    }
  } else if constexpr (MODE == kSmApply) {
    const double t = zz[kSteg] * ldexp(1.0, -jj);
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      D.y[i] += t * D.dy[i];
      D.lam[i] += t * D.dlam[i];
      D.mu[i] += t * D.dmu[i];
      D.s[i] += t * D.ds[i];
    }
    if (threadIdx.x == 0) {
      zz[kZ] += t * zz[kDz];
      zz[kZet] += t * zz[kDzet];
    }
  }
}

// ----------------------------------------------------------- mmasub prep ----

struct PrepArgs {
  int m, n, iteration;
  const double *xval, *xmin, *xmax, *xold1, *xold2, *df0dx, *dfdx;
  double *low, *upp;  // in / out
  const double *asyinit, *asymin, *asymax, *move;  // n-vectors or scalars (stride 0)
  long long s_asyinit, s_asymin, s_asymax, s_move;
  double raa0, albefa, asyincr, asydecr;
  double *alfa, *beta, *p0, *q0, *P, *Q, *uxinv, *xlinv;
};

__global__ __launch_bounds__(256) void mma_prep_kernel(PrepArgs A) {
  const int m = A.m, n = A.n;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (long long)gridDim.x * blockDim.x) {
    const double xv = A.xval[j], xmn = A.xmin[j], xmx = A.xmax[j];
    const double xr = xmx - xmn;
    const double ai = A.asyinit[j * A.s_asyinit];
    double lo, up;
    if (A.iteration < 2) {
      lo = xv - ai * xr;
      up = xv + ai * xr;
    } else {
      const double x1 = A.xold1[j], x2 = A.xold2[j];
      const double zzz = (xv - x1) * (x1 - x2);
      const double f = zzz > 0 ? A.asyincr : (zzz < 0 ? A.asydecr : 1.0);
      lo = xv - f * (x1 - A.low[j]);
      up = xv + f * (A.upp[j] - x1);
      const double amax = A.asymax[j * A.s_asymax], amin = A.asymin[j * A.s_asymin];
      const double lowmin = xv - amax * xr, lowmax = xv - amin * xr;
      const double uppmin = xv + amin * xr, uppmax = xv + amax * xr;
      lo = fmin(fmax(lo, lowmin), lowmax);
      up = fmax(fmin(up, uppmax), uppmin);
    }
    A.low[j] = lo;
    A.upp[j] = up;
    const double mv = A.move[j * A.s_move];
    const double al = fmax(fmax(lo + A.albefa * (xv - lo), xv - mv * xr), xmn);
    const double be = fmin(fmin(up - A.albefa * (up - xv), xv + mv * xr), xmx);
    A.alfa[j] = al;
    A.beta[j] = be;
    const double xmami = fmax(xr, 1e-5);
    const double xmamiinv = 1.0 / xmami;
    const double ux1 = up - xv, xl1 = xv - lo;
    const double ux2 = ux1 * ux1, xl2 = xl1 * xl1;
    A.uxinv[j] = 1.0 / ux1;
    A.xlinv[j] = 1.0 / xl1;
    const double g0 = A.df0dx[j];
    double p0 = fmax(g0, 0.0), q0 = fmax(-g0, 0.0);
    const double pq0 = 0.001 * (p0 + q0) + A.raa0 * xmamiinv;
    A.p0[j] = (p0 + pq0) * ux2;
    A.q0[j] = (q0 + pq0) * xl2;
    for (int i = 0; i < m; ++i) {
      const double g = A.dfdx[(long long)i * n + j];
      double p = fmax(g, 0.0), q = fmax(-g, 0.0);
      const double pq = 0.001 * (p + q) + A.raa0 * xmamiinv;
      A.P[(long long)i * n + j] = (p + pq) * ux2;
      A.Q[(long long)i * n + j] = (q + pq) * xl2;
    }
  }
}

__global__ __launch_bounds__(256) void mma_b_kernel(int m, int nchunk, const double* part,
                                                    const double* fval, double* b) {
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    double s = 0.0;
    for (int ch = 0; ch < nchunk; ++ch) s += part[(long long)i * nchunk + ch];
    b[i] = s - fval[i];
  }
}

int col_blocks(int n) {
  const int b = (n + kCol - 1) / kCol;
  return b < kColBlocks ? b : kColBlocks;
}

}  // namespace

}  // namespace stt

using namespace stt;

// The state and scratch layout of the subsolv workspace.
static void carve_mma(MmaDev& D, double* work) {
  const int m = D.m, n = D.n;
  D.nchunk = n < 4096 ? 1 : (n / 4096 < 256 ? n / 4096 : 256);
  D.gchunk = (n + 1023) / 1024 < 1024 ? (n + 1023) / 1024 : 1024;
  D.colblk = col_blocks(n);
  double* w = work;
  auto take = [&](long long k) {
    double* p = w;
    w += k;
    return p;
  };
  D.x = take(n);
  D.xsi = take(n);
  D.eta = take(n);
  D.y = take(m);
  D.lam = take(m);
  D.mu = take(m);
  D.s = take(m);
  D.zz = take(8);
  D.dx = take(n);
  D.dxsi = take(n);
  D.deta = take(n);
  D.dy = take(m);
  D.dlam = take(m);
  D.dmu = take(m);
  D.ds = take(m);
  D.delx = take(n);
  D.diagx = take(n);
  D.uxinv1 = take(n);
  D.xlinv1 = take(n);
  D.uxinv2 = take(n);
  D.xlinv2 = take(n);
  D.PTl = take(n);
  D.QTl = take(n);
  D.PTd = take(n);
  D.QTd = take(n);
  D.wA = take(n);
  D.wB = take(n);
  D.uxc = take(n);
  D.xlc = take(n);
  D.colpart = take((long long)kColBlocks * kColSlots);
  D.rowpart = take(2LL * m * D.nchunk);
  D.grampart = take((long long)D.gchunk * gram_part_doubles(m));
  D.alam = take((long long)m * m);
  D.gvec = take(m);
  D.ggd = take(m);
  D.out = take(2);
}

extern "C" {

int64_t sttopt_mma_work_count(int32_t m, int32_t n) {
  if (m <= 0 || n <= 0) return 0;
  MmaDev D{};
  D.m = m;
  D.n = n;
  double* base = reinterpret_cast<double*>(4096);
  carve_mma(D, base);
  return D.out + 2 - base;
}

int sttopt_mma_state_offsets(int32_t m, int32_t n, int64_t* offsets) {
  if (m <= 0 || n <= 0 || offsets == nullptr) {
    set_error("sttopt_mma_state_offsets: invalid arguments");
    return STTOPT_EINVAL;
  }
  MmaDev D{};
  D.m = m;
  D.n = n;
  double* base = reinterpret_cast<double*>(4096);
  carve_mma(D, base);
  const double* ptrs[10] = {D.x, D.xsi, D.eta, D.y, D.lam, D.mu, D.s, D.zz + kZ, D.zz + kZet,
                            D.out};
  for (int i = 0; i < 10; ++i) offsets[i] = ptrs[i] - base;
  return STTOPT_OK;
}

int sttopt_mma_subsolv(int32_t m, int32_t n, const double* low, const double* upp,
                       const double* alfa, const double* beta, const double* p0,
                       const double* q0, const double* P, const double* Q, double a0,
                       const double* a, const double* b, const double* c, const double* d,
                       double* work, int32_t op, double epsi, int32_t jj, void* stream) {
  if (m <= 0 || n <= 0 || work == nullptr || m > kMmaMaxM) {
    set_error("sttopt_mma_subsolv: invalid sizes (1 <= m <= 126 constraints, n >= 1)");
    return STTOPT_EINVAL;
  }
  MmaDev D{};
  D.m = m;
  D.n = n;
  D.low = low;
  D.upp = upp;
  D.alfa = alfa;
  D.beta = beta;
  D.p0 = p0;
  D.q0 = q0;
  D.P = P;
  D.Q = Q;
  D.a = a;
  D.b = b;
  D.c = c;
  D.d = d;
  D.a0 = a0;
  carve_mma(D, work);
  hipStream_t s = (hipStream_t)stream;
  const int cb = col_blocks(n);
  const int mlt = m < n;
  const dim3 gv(D.nchunk, (m + kGemvRows - 1) / kGemvRows);
  switch (op) {
    case 0:  // initial point
      hipLaunchKernelGGL(mma_col_kernel<kColInit>, dim3(cb), dim3(kCol), 0, s, D, epsi, 0);
      hipLaunchKernelGGL(mma_small_kernel<kSmInit>, dim3(1), dim3(256), 0, s, D, epsi, 0, mlt);
      break;
    case 1:  // residual at the current point (jj < 0) or candidate steg / 2^jj
      hipLaunchKernelGGL(mma_col_kernel<kColResid>, dim3(cb), dim3(kCol), 0, s, D, epsi,
                         jj < 0 ? -1 : jj);
      hipLaunchKernelGGL(mma_gemv_kernel, gv, dim3(256), 0, s, m, n, P, Q, D.uxc, D.xlc,
                         nullptr, nullptr, 1, D.rowpart, D.nchunk);
      hipLaunchKernelGGL(mma_rows_reduce_kernel, dim3(m, 1), dim3(256), 0, s, m, D.nchunk,
                         D.rowpart, D.gvec, D.ggd);
      hipLaunchKernelGGL(mma_small_kernel<kSmResid>, dim3(1), dim3(256), 0, s, D, epsi,
                         jj < 0 ? -1 : jj, mlt);
      break;
    case 2:  // Newton direction and step bound
      hipLaunchKernelGGL(mma_col_kernel<kColPass1>, dim3(cb), dim3(kCol), 0, s, D, epsi, 0);
      if (mlt) {
        // the Gram pass also forms gvec and GG delx/diagx; at most 2 x 64 x 33 doubles
        // (34 KB) of LDS for every m <= 126
        const size_t lds = sizeof(double) * gram_lds_doubles(m);
        const int nt = gram_tiles(m);
        if (m + 1 <= 20)
          hipLaunchKernelGGL((mma_gram_kernel<64, 1, 5>), dim3(D.gchunk), dim3(256), lds, s, D);
        else if (nt <= 2)
          hipLaunchKernelGGL((mma_gram_kernel<64, 1, 8>), dim3(D.gchunk), dim3(256), lds, s, D);
        else if (nt <= 4)
          hipLaunchKernelGGL((mma_gram_kernel<32, 3, 8>), dim3(D.gchunk), dim3(256), lds, s, D);
        else
          hipLaunchKernelGGL((mma_gram_kernel<16, 9, 8>), dim3(D.gchunk), dim3(256), lds, s, D);
        hipLaunchKernelGGL(mma_gram_reduce_kernel, dim3(m * m + 2 * m), dim3(256), 0, s, D);
      } else {
        hipLaunchKernelGGL(mma_gemv_kernel, gv, dim3(256), 0, s, m, n, P, Q, D.uxinv1,
                           D.xlinv1, D.wA, D.wB, 2, D.rowpart, D.nchunk);
        hipLaunchKernelGGL(mma_rows_reduce_kernel, dim3(m, 2), dim3(256), 0, s, m, D.nchunk,
                           D.rowpart, D.gvec, D.ggd);
      }
      hipLaunchKernelGGL(mma_small_kernel<kSmSolve>, dim3(1), dim3(256), 0, s, D, epsi, 0, mlt);
      if (mlt)
        hipLaunchKernelGGL(mma_col_kernel<kColStep>, dim3(cb), dim3(kCol), 0, s, D, epsi, 0);
      else
        hipLaunchKernelGGL(mma_col_kernel<kColStepGivenDx>, dim3(cb), dim3(kCol), 0, s, D, epsi, 0);
      hipLaunchKernelGGL(mma_small_kernel<kSmSteg>, dim3(1), dim3(256), 0, s, D, epsi, 0, mlt);
      break;
    case 3:  // accept candidate jj
      hipLaunchKernelGGL(mma_col_kernel<kColApply>, dim3(cb), dim3(kCol), 0, s, D, epsi, jj);
      hipLaunchKernelGGL(mma_small_kernel<kSmApply>, dim3(1), dim3(256), 0, s, D, epsi, jj, mlt);
      break;
    default:
      set_error("sttopt_mma_subsolv: unknown op");
      return STTOPT_EINVAL;
  }
  STT_CHECK_LAUNCH("mma kernels");
  return STTOPT_OK;
}

int sttopt_mma_prepare(int32_t m, int32_t n, int32_t iteration, const double* xval,
                       const double* xmin, const double* xmax, const double* xold1,
                       const double* xold2, const double* df0dx, const double* fval,
                       const double* dfdx, double* low, double* upp, const double* asyinit,
                       int64_t asyinit_stride, const double* asymin, int64_t asymin_stride,
                       const double* asymax, int64_t asymax_stride, const double* move,
                       int64_t move_stride, double raa0, double albefa, double asyincr,
                       double asydecr, double* alfa, double* beta, double* p0, double* q0,
                       double* P, double* Q, double* b, double* work, void* stream) {
  if (m <= 0 || n <= 0 || work == nullptr) {
    set_error("sttopt_mma_prepare: invalid sizes");
    return STTOPT_EINVAL;
  }
  PrepArgs A{m, n, iteration, xval, xmin, xmax, xold1, xold2, df0dx, dfdx, low, upp,
             asyinit, asymin, asymax, move, asyinit_stride, asymin_stride, asymax_stride,
             move_stride, raa0, albefa, asyincr, asydecr, alfa, beta, p0, q0, P, Q,
             work, work + n};
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(mma_prep_kernel, dim3(col_blocks(n)), dim3(256), 0, s, A);
  const int nchunk = n < 4096 ? 1 : (n / 4096 < 256 ? n / 4096 : 256);
  double* part = work + 2LL * n;
  hipLaunchKernelGGL(mma_gemv_kernel, dim3(nchunk, (m + kGemvRows - 1) / kGemvRows), dim3(256),
                     0, s, m, n, P, Q, work,
                     work + n, nullptr, nullptr, 1, part, nchunk);
  hipLaunchKernelGGL(mma_b_kernel, dim3(1), dim3(256), 0, s, m, nchunk, part, fval, b);
  STT_CHECK_LAUNCH("mma_prepare kernels");
  return STTOPT_OK;
}

int64_t sttopt_mma_prepare_work_count(int32_t m, int32_t n) {
  if (m <= 0 || n <= 0) return 0;
  const long long nchunk = n < 4096 ? 1 : (n / 4096 < 256 ? n / 4096 : 256);
  return 2LL * n + (long long)m * nchunk;
}

}  // extern "C"
