// Gauss-point gradient fields of the scalar (one dof per node) Q4 mesh behind the
// virtual-heat time fields, and their adjoints (sttopt/virtual_heat.py:290-347):
//
//   gauss gradients  G[e][g][d] = sum_i B[g][d][i] u[node_i(e)]
//   direction field  Y = G / sqrt(|G|^2 + DBL_MIN) per Gauss point
//   divergence load  F[n] = 1/4 sum_{e at n} w_e sum_{g,d} X[e][g][d] B[g][d][loc_n(e)]
//
// B holds the bilinear shape-function gradients at the 2 x 2 Gauss points in the local
// node order BL, BR, TR, TL.  The gradients run one element per thread with their
// stores transposed through LDS; the adjoint kernels one thread per (element, Gauss
// point), so a wavefront's 16-byte accesses are contiguous; the load is a node-centric
// gather over the (up to) four elements around a node, in a fixed order, from per-corner
// element contributions staged in LDS, so it needs no atomics and is deterministic.
#include <cfloat>

#include "stt_kernels.h"

namespace stt {

namespace {

struct Bmat {
  double v[4][2][4];  // [gauss point][x, y][local node]
};

// The four nodal values of element e = (r, c): BL, BR, TR, TL.
__device__ __forceinline__ void element_nodes(const double* __restrict__ u, int nnx, int r,
                                              int c, double (&ue)[4]) {
  const long long tl = (long long)r * nnx + c;
  ue[0] = u[tl + nnx];
  ue[1] = u[tl + nnx + 1];
  ue[2] = u[tl + 1];
  ue[3] = u[tl];
}

// One thread per (element, Gauss point): thread q handles Gauss point q % 4 of element
// q / 4, so the 16-byte results of a wavefront are contiguous in the [nel][4][2] field.
__device__ __forceinline__ bool gauss_thread(int nelx, int nely, long long& e, int& g, int& r,
                                             int& c) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  e = q >> 2;
  g = (int)(q & 3);
  if (e >= (long long)nelx * nely) return false;
  r = (int)(e / nelx);
  c = (int)(e - (long long)r * nelx);
  return true;
}

__device__ __forceinline__ double2 grad_at(const Bmat& B, const double (&ue)[4], int g) {
  double2 v;
  double sx = 0.0, sy = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) sx += B.v[g][0][i] * ue[i];
#pragma unroll
  for (int i = 0; i < 4; ++i) sy += B.v[g][1][i] * ue[i];
  v.x = sx;
  v.y = sy;
  return v;
}

// out = scale * escale_e * G (escale may be null), or the direction field when NORM.
// One element per thread; the block's 256 x 8 results are transposed through LDS
// (element stride 9 doubles against bank conflicts) and stored as 8 fully coalesced
// passes over the block's contiguous output.
template <bool NORM>
__global__ __launch_bounds__(256) void gauss_gradients_kernel(
    int nelx, int nely, Bmat B, const double* __restrict__ u, long long su,
    const double* __restrict__ escale, long long ses, double scale, double* __restrict__ out,
    long long so) {
  __shared__ double sO[256 * 9];
  const int row = blockIdx.y;
  const long long nel = (long long)nelx * nely;
  const long long e0 = (long long)blockIdx.x * blockDim.x;
  const long long e = e0 + threadIdx.x;
  if (e < nel) {
    const int r = (int)(e / nelx), c = (int)(e - (long long)r * nelx);
    double ue[4];
    element_nodes(u + row * su, nelx + 1, r, c, ue);
    double f = scale;
    if (!NORM && escale != nullptr) f *= escale[row * ses + e];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const double2 gv = grad_at(B, ue, g);
      double vx, vy;
      if (NORM) {
        const double inv = 1.0 / sqrt(gv.x * gv.x + gv.y * gv.y + DBL_MIN);
        vx = gv.x * inv;
        vy = gv.y * inv;
      } else {
        vx = f * gv.x;
        vy = f * gv.y;
      }
      sO[threadIdx.x * 9 + 2 * g] = vx;
      sO[threadIdx.x * 9 + 2 * g + 1] = vy;
    }
  }
  __syncthreads();
  const long long n8 = (nel - e0 < 256 ? nel - e0 : 256) * 8;  // doubles of this block
  double* o = out + row * so + 8 * e0;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int j = threadIdx.x + 256 * m;
    if (j < n8) o[j] = sO[(j >> 3) * 9 + (j & 7)];
  }
}

// F[n] = scale * sum over the elements around node n of w_e sum_{g,d} X[e][g][d]
// B[g][d][loc]; w may be null (1).  A workgroup owns 64 x 4 nodes: it first forms the
// four per-corner contributions of the 65 x 5 elements around them in LDS (each element's
// X read once), then every node adds those of its up-left, up-right, down-left and
// down-right element (where it is the BR, BL, TR, TL corner), in that order.
constexpr int kLX = 64, kLY = 4;
__global__ __launch_bounds__(256) void divergence_load_kernel(
    int nelx, int nely, Bmat B, const double* __restrict__ w, long long sw,
    const double* __restrict__ X, long long sx, double scale, double* __restrict__ F,
    long long sf, int ntx, int tiles) {
  __shared__ double sC[kLY + 1][kLX + 1][4];
  const int row = blockIdx.x / tiles;
  const int t = blockIdx.x - row * tiles;
  const int by = t / ntx, bx = t - by * ntx;
  const int r0 = by * kLY, c0 = bx * kLX;
  const double* Xr = X + row * sx;
  constexpr int kE = (kLY + 1) * (kLX + 1);
  for (int i = threadIdx.x; i < kE; i += blockDim.x) {
    const int ly = i / (kLX + 1), lx = i - ly * (kLX + 1);
    const int er = r0 - 1 + ly, ec = c0 - 1 + lx;
    double cc[4] = {0.0, 0.0, 0.0, 0.0};
    if (er >= 0 && er < nely && ec >= 0 && ec < nelx) {
      const long long e = (long long)er * nelx + ec;
      const double2* xe = reinterpret_cast<const double2*>(Xr + 8 * e);
      double2 xv[4];
#pragma unroll
      for (int g = 0; g < 4; ++g) xv[g] = xe[g];
      const double we = w != nullptr ? w[row * sw + e] : 1.0;
#pragma unroll
      for (int loc = 0; loc < 4; ++loc) {
        double s = 0.0;
#pragma unroll
        for (int g = 0; g < 4; ++g) s += xv[g].x * B.v[g][0][loc] + xv[g].y * B.v[g][1][loc];
        cc[loc] = w != nullptr ? s * we : s;
      }
    }
#pragma unroll
    for (int loc = 0; loc < 4; ++loc) sC[ly][lx][loc] = cc[loc];
  }
  __syncthreads();
  const int lane = threadIdx.x & (kLX - 1), wv = threadIdx.x / kLX;
  const int r = r0 + wv, c = c0 + lane;
  if (r > nely || c > nelx) return;
  double acc = 0.0;
  acc += sC[wv][lane][1];          // up-left element: the node is its BR corner
  acc += sC[wv][lane + 1][0];      // up-right: BL
  acc += sC[wv + 1][lane][2];      // down-left: TR
  acc += sC[wv + 1][lane + 1][3];  // down-right: TL
  F[row * sf + (long long)r * (nelx + 1) + c] = scale * acc;
}

// out_e = scale * sum_{g,d} X[e][g][d] G(v)[e][g][d]: the weight gradient of the load.
// One thread per Gauss point; the four of an element sum through lane shuffles.
__global__ __launch_bounds__(256) void gauss_dot_kernel(int nelx, int nely, Bmat B,
                                                        const double* __restrict__ X,
                                                        long long sx,
                                                        const double* __restrict__ v,
                                                        long long sv, double scale,
                                                        double* __restrict__ out,
                                                        long long so) {
  const int row = blockIdx.y;
  long long e;
  int g, r, c;
  const bool ok = gauss_thread(nelx, nely, e, g, r, c);
  double s = 0.0;
  if (ok) {
    double ve[4];
    element_nodes(v + row * sv, nelx + 1, r, c, ve);
    const double2 gv = grad_at(B, ve, g);
    const double2 x = reinterpret_cast<const double2*>(X + row * sx)[4 * e + g];
    s = x.x * gv.x + x.y * gv.y;
  }
  // the 4 Gauss points of an element are 4 consecutive lanes
  s += __shfl_xor(s, 1, kWave);
  s += __shfl_xor(s, 2, kWave);
  if (ok && g == 0) out[row * so + e] = scale * s;
}

// Adjoint of the direction field at the Gauss points: with G the gradients of s,
// n = sqrt(|G|^2 + DBL_MIN) and Y = G / n, gG = (gY - Y (Y . gY)) / n.
__global__ __launch_bounds__(256) void direction_backward_kernel(
    int nelx, int nely, Bmat B, const double* __restrict__ s, long long ss,
    const double* __restrict__ gY, long long sgy, double* __restrict__ gG, long long sgg) {
  const int row = blockIdx.y;
  long long e;
  int g, r, c;
  if (!gauss_thread(nelx, nely, e, g, r, c)) return;
  double se[4];
  element_nodes(s + row * ss, nelx + 1, r, c, se);
  const double2 gv = grad_at(B, se, g);
  const double inv = 1.0 / sqrt(gv.x * gv.x + gv.y * gv.y + DBL_MIN);
  const double yx = gv.x * inv, yy = gv.y * inv;
  const double2 b = reinterpret_cast<const double2*>(gY + row * sgy)[4 * e + g];
  const double dot = yx * b.x + yy * b.y;
  double2 o;
  o.x = (b.x - yx * dot) * inv;
  o.y = (b.y - yy * dot) * inv;
  reinterpret_cast<double2*>(gG + row * sgg)[4 * e + g] = o;
}

Bmat bmat(const double* B) {
  Bmat b;
  for (int q = 0; q < 4; ++q)
    for (int d = 0; d < 2; ++d)
      for (int i = 0; i < 4; ++i) b.v[q][d][i] = B[(q * 2 + d) * 4 + i];
  return b;
}

bool aligned16(const void* p, long long stride) {
  return (reinterpret_cast<uintptr_t>(p) & 15) == 0 && stride % 2 == 0;
}

int check(int nelx, int nely, int rows, const double* B, const char* who) {
  if (nelx <= 0 || nely <= 0 || rows < 0 || B == nullptr) {
    set_error(std::string(who) + ": invalid grid or null B");
    return STTOPT_EINVAL;
  }
  return STTOPT_OK;
}

}  // namespace

}  // namespace stt

using namespace stt;

extern "C" {

int sttopt_gauss_gradients(int32_t nelx, int32_t nely, int32_t rows, const double* B,
                           const double* u, int64_t u_stride, const double* escale,
                           int64_t escale_stride, double scale, int32_t normalize,
                           double* out, int64_t out_stride, void* stream) {
  if (int e = check(nelx, nely, rows, B, "sttopt_gauss_gradients")) return e;
  if (rows == 0) return STTOPT_OK;
  if (u == nullptr || out == nullptr || !aligned16(out, out_stride)) {
    set_error("sttopt_gauss_gradients: null vector or out not 16-byte aligned");
    return STTOPT_EINVAL;
  }
  const long long nel = (long long)nelx * nely;
  const dim3 grid(ceil_div(nel, 256), rows), block(256);
  hipStream_t s = (hipStream_t)stream;
  if (normalize)
    hipLaunchKernelGGL(gauss_gradients_kernel<true>, grid, block, 0, s, nelx, nely, bmat(B), u,
                       u_stride, nullptr, 0, 1.0, out, out_stride);
  else
    hipLaunchKernelGGL(gauss_gradients_kernel<false>, grid, block, 0, s, nelx, nely, bmat(B),
                       u, u_stride, escale, escale_stride, scale, out, out_stride);
  STT_CHECK_LAUNCH("gauss_gradients_kernel");
  return STTOPT_OK;
}

int sttopt_divergence_load(int32_t nelx, int32_t nely, int32_t rows, const double* B,
                           const double* w, int64_t w_stride, const double* X,
                           int64_t X_stride, double scale, double* F, int64_t F_stride,
                           void* stream) {
  if (int e = check(nelx, nely, rows, B, "sttopt_divergence_load")) return e;
  if (rows == 0) return STTOPT_OK;
  if (X == nullptr || F == nullptr || !aligned16(X, X_stride)) {
    set_error("sttopt_divergence_load: null vector or X not 16-byte aligned");
    return STTOPT_EINVAL;
  }
  const int ntx = (nelx + 1 + kLX - 1) / kLX, nty = (nely + 1 + kLY - 1) / kLY;
  hipLaunchKernelGGL(divergence_load_kernel, dim3(ntx * nty * rows), dim3(256), 0,
                     (hipStream_t)stream, nelx, nely, bmat(B), w, w_stride, X, X_stride, scale,
                     F, F_stride, ntx, ntx * nty);
  STT_CHECK_LAUNCH("divergence_load_kernel");
  return STTOPT_OK;
}

int sttopt_gauss_dot(int32_t nelx, int32_t nely, int32_t rows, const double* B,
                     const double* X, int64_t X_stride, const double* v, int64_t v_stride,
                     double scale, double* out, int64_t out_stride, void* stream) {
  if (int e = check(nelx, nely, rows, B, "sttopt_gauss_dot")) return e;
  if (rows == 0) return STTOPT_OK;
  if (X == nullptr || v == nullptr || out == nullptr || !aligned16(X, X_stride)) {
    set_error("sttopt_gauss_dot: null vector or X not 16-byte aligned");
    return STTOPT_EINVAL;
  }
  const long long nel = (long long)nelx * nely;
  hipLaunchKernelGGL(gauss_dot_kernel, dim3(ceil_div(4 * nel, 256), rows), dim3(256), 0,
                     (hipStream_t)stream, nelx, nely, bmat(B), X, X_stride, v, v_stride, scale,
                     out, out_stride);
  STT_CHECK_LAUNCH("gauss_dot_kernel");
  return STTOPT_OK;
}

int sttopt_direction_field_backward(int32_t nelx, int32_t nely, int32_t rows,
                                    const double* B, const double* s, int64_t s_stride,
                                    const double* gY, int64_t gY_stride, double* gG,
                                    int64_t gG_stride, void* stream) {
  if (int e = check(nelx, nely, rows, B, "sttopt_direction_field_backward")) return e;
  if (rows == 0) return STTOPT_OK;
  if (s == nullptr || gY == nullptr || gG == nullptr || !aligned16(gY, gY_stride) ||
      !aligned16(gG, gG_stride)) {
    set_error("sttopt_direction_field_backward: null vector or gradient fields not 16-byte aligned");
    return STTOPT_EINVAL;
  }
  const long long nel = (long long)nelx * nely;
  hipLaunchKernelGGL(direction_backward_kernel, dim3(ceil_div(4 * nel, 256), rows), dim3(256), 0,
                     (hipStream_t)stream, nelx, nely, bmat(B), s, s_stride, gY, gY_stride, gG,
                     gG_stride);
  STT_CHECK_LAUNCH("direction_backward_kernel");
  return STTOPT_OK;
}

}  // extern "C"
