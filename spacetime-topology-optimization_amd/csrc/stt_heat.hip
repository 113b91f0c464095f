// Gauss-point gradient fields of the scalar (one dof per node) Q4 mesh behind the
// virtual-heat time fields, and their adjoints (sttopt/virtual_heat.py:290-347):
//
//   gauss gradients  G[e][g][d] = sum_i B[g][d][i] u[node_i(e)]
//   direction field  Y = G / sqrt(|G|^2 + DBL_MIN) per Gauss point
//   divergence load  F[n] = 1/4 sum_{e at n} w_e sum_{g,d} X[e][g][d] B[g][d][loc_n(e)]
//
// B holds the bilinear shape-function gradients at the 2 x 2 Gauss points in the local
// node order BL, BR, TR, TL.  Element kernels run one element per thread (consecutive
// threads on consecutive elements of a row, so the two node rows an element reads are
// coalesced); the load is a node-centric gather over the (up to) four elements around
// a node, in a fixed order, so it needs no atomics and is deterministic.
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

__device__ __forceinline__ void gauss_grad(const Bmat& B, const double (&ue)[4],
                                           double (&g)[4][2]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) s += B.v[q][d][i] * ue[i];
      g[q][d] = s;
    }
  }
}

// out = scale * escale_e * G (escale may be null), or the direction field when NORM.
template <bool NORM>
__global__ __launch_bounds__(256) void gauss_gradients_kernel(
    int nelx, int nely, Bmat B, const double* __restrict__ u, long long su,
    const double* __restrict__ escale, long long ses, double scale, double* __restrict__ out,
    long long so) {
  const int row = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nelx * nely) return;
  const int r = (int)(e / nelx), c = (int)(e - (long long)r * nelx);
  double ue[4], g[4][2];
  element_nodes(u + row * su, nelx + 1, r, c, ue);
  gauss_grad(B, ue, g);
  double f = scale;
  if (!NORM && escale != nullptr) f *= escale[row * ses + e];
  double2* o = reinterpret_cast<double2*>(out + row * so + 8 * e);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double2 v;
    if (NORM) {
      const double inv = 1.0 / sqrt(g[q][0] * g[q][0] + g[q][1] * g[q][1] + DBL_MIN);
      v.x = g[q][0] * inv;
      v.y = g[q][1] * inv;
    } else {
      v.x = f * g[q][0];
      v.y = f * g[q][1];
    }
    o[q] = v;
  }
}

// F[n] = scale * sum over the elements around node n of w_e sum_{g,d} X[e][g][d]
// B[g][d][loc]; w may be null (1).  Elements in the order up-left, up-right,
// down-left, down-right of the node, where the node is their BR, BL, TR, TL corner.
__global__ __launch_bounds__(256) void divergence_load_kernel(
    int nelx, int nely, Bmat B, const double* __restrict__ w, long long sw,
    const double* __restrict__ X, long long sx, double scale, double* __restrict__ F,
    long long sf) {
  const int row = blockIdx.y;
  const int nnx = nelx + 1;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (long long)nnx * (nely + 1)) return;
  const int r = (int)(n / nnx), c = (int)(n - (long long)r * nnx);
  const double* Xr = X + row * sx;
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int er = r - 1 + (q >> 1), ec = c - 1 + (q & 1);
    if (er < 0 || er >= nely || ec < 0 || ec >= nelx) continue;
    constexpr int locs[4] = {1, 0, 2, 3};  // BR of up-left, BL of up-right, ...
    const int loc = locs[q];
    const long long e = (long long)er * nelx + ec;
    const double2* xe = reinterpret_cast<const double2*>(Xr + 8 * e);
    double s = 0.0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const double2 v = xe[g];
      s += v.x * B.v[g][0][loc] + v.y * B.v[g][1][loc];
    }
    if (w != nullptr) s *= w[row * sw + e];
    acc += s;
  }
  F[row * sf + n] = scale * acc;
}

// out_e = scale * sum_{g,d} X[e][g][d] G(v)[e][g][d]: the weight gradient of the load.
__global__ __launch_bounds__(256) void gauss_dot_kernel(int nelx, int nely, Bmat B,
                                                        const double* __restrict__ X,
                                                        long long sx,
                                                        const double* __restrict__ v,
                                                        long long sv, double scale,
                                                        double* __restrict__ out,
                                                        long long so) {
  const int row = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nelx * nely) return;
  const int r = (int)(e / nelx), c = (int)(e - (long long)r * nelx);
  double ve[4], g[4][2];
  element_nodes(v + row * sv, nelx + 1, r, c, ve);
  gauss_grad(B, ve, g);
  const double2* xe = reinterpret_cast<const double2*>(X + row * sx + 8 * e);
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double2 x = xe[q];
    s += x.x * g[q][0] + x.y * g[q][1];
  }
  out[row * so + e] = scale * s;
}

// Adjoint of the direction field at the Gauss-point level: with G the gradients of s,
// n = sqrt(|G|^2 + DBL_MIN) and Y = G / n, gG = (gY - Y (Y . gY)) / n.
__global__ __launch_bounds__(256) void direction_backward_kernel(
    int nelx, int nely, Bmat B, const double* __restrict__ s, long long ss,
    const double* __restrict__ gY, long long sgy, double* __restrict__ gG, long long sgg) {
  const int row = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nelx * nely) return;
  const int r = (int)(e / nelx), c = (int)(e - (long long)r * nelx);
  double se[4], g[4][2];
  element_nodes(s + row * ss, nelx + 1, r, c, se);
  gauss_grad(B, se, g);
  const double2* gy = reinterpret_cast<const double2*>(gY + row * sgy + 8 * e);
  double2* o = reinterpret_cast<double2*>(gG + row * sgg + 8 * e);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double inv = 1.0 / sqrt(g[q][0] * g[q][0] + g[q][1] * g[q][1] + DBL_MIN);
    const double yx = g[q][0] * inv, yy = g[q][1] * inv;
    const double2 b = gy[q];
    const double dot = yx * b.x + yy * b.y;
    double2 v;
    v.x = (b.x - yx * dot) * inv;
    v.y = (b.y - yy * dot) * inv;
    o[q] = v;
  }
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
  const long long nn = (long long)(nelx + 1) * (nely + 1);
  hipLaunchKernelGGL(divergence_load_kernel, dim3(ceil_div(nn, 256), rows), dim3(256), 0,
                     (hipStream_t)stream, nelx, nely, bmat(B), w, w_stride, X, X_stride, scale,
                     F, F_stride);
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
  hipLaunchKernelGGL(gauss_dot_kernel, dim3(ceil_div(nel, 256), rows), dim3(256), 0,
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
  hipLaunchKernelGGL(direction_backward_kernel, dim3(ceil_div(nel, 256), rows), dim3(256), 0,
                     (hipStream_t)stream, nelx, nely, bmat(B), s, s_stride, gY, gY_stride, gG,
                     gG_stride);
  STT_CHECK_LAUNCH("direction_backward_kernel");
  return STTOPT_OK;
}

}  // extern "C"
