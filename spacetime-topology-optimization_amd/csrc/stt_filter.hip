// Density filter, symmetric sparse products and the Heaviside projection of the
// density field (sttopt/filters.py:38-76,108-131; sttopt/torch_util.py:157-188).
//
// The density filter H[e1, e2] = max(0, rmin - |e1 - e2|) over the square of offsets
// |di|, |dj| <= ceil(rmin) - 1 is translation invariant, so it runs matrix-free as a
// radius-R stencil on the element grid: a workgroup stages a 64 x 16 element tile and
// its R-wide halo in LDS (zeros outside the grid, which drops the same neighbours the
// matrix does) together with the (2R+1)^2 weights, and each thread sums 4 elements in
// the matrix's column order (dj outer, di inner).  General symmetric CSR matrices go
// through a row-per-thread CSR product.  Both take an optional per-element scale s:
// y = (H x) / s after the product or y = H (x / s) before it (the adjoint of the
// first for symmetric H).
#include <cfloat>
#include <cmath>

#include "stt_kernels.h"

namespace stt {

namespace {

constexpr int kFX = 64, kFY = 16, kFRows = 4;  // tile; rows per thread
constexpr int kMaxR = 8;                       // largest stencil radius staged in LDS

enum { kScaleNone = 0, kScalePost = 1, kScalePre = 2 };

__global__ __launch_bounds__(256) void density_filter_kernel(
    int nelx, int nely, int R, double rmin, const double* __restrict__ x, long long sx,
    const double* __restrict__ sc, int mode, double* __restrict__ y, long long sy, int ntx,
    int tiles) {
  __shared__ double sW[(2 * kMaxR + 1) * (2 * kMaxR + 1)];
  __shared__ double sX[kFY + 2 * kMaxR][kFX + 2 * kMaxR];
  const int row = blockIdx.x / tiles;
  const int t = blockIdx.x - row * tiles;
  const int by = t / ntx, bx = t - by * ntx;
  const int x0 = bx * kFX, y0 = by * kFY;
  const int W = kFX + 2 * R, H = kFY + 2 * R, K = 2 * R + 1;
  const double* xr = x + row * sx;
  for (int i = threadIdx.x; i < K * K; i += blockDim.x) {
    const int dj = i / K - R, di = i % K - R;
    sW[i] = fmax(0.0, rmin - sqrt((double)(di * di + dj * dj)));
  }
  // staging: wavefront w takes halo rows w, w + 4, ...; lane l columns l and 64 + l
  // (the latter for l < 2R)
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int ly = w; ly < H; ly += 4) {
    const int gy = y0 - R + ly;
    const bool rok = gy >= 0 && gy < nely;
    for (int lx = lane; lx < W; lx += 64) {
      const int gx = x0 - R + lx;
      double v = 0.0;
      if (rok && gx >= 0 && gx < nelx) {
        const long long e = (long long)gy * nelx + gx;
        v = xr[e];
        if (mode == kScalePre) v /= sc[e];
      }
      sX[ly][lx] = v;
    }
  }
  __syncthreads();
  const int gx = x0 + lane;
  if (gx >= nelx) return;
#pragma unroll
  for (int j = 0; j < kFRows; ++j) {
    const int ly = w * kFRows + j, gy = y0 + ly;
    if (gy >= nely) break;
    double acc = 0.0;
    for (int dj = -R; dj <= R; ++dj) {
      const double* wr = sW + (dj + R) * K + R;
      const double* xrow = &sX[ly + R + dj][lane + R];
      for (int di = -R; di <= R; ++di) acc += wr[di] * xrow[di];
    }
    const long long e = (long long)gy * nelx + gx;
    if (mode == kScalePost) acc /= sc[e];
    y[row * sy + e] = acc;
  }
}

// Radius R known at compile time: each thread holds its (4 + 2R) x (2R + 1) input window
// in registers (read once from LDS) and the (2R + 1)^2 weights come from the kernel
// arguments (uniform, scalar registers), so a tap costs one FMA and no LDS read.
template <int R>
struct FilterWeights {
  double v[(2 * R + 1) * (2 * R + 1)];
};

template <int R>
__global__ __launch_bounds__(256) void density_filter_r_kernel(
    int nelx, int nely, FilterWeights<R> fw, const double* __restrict__ x, long long sx,
    const double* __restrict__ sc, int mode, double* __restrict__ y, long long sy, int ntx,
    int tiles) {
  constexpr int W = kFX + 2 * R, H = kFY + 2 * R, K = 2 * R + 1;
  __shared__ double sX[H][W];
  const int row = blockIdx.x / tiles;
  const int t = blockIdx.x - row * tiles;
  const int by = t / ntx, bx = t - by * ntx;
  const int x0 = bx * kFX, y0 = by * kFY;
  const double* xr = x + row * sx;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int ly = w; ly < H; ly += 4) {
    const int gy = y0 - R + ly;
    const bool rok = gy >= 0 && gy < nely;
#pragma unroll
    for (int lx = lane; lx < W; lx += 64) {
      const int gx = x0 - R + lx;
      double v = 0.0;
      if (rok && gx >= 0 && gx < nelx) {
        const long long e = (long long)gy * nelx + gx;
        v = xr[e];
        if (mode == kScalePre) v /= sc[e];
      }
      sX[ly][lx] = v;
    }
  }
  __syncthreads();
  const int gx = x0 + lane;
  const int i0 = w * kFRows;
  double win[kFRows + 2 * R][K];
#pragma unroll
  for (int r = 0; r < kFRows + 2 * R; ++r) {
#pragma unroll
    for (int c = 0; c < K; ++c) win[r][c] = sX[i0 + r][lane + c];
  }
  if (gx >= nelx) return;
#pragma unroll
  for (int j = 0; j < kFRows; ++j) {
    const int gy = y0 + i0 + j;
    if (gy >= nely) break;
    double acc = 0.0;
#pragma unroll
    for (int dj = 0; dj < K; ++dj) {
#pragma unroll
      for (int di = 0; di < K; ++di) acc += fw.v[dj * K + di] * win[j + dj][di];
    }
    const long long e = (long long)gy * nelx + gx;
    if (mode == kScalePost) acc /= sc[e];
    y[row * sy + e] = acc;
  }
}

template <int R>
void launch_filter_r(int nelx, int nely, int rows, double rmin, const double* x, long long sx,
                     const double* sc, int mode, double* y, long long sy, hipStream_t s) {
  FilterWeights<R> fw;
  constexpr int K = 2 * R + 1;
  for (int dj = -R; dj <= R; ++dj)
    for (int di = -R; di <= R; ++di)
      fw.v[(dj + R) * K + di + R] = std::fmax(0.0, rmin - std::sqrt((double)(di * di + dj * dj)));
  const int ntx = (nelx + kFX - 1) / kFX, nty = (nely + kFY - 1) / kFY;
  hipLaunchKernelGGL(density_filter_r_kernel<R>, dim3(ntx * nty * rows), dim3(256), 0, s, nelx,
                     nely, fw, x, sx, sc, mode, y, sy, ntx, ntx * nty);
}

// Radius beyond the LDS tile: every tap gathered from global memory (cache served).
__global__ __launch_bounds__(256) void density_filter_direct_kernel(
    int nelx, int nely, int R, double rmin, const double* __restrict__ x, long long sx,
    const double* __restrict__ sc, int mode, double* __restrict__ y, long long sy) {
  const int row = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nelx * nely) return;
  const int gy = (int)(e / nelx), gx = (int)(e - (long long)gy * nelx);
  const double* xr = x + row * sx;
  double acc = 0.0;
  for (int dj = -R; dj <= R; ++dj) {
    const int yy = gy + dj;
    if (yy < 0 || yy >= nely) continue;
    for (int di = -R; di <= R; ++di) {
      const int xx = gx + di;
      if (xx < 0 || xx >= nelx) continue;
      const double wgt = fmax(0.0, rmin - sqrt((double)(di * di + dj * dj)));
      const long long f = (long long)yy * nelx + xx;
      double v = xr[f];
      if (mode == kScalePre) v /= sc[f];
      acc += wgt * v;
    }
  }
  if (mode == kScalePost) acc /= sc[e];
  y[row * sy + e] = acc;
}

// y = A x for a CSR matrix (one row per thread, entries in stored order).
__global__ __launch_bounds__(256) void csr_matvec_kernel(
    long long n, const long long* __restrict__ ptr, const long long* __restrict__ col,
    const double* __restrict__ val, const double* __restrict__ x, long long sx,
    const double* __restrict__ sc, int mode, double* __restrict__ y, long long sy) {
  const int row = blockIdx.y;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* xr = x + row * sx;
  double acc = 0.0;
  for (long long k = ptr[i]; k < ptr[i + 1]; ++k) {
    const long long j = col[k];
    double v = xr[j];
    if (mode == kScalePre) v /= sc[j];
    acc += val[k] * v;
  }
  if (mode == kScalePost) acc /= sc[i];
  y[row * sy + i] = acc;
}

// Smoothed threshold (tanh(b eta) + tanh(b (x - eta))) / (tanh(b eta) + tanh(b (1 - eta)))
// or its derivative b (1 - tanh^2(b (x - eta))) / denominator; beta = inf: the step
// with value 1/2 at eta (the derivative is not defined there and is not offered).
// te = tanh(beta eta) and denom = te + tanh(beta (1 - eta)) come from the host.
__global__ __launch_bounds__(256) void heaviside_kernel(long long n, double beta, double eta,
                                                        double te, double denom,
                                                        int derivative,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ g,
                                                        double* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = x[i];
  double out;
  if (isinf(beta)) {
    out = v > eta ? 1.0 : (v < eta ? 0.0 : 0.5);
  } else {
    const double th = tanh_expm1(beta * (v - eta));
    out = derivative ? beta * (1.0 - th * th) / denom : (te + th) / denom;
  }
  if (g != nullptr) out *= g[i];
  y[i] = out;
}

}  // namespace

}  // namespace stt

using namespace stt;

extern "C" {

int sttopt_density_filter(int32_t nelx, int32_t nely, int32_t rows, double rmin,
                          const double* x, int64_t x_stride, const double* scale,
                          int32_t scale_mode, double* y, int64_t y_stride, void* stream) {
  if (nelx <= 0 || nely <= 0 || rows < 0 || !(rmin > 0.0) || x == nullptr || y == nullptr ||
      scale_mode < 0 || scale_mode > 2 || (scale_mode != 0 && scale == nullptr)) {
    set_error("sttopt_density_filter: invalid arguments");
    return STTOPT_EINVAL;
  }
  if (rows == 0) return STTOPT_OK;
  const int R = (int)std::ceil(rmin) - 1;
  hipStream_t s = (hipStream_t)stream;
  if (R >= 1 && R <= 4) {
    switch (R) {
      case 1: launch_filter_r<1>(nelx, nely, rows, rmin, x, x_stride, scale, scale_mode, y, y_stride, s); break;
      case 2: launch_filter_r<2>(nelx, nely, rows, rmin, x, x_stride, scale, scale_mode, y, y_stride, s); break;
      case 3: launch_filter_r<3>(nelx, nely, rows, rmin, x, x_stride, scale, scale_mode, y, y_stride, s); break;
      default: launch_filter_r<4>(nelx, nely, rows, rmin, x, x_stride, scale, scale_mode, y, y_stride, s); break;
    }
  } else if (R <= kMaxR) {
    const int ntx = (nelx + kFX - 1) / kFX, nty = (nely + kFY - 1) / kFY;
    hipLaunchKernelGGL(density_filter_kernel, dim3(ntx * nty * rows), dim3(256), 0, s, nelx,
                       nely, R, rmin, x, x_stride, scale, scale_mode, y, y_stride, ntx,
                       ntx * nty);
  } else {
    const long long nel = (long long)nelx * nely;
    hipLaunchKernelGGL(density_filter_direct_kernel, dim3(ceil_div(nel, 256), rows), dim3(256),
                       0, s, nelx, nely, R, rmin, x, x_stride, scale, scale_mode, y, y_stride);
  }
  STT_CHECK_LAUNCH("density_filter_kernel");
  return STTOPT_OK;
}

int sttopt_csr_matvec(int64_t n, const int64_t* row_ptr, const int64_t* col_idx,
                      const double* values, int32_t rows, const double* x, int64_t x_stride,
                      const double* scale, int32_t scale_mode, double* y, int64_t y_stride,
                      void* stream) {
  if (n < 0 || rows < 0 || scale_mode < 0 || scale_mode > 2 ||
      (scale_mode != 0 && scale == nullptr)) {
    set_error("sttopt_csr_matvec: invalid arguments");
    return STTOPT_EINVAL;
  }
  if (n == 0 || rows == 0) return STTOPT_OK;
  if (row_ptr == nullptr || col_idx == nullptr || values == nullptr || x == nullptr ||
      y == nullptr) {
    set_error("sttopt_csr_matvec: null pointer");
    return STTOPT_EINVAL;
  }
  hipLaunchKernelGGL(csr_matvec_kernel, dim3(ceil_div(n, 256), rows), dim3(256), 0,
                     (hipStream_t)stream, (long long)n,
                     reinterpret_cast<const long long*>(row_ptr),
                     reinterpret_cast<const long long*>(col_idx), values, x, x_stride, scale,
                     scale_mode, y, y_stride);
  STT_CHECK_LAUNCH("csr_matvec_kernel");
  return STTOPT_OK;
}

int sttopt_heaviside(int64_t n, double beta, double eta, int32_t derivative, const double* x,
                     const double* g, double* y, void* stream) {
  if (n < 0 || x == nullptr || y == nullptr || (derivative && std::isinf(beta))) {
    set_error("sttopt_heaviside: invalid arguments");
    return STTOPT_EINVAL;
  }
  if (n == 0) return STTOPT_OK;
  const double te = std::isinf(beta) ? 0.0 : std::tanh(beta * eta);
  const double denom = std::isinf(beta) ? 1.0 : te + std::tanh(beta * (1.0 - eta));
  hipLaunchKernelGGL(heaviside_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, (long long)n, beta, eta, te, denom, derivative, x, g,
                     y);
  STT_CHECK_LAUNCH("heaviside_kernel");
  return STTOPT_OK;
}

}  // extern "C"
