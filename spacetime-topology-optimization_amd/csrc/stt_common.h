// Shared device helpers for the gfx950 solve-and-sensitivity kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/sttopt_hip.h"

namespace stt {

constexpr int kWave = 64;
// Level-0 node tiles: 64 nodes along x (one wavefront) by 16 node rows; each of the
// block's four wavefronts owns 4 consecutive node rows (4 nodes per thread).  Every
// level-0 kernel that writes per-tile reduction partials uses this tile.
constexpr int kTileX = 64;
constexpr int kWaves = 4;
constexpr int kRowsPerThread = 4;
constexpr int kTileY = kWaves * kRowsPerThread;
constexpr int kBlock = kTileX * kWaves;
// Coarse-level stencils: one node per thread, 64 x 4 node tiles.
constexpr int kSmallTileY = 4;

void set_error(const std::string& msg);
int hip_status(hipError_t e, const char* what);

#define STT_CHECK_LAUNCH(what)                                    \
  do {                                                            \
    hipError_t e__ = hipGetLastError();                           \
    if (e__ != hipSuccess) return ::stt::hip_status(e__, what);   \
  } while (0)

// (4d) x (4d) reference element matrix, passed by value in the kernel arguments.
template <typename T>
struct KE {
  T v[64];
};

template <typename T>
inline KE<T> make_ke(const double* ke, int d) {
  KE<T> k{};
  const int n = 4 * d;
  for (int i = 0; i < n * n; ++i) k.v[i] = static_cast<T>(ke[i]);
  return k;
}

__device__ __forceinline__ bool mask_free(const uint32_t* bits, long long dof) {
  return bits == nullptr || ((bits[dof >> 5] >> (dof & 31)) & 1u);
}

// XCD-aware workgroup order.  Workgroups are handed to the eight XCDs
// round-robin; this remaps the launch index so that each XCD walks one contiguous
// run of tiles, which keeps the halo rows a tile shares with its neighbours in
// that XCD's L2.
__device__ __forceinline__ int xcd_remap(int L, int total) {
  const int xcd = L & 7, local = L >> 3;
  const int q = total >> 3, r = total & 7;
  return xcd * q + min(xcd, r) + local;
}

// Row and tile of (remapped) launch index t for launches over rows x tiles, tile-major:
// the rows of one tile are consecutive, so each XCD's band covers every row of a range
// of tiles, and rows that stop early (converged) leave the XCDs evenly loaded.
__device__ __forceinline__ void row_tile(int t, int tiles, int total, int& row, int& tt) {
  const int rows = total / tiles;
  tt = t / rows;
  row = t - tt * rows;
}

// The D dofs of a node, moved as one D-vector (D = 2: 8- or 16-byte aligned pairs).
template <typename T, int D>
struct Vec {
  T v[D];
};

template <typename T, int D>
__device__ __forceinline__ Vec<T, D> load_node(const T* base, long long node) {
  Vec<T, D> out;
  if constexpr (D == 2) {
    using V2 = HIP_vector_type<T, 2>;
    const V2 w = *reinterpret_cast<const V2*>(base + 2 * node);
    out.v[0] = w.x;
    out.v[1] = w.y;
  } else {
    out.v[0] = base[node];
  }
  return out;
}

template <typename T, int D>
__device__ __forceinline__ void store_node(T* base, long long node, const Vec<T, D>& val) {
  if constexpr (D == 2) {
    using V2 = HIP_vector_type<T, 2>;
    V2 w;
    w.x = val.v[0];
    w.y = val.v[1];
    *reinterpret_cast<V2*>(base + 2 * node) = w;
  } else {
    base[node] = val.v[0];
  }
}

// Element-local node order BL, BR, TR, TL as (row, col) offsets from the
// element's top-left node.
__device__ __forceinline__ int loc_dr(int j) { return (j == 0 || j == 1) ? 1 : 0; }
__device__ __forceinline__ int loc_dc(int j) { return (j == 1 || j == 2) ? 1 : 0; }

// Wave and block reductions in fp64 (fixed order: deterministic).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  return v;
}

// tanh(y) from the identity tanh a = (1 - e^-2a) / (1 + e^-2a) = -m / (2 + m), with
// m = expm1(-2a), a = |y| and the sign of y: one expm1 and one division (74 VALU
// instructions against 171 for the library tanh).  expm1 keeps the relative accuracy
// near 0; beyond |y| ~ 19 the result is exactly +-1.
__device__ __forceinline__ double tanh_expm1(double y) {
  const double m = expm1(-2.0 * fabs(y));
  return copysign(-m / (2.0 + m), y);
}

// Sum over a block of kBlock threads; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds /* >= kBlock/kWave */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i) s += lds[i];
  }
  return s;
}

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// A batched vector whose rows can be moved as whole 2-vectors: null, or base pointer
// aligned to the pair and an even row stride.
inline bool pair_aligned(const void* p, long long stride, size_t elem) {
  return p == nullptr ||
         (reinterpret_cast<uintptr_t>(p) % (2 * elem) == 0 && stride % 2 == 0);
}

}  // namespace stt
