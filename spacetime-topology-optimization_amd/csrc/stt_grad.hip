// Element-wise adjoint and objective kernels: the FemSolve.backward element terms
// (torch_solve.py:144-152), the strain energy (compliance.py:49-59), and the fused
// space-time compliance prologue / epilogue with its hand-derived adjoint
// (compliance.py:33-46, 137-161, 198-253; tests/reference/compliance.py:24-110).
#include <mutex>

#include "stt_kernels.h"

namespace stt {

// Gathers the 4d element values of element (er, ec) in local order BL, BR, TR, TL, one
// vector load per node (u node-aligned: the d = 2 entry points check it).
template <int D>
__device__ __forceinline__ void gather_elem(const double* u, int nnx, int er, int ec,
                                            double* ue) {
  const long long tl = (long long)er * nnx + ec;
  const long long nodes[4] = {tl + nnx, tl + nnx + 1, tl + 1, tl};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const Vec<double, D> v = load_node<double, D>(u, nodes[j]);
#pragma unroll
    for (int k = 0; k < D; ++k) ue[j * D + k] = v.v[k];
  }
}

// Copies the (4d)^2 KE entries into LDS; the caller synchronizes before reading them.
// Every lane of a wavefront then reads the same entry at once (an LDS broadcast):
// held in scalar registers, the 64 fp64 entries do not fit alongside the rest of the
// kernel, and the compiler spills them to vector lanes and reads them back one
// v_readlane at a time.
template <int D>
__device__ __forceinline__ void stage_ke(const KE<double>& ke, double* ske) {
  constexpr int K = 4 * D;
  for (int i = threadIdx.x; i < K * K; i += blockDim.x) ske[i] = ke.v[i];
}

// sum_i (sum_j a_j KE[j][i]) b_i  -- the reference's (a_e @ KE) . b_e; ke in LDS
template <int D>
__device__ __forceinline__ double ke_form(const double* ke, const double* a,
                                          const double* b) {
  constexpr int K = 4 * D;
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < K; ++j) t += a[j] * ke[j * K + i];
    s += t * b[i];
  }
  return s;
}

template <int D>
__global__ __launch_bounds__(256) void adjoint_grad_kernel(int nelx, int nely,
                                                           const double* u, long long su,
                                                           const double* lam, long long sl,
                                                           KE<double> ke, double* gd,
                                                           long long sgd, double* gr,
                                                           long long sgr) {
  constexpr int K = 4 * D;
  __shared__ double ske[K * K];
  stage_ke<D>(ke, ske);
  __syncthreads();
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  if (e >= (long long)nelx * nely) return;
  const int er = (int)(e / nelx), ec = (int)(e - (long long)er * nelx);
  double ue[K], le[K];
  gather_elem<D>(u + row * su, nelx + 1, er, ec, ue);
  gather_elem<D>(lam + row * sl, nelx + 1, er, ec, le);
  if (gd != nullptr) gd[row * sgd + e] = -ke_form<D>(ske, le, ue);
  if (gr != nullptr) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < K; ++i) s += le[i] * ue[i];
    gr[row * sgr + e] = -s / 4.0;
  }
}

template <int D>
__global__ __launch_bounds__(256) void strain_energy_kernel(int nelx, int nely,
                                                            const double* u, long long su,
                                                            KE<double> ke, double* ce,
                                                            long long sce) {
  constexpr int K = 4 * D;
  __shared__ double ske[K * K];
  stage_ke<D>(ke, ske);
  __syncthreads();
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  if (e >= (long long)nelx * nely) return;
  const int er = (int)(e / nelx), ec = (int)(e - (long long)er * nelx);
  double ue[K];
  gather_elem<D>(u + row * su, nelx + 1, er, ec, ue);
  ce[row * sce + e] = ke_form<D>(ske, ue, ue);
}

// ------------------------------------------------------- space-time stages --

struct StageParams {
  int nelx, nely, nstage;
  double beta, emin, emax, penal, fe;
};

// Per-stage constants of H_beta: a = tanh(beta t_k), den = a + tanh(beta (1 - t_k)).
struct MaskRow {
  double tk, a, den;
};

__device__ __forceinline__ MaskRow mask_row(double tk, double beta) {
  MaskRow r;
  r.tk = tk;
  // the same tanh as stage_mask's: at t = 1 its numerator then equals den exactly, so the
  // mask is exactly 0 there and the clamp does not act
  r.a = tanh_expm1(beta * tk);
  r.den = r.a + tanh_expm1(beta * (1.0 - tk));
  return r;
}

// H_beta(t_k - t) (compliance.time_mask) and its t-derivative
// (tests/reference/compliance.time_mask_derivative), zero where the clamp acts.
__device__ __forceinline__ void stage_mask(double t, const MaskRow& k, double beta, double& m,
                                           double& dm) {
  const double th = tanh_expm1(beta * (t - k.tk));
  const double v = 1.0 - (k.a + th) / k.den;
  if (v >= 0.0) {
    m = v;
    dm = -beta * (1.0 - th * th) / k.den;
  } else {
    m = 0.0;
    dm = 0.0;
  }
}

// x^p, by multiplication for the small integer exponents SIMP uses (p = 3 in practice).
__device__ __forceinline__ double pow_p(double x, double p) {
  if (p == 3.0) return x * x * x;
  if (p == 2.0) return x * x;
  if (p == 1.0) return x;
  if (p == 4.0) {
    const double y = x * x;
    return y * y;
  }
  if (p == 0.0) return 1.0;
  return pow(x, p);
}

__device__ __forceinline__ double simp(double x, const StageParams& p) {
  return p.emin + pow_p(x, p.penal) * (p.emax - p.emin);
}

// Stage densities and self-weight loads, one 64 x 4 node tile of row blockIdx.y per
// block.  x*m of the 65 x 5 elements around the tile is formed once, in LDS; the block
// writes the SIMP densities of the elements whose top-left node lies in the tile and
// the loads of the tile's nodes, F_y(node) = -fe/4 * sum of x*m over the node's
// elements, F_x = 0.  Row 0 (the whole structure) takes x unmasked and copies the
// external load.  The stage constants of the row are formed once per block.
__global__ __launch_bounds__(256) void stage_prologue_kernel(StageParams p, const double* x,
                                                             const double* t,
                                                             const double* times,
                                                             const double* fext,
                                                             double* density, double* load) {
  __shared__ double sxm[5][65];
  __shared__ MaskRow sk;
  const int nnx = p.nelx + 1, nny = p.nely + 1;
  const long long ndof = 2LL * nnx * nny;
  const long long nel = (long long)p.nelx * p.nely;
  const int tnx = (nnx + 63) / 64;
  const int bx = blockIdx.x % tnx, by = blockIdx.x / tnx;
  const int c0 = bx * 64, r0 = by * 4;
  const int lr = threadIdx.x >> 6, lc = threadIdx.x & 63;
  const int c = c0 + lc, r = r0 + lr;
  const int row = blockIdx.y;
  const bool own = r < nny && c < nnx;
  const long long node = (long long)r * nnx + c;
  const bool own_e = r < p.nely && c < p.nelx;  // the element whose top-left node is (r, c)
  const long long e = (long long)r * p.nelx + c;
  if (row == 0) {
    if (own) {
      load[2 * node] = fext[2 * node];
      load[2 * node + 1] = fext[2 * node + 1];
    }
    if (own_e) density[e] = simp(x[e], p);
    return;  // uniform: the whole block is row 0
  }
  if (threadIdx.x == 0) sk = mask_row(times[row - 1], p.beta);
  __syncthreads();
  const MaskRow k = sk;
  for (int i = threadIdx.x; i < 5 * 65; i += 256) {
    const int er = r0 - 1 + i / 65, ec = c0 - 1 + i % 65;
    double v = 0.0;
    if (er >= 0 && er < p.nely && ec >= 0 && ec < p.nelx) {
      const long long ei = (long long)er * p.nelx + ec;
      double m, dm;
      stage_mask(t[ei], k, p.beta, m, dm);
      v = x[ei] * m;
    }
    sxm[i / 65][i % 65] = v;
  }
  __syncthreads();
  if (own_e) density[row * nel + e] = simp(sxm[lr + 1][lc + 1], p);
  if (!own) return;
  const double acc = sxm[lr][lc] + sxm[lr][lc + 1] + sxm[lr + 1][lc] + sxm[lr + 1][lc + 1];
  load[row * ndof + 2 * node] = 0.0;
  load[row * ndof + 2 * node + 1] = -(p.fe / 4.0) * acc;
}

// Per-row compliance partials: sum_e density_e * (U_e^T KE U_e), with the SIMP
// densities of the prologue (SIMP(x m) of the row's stage mask m).
__global__ __launch_bounds__(256) void stage_comp_kernel(StageParams p,
                                                         const double* __restrict__ density,
                                                         KE<double> ke, const double* u,
                                                         double* part, int nblk) {
  __shared__ double red[4];
  __shared__ double ske[64];
  const long long nel = (long long)p.nelx * p.nely;
  const long long ndof = 2LL * (p.nelx + 1) * (p.nely + 1);
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  stage_ke<2>(ke, ske);
  __syncthreads();
  double v = 0.0;
  if (e < nel) {
    const int er = (int)(e / p.nelx), ec = (int)(e - (long long)er * p.nelx);
    double ue[8];
    gather_elem<2>(u + row * ndof, p.nelx + 1, er, ec, ue);
    v = density[row * nel + e] * ke_form<2>(ske, ue, ue);
  }
  const double s = block_sum(v, red);
  if (threadIdx.x == 0) part[(long long)row * nblk + blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void row_finalize_kernel(const double* part, int nblk,
                                                           double* out) {
  __shared__ double red[4];
  const int row = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) acc += part[(long long)row * nblk + i];
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) out[row] = s;
}

// Hand-derived adjoint of sum_row g[row] * c_row with lambda = 2 U:
//   row 0:  dc/dx_e = -p dE x^(p-1) c_e
//   row k:  with xt = x m, A = -p dE xt^(p-1) c_e, G = -2 (C^T U_y)_e,
//           dc/dx_e = (A + G) m,  dc/dt_e = (A + G) x m'
__global__ __launch_bounds__(256) void stage_adjoint_kernel(StageParams p, const double* x,
                                                            const double* t,
                                                            const double* times,
                                                            KE<double> ke, const double* u,
                                                            const double* g, double* dx,
                                                            double* dt) {
  constexpr int kChunk = 64;  // stage constants staged in LDS per chunk of stages
  __shared__ MaskRow rows[kChunk];
  __shared__ double ske[64];
  stage_ke<2>(ke, ske);
  __syncthreads();
  const long long nel = (long long)p.nelx * p.nely;
  const long long ndof = 2LL * (p.nelx + 1) * (p.nely + 1);
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool own = e < nel;
  const long long ee = own ? e : 0;
  const int er = (int)(ee / p.nelx), ec = (int)(ee - (long long)er * p.nelx);
  const double de = p.emax - p.emin;
  const double xe = x[ee], te = p.nstage > 0 ? t[ee] : 0.0;
  double ue[8];
  gather_elem<2>(u, p.nelx + 1, er, ec, ue);
  double gx = g[0] * (-p.penal * de * pow_p(xe, p.penal - 1.0) * ke_form<2>(ske, ue, ue));
  double gt = 0.0;
  for (int k0 = 1; k0 <= p.nstage; k0 += kChunk) {
    const int nk = min(kChunk, p.nstage - k0 + 1);
    __syncthreads();  // the previous chunk's constants are consumed
    if (threadIdx.x < nk) rows[threadIdx.x] = mask_row(times[k0 - 1 + threadIdx.x], p.beta);
    __syncthreads();
    for (int j = 0; j < nk; ++j) {
      const int k = k0 + j;
      const double gk = g[k];
      gather_elem<2>(u + k * ndof, p.nelx + 1, er, ec, ue);
      double m, dm;
      stage_mask(te, rows[j], p.beta, m, dm);
      const double xt = xe * m;
      const double A = -p.penal * de * pow_p(xt, p.penal - 1.0) * ke_form<2>(ske, ue, ue);
      // (C^T U_y)_e = fe/4 * sum of the element's four nodal y-displacements
      const double cu = (p.fe / 4.0) * (ue[1] + ue[3] + ue[5] + ue[7]);
      const double s = A - 2.0 * cu;
      gx += gk * s * m;
      gt += gk * s * xe * dm;
    }
  }
  if (own) {
    dx[e] = gx;
    dt[e] = gt;
  }
}

// Device scratch of sttopt_stage_epilogue_noalloc: the per-block compliance partials
// live in this fixed array instead of a caller buffer, so that entry point allocates
// nothing.  Every GPU holds its own copy at its own address until the process ends.
// In host code the array's name is only a stand-in for the symbol: kernels receive the
// device address that hipGetSymbolAddress returns for it.
static __device__ __align__(256) double noalloc_scratch[STTOPT_NOALLOC_SCRATCH_DOUBLES];

constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static double* g_noalloc_ptr[kMaxDevices] = {};

// Device address of noalloc_scratch on the current device, looked up on the first call
// per device (that device current) and kept.
static int noalloc_scratch_ptr(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_status(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("sttopt_stage_epilogue_noalloc: device index beyond the supported 64");
    return STTOPT_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
    if (e != hipSuccess) return hip_status(e, "hipGetSymbolAddress(noalloc_scratch)");
    if (p == nullptr) {
      set_error("sttopt_stage_epilogue_noalloc: no device address for the static scratch");
      return STTOPT_EHIP;
    }
    g_noalloc_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_noalloc_ptr[dev];
  return STTOPT_OK;
}

}  // namespace stt

using namespace stt;

extern "C" {

int sttopt_adjoint_grad(int32_t d, int32_t nelx, int32_t nely, int32_t rows, const double* u,
                        int64_t u_stride, const double* lam, int64_t lam_stride,
                        const double* ke, double* grad_density, int64_t gd_stride,
                        double* grad_reaction, int64_t gr_stride, void* stream) {
  if ((d != 1 && d != 2) || nelx <= 0 || nely <= 0 || u == nullptr || lam == nullptr ||
      ke == nullptr || rows < 0) {
    set_error("sttopt_adjoint_grad: invalid arguments");
    return STTOPT_EINVAL;
  }
  if (d == 2 && !(pair_aligned(u, u_stride, sizeof(double)) &&
                  pair_aligned(lam, lam_stride, sizeof(double)))) {
    set_error("sttopt_adjoint_grad: d = 2 vectors need node-aligned pointers and even row strides");
    return STTOPT_EINVAL;
  }
  if (rows == 0 || (grad_density == nullptr && grad_reaction == nullptr)) return STTOPT_OK;
  const long long nel = (long long)nelx * nely;
  dim3 grid(ceil_div(nel, 256), rows), block(256);
  const KE<double> k = make_ke<double>(ke, d);
  hipStream_t s = (hipStream_t)stream;
  if (d == 2)
    hipLaunchKernelGGL(adjoint_grad_kernel<2>, grid, block, 0, s, nelx, nely, u, u_stride,
                       lam, lam_stride, k, grad_density, gd_stride, grad_reaction, gr_stride);
  else
    hipLaunchKernelGGL(adjoint_grad_kernel<1>, grid, block, 0, s, nelx, nely, u, u_stride,
                       lam, lam_stride, k, grad_density, gd_stride, grad_reaction, gr_stride);
  STT_CHECK_LAUNCH("adjoint_grad_kernel");
  return STTOPT_OK;
}

int sttopt_strain_energy(int32_t d, int32_t nelx, int32_t nely, int32_t rows, const double* u,
                         int64_t u_stride, const double* ke, double* ce, int64_t ce_stride,
                         void* stream) {
  if ((d != 1 && d != 2) || nelx <= 0 || nely <= 0 || u == nullptr || ke == nullptr ||
      ce == nullptr || rows < 0) {
    set_error("sttopt_strain_energy: invalid arguments");
    return STTOPT_EINVAL;
  }
  if (d == 2 && !pair_aligned(u, u_stride, sizeof(double))) {
    set_error("sttopt_strain_energy: d = 2 vectors need node-aligned pointers and even row strides");
    return STTOPT_EINVAL;
  }
  if (rows == 0) return STTOPT_OK;
  const long long nel = (long long)nelx * nely;
  dim3 grid(ceil_div(nel, 256), rows), block(256);
  const KE<double> k = make_ke<double>(ke, d);
  hipStream_t s = (hipStream_t)stream;
  if (d == 2)
    hipLaunchKernelGGL(strain_energy_kernel<2>, grid, block, 0, s, nelx, nely, u, u_stride, k,
                       ce, ce_stride);
  else
    hipLaunchKernelGGL(strain_energy_kernel<1>, grid, block, 0, s, nelx, nely, u, u_stride, k,
                       ce, ce_stride);
  STT_CHECK_LAUNCH("strain_energy_kernel");
  return STTOPT_OK;
}

int sttopt_stage_prologue(int32_t nelx, int32_t nely, int32_t nstage, const double* x,
                          const double* t, const double* stage_times, double beta,
                          double emin, double emax, double penal, double fe,
                          const double* f_ext, double* density, double* load, void* stream) {
  if (nelx <= 0 || nely <= 0 || nstage < 0 || !x || !f_ext || !density || !load ||
      (nstage > 0 && (!t || !stage_times))) {
    set_error("sttopt_stage_prologue: invalid arguments");
    return STTOPT_EINVAL;
  }
  StageParams p{nelx, nely, nstage, beta, emin, emax, penal, fe};
  hipStream_t s = (hipStream_t)stream;
  const int tiles = ceil_div(nelx + 1, 64) * ceil_div(nely + 1, 4);
  hipLaunchKernelGGL(stage_prologue_kernel, dim3(tiles, 1 + nstage), dim3(256), 0, s, p, x,
                     t, stage_times, f_ext, density, load);
  STT_CHECK_LAUNCH("stage prologue");
  return STTOPT_OK;
}

int64_t sttopt_stage_scratch_count(int32_t nelx, int32_t nely, int32_t nstage) {
  const long long nel = (long long)nelx * nely;
  return (long long)(1 + nstage) * ceil_div(nel, 256);
}

int sttopt_stage_epilogue(int32_t nelx, int32_t nely, int32_t nstage, const double* x,
                          const double* t, const double* stage_times, double beta,
                          double emin, double emax, double penal, double fe, const double* ke,
                          const double* u, const double* density, const double* g,
                          double* comp, double* dx, double* dt, double* scratch,
                          void* stream) {
  if (nelx <= 0 || nely <= 0 || nstage < 0 || !x || !ke || !u ||
      (nstage > 0 && (!t || !stage_times))) {
    set_error("sttopt_stage_epilogue: invalid arguments");
    return STTOPT_EINVAL;
  }
  if (!pair_aligned(u, 2, sizeof(double))) {
    set_error("sttopt_stage_epilogue: u needs a node-aligned (16-byte) pointer");
    return STTOPT_EINVAL;
  }
  StageParams p{nelx, nely, nstage, beta, emin, emax, penal, fe};
  const KE<double> k = make_ke<double>(ke, 2);
  const long long nel = (long long)nelx * nely;
  const int nblk = ceil_div(nel, 256);
  hipStream_t s = (hipStream_t)stream;
  if (comp != nullptr) {
    if (scratch == nullptr || density == nullptr) {
      set_error("sttopt_stage_epilogue: compliance needs scratch and the stage densities");
      return STTOPT_EINVAL;
    }
    hipLaunchKernelGGL(stage_comp_kernel, dim3(nblk, 1 + nstage), dim3(256), 0, s, p, density,
                       k, u, scratch, nblk);
    hipLaunchKernelGGL(row_finalize_kernel, dim3(1 + nstage), dim3(256), 0, s, scratch, nblk,
                       comp);
    STT_CHECK_LAUNCH("stage compliance");
  }
  if (dx != nullptr || dt != nullptr) {
    if (g == nullptr || dx == nullptr || dt == nullptr) {
      set_error("sttopt_stage_epilogue: the adjoint needs g, dx and dt");
      return STTOPT_EINVAL;
    }
    hipLaunchKernelGGL(stage_adjoint_kernel, dim3(nblk), dim3(256), 0, s, p, x, t,
                       stage_times, k, u, g, dx, dt);
    STT_CHECK_LAUNCH("stage adjoint");
  }
  return STTOPT_OK;
}

int sttopt_stage_epilogue_noalloc(int32_t nelx, int32_t nely, int32_t nstage,
                                  const double* x, const double* t,
                                  const double* stage_times, double beta, double emin,
                                  double emax, double penal, double fe, const double* ke,
                                  const double* u, const double* density, const double* g,
                                  double* comp, double* dx, double* dt, void* stream) {
  double* scratch = nullptr;
  if (comp != nullptr) {
    if (nelx <= 0 || nely <= 0 || nstage < 0) {
      set_error("sttopt_stage_epilogue_noalloc: invalid arguments");
      return STTOPT_EINVAL;
    }
    const int64_t need = sttopt_stage_scratch_count(nelx, nely, nstage);
    if (need > STTOPT_NOALLOC_SCRATCH_DOUBLES) {
      set_error("sttopt_stage_epilogue_noalloc: " + std::to_string(nelx) + " x " +
                std::to_string(nely) + " with " + std::to_string(nstage) + " stages needs " +
                std::to_string(need) + " doubles of scratch; the static buffer holds " +
                std::to_string(STTOPT_NOALLOC_SCRATCH_DOUBLES) +
                " (use sttopt_stage_epilogue with a caller buffer)");
      return STTOPT_EINVAL;
    }
    const int st = noalloc_scratch_ptr(&scratch);
    if (st != STTOPT_OK) return st;
  }
  // every partial the compliance kernel reads is written by the same call first, so
  // the array needs no clearing between calls
  return sttopt_stage_epilogue(nelx, nely, nstage, x, t, stage_times, beta, emin, emax,
                               penal, fe, ke, u, density, g, comp, dx, dt, scratch, stream);
}

}  // extern "C"
