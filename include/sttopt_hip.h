/*
 * sttopt_hip.h -- C ABI of the MI355X (gfx950) solve-and-sensitivity library.
 *
 * Every entry point takes plain device pointers, element counts and a HIP stream
 * (passed as void*), and returns an int status (STTOPT_OK or a negative code; the
 * message of the last failure on the calling thread is sttopt_last_error()).  No
 * entry point allocates device memory: the caller owns every buffer, sized with the
 * *_bytes / *_count helpers below.  All vectors are batched: row b of a batched
 * array starts at base + b * stride, and a stride of 0 shares one row across the
 * batch.
 *
 * Mesh conventions (sttopt/fem.py:60-95, sttopt/conventions.md:24-53): an
 * nelx x nely grid of unit Q4 elements, element e at (row e / nelx, col e % nelx)
 * with row 0 at the top, node (r, c) numbered r * (nelx + 1) + c, d dofs per node
 * interleaved (dof = d * node + component), element-local node order
 * bottom-left, bottom-right, top-right, top-left.  Free-dof masks are packed bits:
 * dof i is free when (mask[i >> 5] >> (i & 31)) & 1.  With d = 2 the kernels move
 * a node's two dofs as one vector: per-dof vectors (u, y, b, rdiag, PCG vectors and
 * work) need pointers aligned to 2 * sizeof(element) and even row strides
 * (STTOPT_EINVAL otherwise).
 *
 * Each declaration names the function of the reference (a pure PyTorch package)
 * whose work it replaces.  The reference has no FFI of its own; the Python modules
 * of this package bind these entry points with ctypes behind the reference's own
 * call signatures.
 */
#ifndef STTOPT_HIP_H
#define STTOPT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STTOPT_OK 0
#define STTOPT_EINVAL (-1)     /* bad sizes, null pointers, unsupported mode */
#define STTOPT_EHIP (-2)       /* a HIP runtime call failed */
#define STTOPT_ENOCONV (-3)    /* PCG reached max_iter with rows above rtol */
#define STTOPT_ENOTSPD (-4)    /* a dense matrix is not positive definite */

#define STTOPT_MAX_LEVELS 16

/* dtype codes */
#define STTOPT_F64 0
#define STTOPT_F32 1

/* layouts of coarse element matrices (entry-major, [rows][entries][nel]):
 * FULL: the (4d)^2 entries a * 4d + b; UPPER: the (4d)(4d+1)/2 entries a <= b of the
 * symmetric matrix, row-major over a (entry index a * 4d - a (a - 1) / 2 + b - a).
 * NODE: the assembled operator as a node stencil, one record per node
 * [rows][nodes][4 + 4 d^2]: the upper triangle of the node's own d x d block (row-major,
 * zero padded to 4 values), then its d x d blocks (row: the node's dof, column: the
 * neighbour's) to the neighbours at (node row, column) + (0, 1), (1, -1), (1, 0),
 * (1, 1); the other four neighbours' blocks are theirs, transposed (20 values per node
 * at d = 2; records and blocks 16-byte aligned for vector loads) */
#define STTOPT_KEFF_FULL 0
#define STTOPT_KEFF_UPPER 1
#define STTOPT_KEFF_NODE 2

/* sttopt_level.flags */
#define STTOPT_SEPARATE_SWEEPS 1 /* level 0: one pass per smoothing sweep at every level
                                    instead of the fused passes (same results; for
                                    testing) */
#define STTOPT_LEVEL_SEPARATE 2  /* a node-stencil level of the fp32 cycle: one pass per
                                    sweep, residual and transfer at this level instead of
                                    its two passes (same results) */
#define STTOPT_TILE_VARIANT_SHIFT 8 /* bits 8-11: the tile schedule of this level's two
                                       passes (0: 32 x 4 coarse / 64 x 8 fine nodes per
                                       512-thread workgroup, 1: 16 x 4 / 32 x 8 per 256) */

/* operator modes for sttopt_apply */
#define STTOPT_MODE_MATVEC 0   /* y = K u (no projection)                 */
#define STTOPT_MODE_OPERATOR 1 /* y = P (K P u + rdiag * P u)              */
#define STTOPT_MODE_OPERATOR_PROJECTED 2 /* y = P (K u + rdiag * u) for a u that is
                                            already 0 at fixed dofs (a PCG direction):
                                            the outer iteration's apply.  No
                                            counterpart in the reference, whose pcg calls
                                            operator() (the input projection is then a
                                            no-op). */

const char* sttopt_last_error(void);
int sttopt_version(void);
/* sizeof(sttopt_level) and sizeof(sttopt_mg), for binding checks. */
int sttopt_abi_sizes(int64_t* level_bytes, int64_t* mg_bytes);

/* ---------------------------------------------------------------- masks ---- */

/* Pack a bool/uint8 free-dof mask (1 = free) into bits.  Words = (ndof+31)/32.
 * Replaces torch_fem.free_mask's bool tensor on the device (torch_fem.py:101-107). */
int sttopt_pack_mask(const uint8_t* mask, int64_t ndof, uint32_t* bits, void* stream);

/* Coarse mask = fine mask sampled at every kx-th node column / ky-th node row
 * (torch_mg.py:376). */
int sttopt_coarsen_mask(const uint32_t* fine_bits, int32_t nelx, int32_t nely,
                        int32_t d, int32_t kx, int32_t ky, uint32_t* coarse_bits,
                        void* stream);

/* ------------------------------------------------------------ operators ---- */

/* Fine-level matrix-free apply y = sum_e density_e A_e^T KE A_e u, optionally
 * projected and with a lumped reaction diagonal.
 * Replaces torch_fem.matvec (torch_fem.py:62-83) in mode MATVEC and
 * torch_fem.operator (torch_fem.py:135-159) in mode OPERATOR.
 * ke: (4d)x(4d) row-major doubles (host memory).  mask/rdiag may be NULL. */
int sttopt_apply(int32_t dtype, int32_t mode, int32_t d, int32_t nelx, int32_t nely,
                 int32_t rows, const void* u, int64_t u_stride, const void* density,
                 int64_t density_stride, const double* ke, const uint32_t* mask,
                 const void* rdiag, int64_t rdiag_stride, void* y, int64_t y_stride,
                 void* stream);

/* diag(K) (+ rdiag), with 1 at fixed dofs when mask != NULL.
 * Replaces torch_fem.matvec_diagonal / jacobi_preconditioner_diag
 * (torch_fem.py:86-98, 162-181). */
int sttopt_diagonal(int32_t dtype, int32_t d, int32_t nelx, int32_t nely, int32_t rows,
                    const void* density, int64_t density_stride, const double* ke,
                    const uint32_t* mask, const void* rdiag, int64_t rdiag_stride,
                    void* out, int64_t out_stride, void* stream);

/* Lumped reaction diagonal: each element adds reaction_e / 4 to each of its dofs.
 * Replaces torch_fem.lumped_reaction (torch_fem.py:117-132). */
int sttopt_lumped_reaction(int32_t dtype, int32_t d, int32_t nelx, int32_t nely,
                           int32_t rows, const void* reaction, int64_t reaction_stride,
                           void* out, int64_t out_stride, void* stream);

/* Coarse-level apply with one explicit (4d)x(4d) matrix per element, stored
 * entry-major (structure of arrays): keff[row][a * 4d + b][e].
 * Replaces torch_mg.keff_matvec + _Level.__call__ (torch_mg.py:179-193, 239-255). */
int sttopt_apply_keff(int32_t dtype, int32_t d, int32_t nelx, int32_t nely,
                      int32_t rows, const void* u, int64_t u_stride, const void* keff,
                      int64_t keff_stride, const uint32_t* mask, void* y,
                      int64_t y_stride, void* stream);

/* --------------------------------------------------------- multigrid -------- */

typedef struct sttopt_level {
  int32_t nelx, nely;     /* elements of this level                              */
  int32_t kx, ky;         /* coarsening factors to the next level; 1,1 at last   */
  const uint32_t* mask;   /* packed free-dof bits of this level                  */
  /* level 0: SIMP density (+ optional lumped reaction diagonal), in the V-cycle
   * precision; coarser levels: entry-major element matrices keff                */
  const void* density;
  int64_t density_stride;
  const void* rdiag;
  int64_t rdiag_stride;
  const void* keff;
  int64_t keff_stride;
  int32_t keff_layout;    /* STTOPT_KEFF_FULL, STTOPT_KEFF_UPPER or STTOPT_KEFF_NODE */
  int32_t flags;          /* STTOPT_SEPARATE_SWEEPS (level 0), STTOPT_LEVEL_SEPARATE,
                             tile variant << STTOPT_TILE_VARIANT_SHIFT; 0 default   */
  /* coarser levels, alternative to keff: keff_E = sum over the pkx x pky level-0
   * elements g under E of density0_g * ptable[g], ptable: (pkx*pky) x (4d)^2 in the
   * V-cycle precision (device); the V-cycle then reads only level 0's density      */
  const void* ptable;
  int32_t pkx, pky;
  const void* invdiag;    /* 1 / diag at free dofs, 0 at fixed, V-cycle precision */
  int64_t invdiag_stride;
  const double* cinv;     /* coarsest level: dense inverse [rows][n][n], fp64     */
  int64_t cinv_stride;
  /* per-level V-cycle work vectors, each [rows][ndof] in the V-cycle precision   */
  void* vb;
  void* vx;
  void* vx2;
  void* vr;
  void* ve;               /* ve, vw: only needed when gamma > 1                   */
  void* vw;
} sttopt_level;

typedef struct sttopt_mg {
  int32_t rows;           /* batch rows                                           */
  int32_t d;              /* dofs per node (2 elasticity, 1 scalar)               */
  int32_t nlevels;
  int32_t vdtype;         /* V-cycle precision: STTOPT_F64 or STTOPT_F32          */
  int32_t n_smooth;
  int32_t gamma;
  double omega;
  double ke[64];          /* reference element matrix, (4d)x(4d) row-major        */
  /* the fp64 outer operator (level 0)                                            */
  const double* density64;
  int64_t density64_stride;
  const double* rdiag64;
  int64_t rdiag64_stride;
  sttopt_level levels[STTOPT_MAX_LEVELS];
  /* device scratch sized by sttopt_mg_scratch_bytes                              */
  void* scratch;
  /* optional device word of sttopt_spd_inverse_band_async for the coarsest inverse:
   * sttopt_mgpcg and sttopt_mgpcg_batch read it with their final readback and return
   * STTOPT_ENOTSPD when it is nonzero                                              */
  const int32_t* cinv_status;
  /* optional hipEvent_t recorded after the coarsest inverse's launches when they were
   * queued on another stream: every V-cycle makes its stream wait for it before its
   * first use of the inverse, so the level passes above the coarsest run beside the
   * inverse.  Null: the inverse is ordered on the solve's stream already.          */
  void* cinv_ready;
} sttopt_mg;

/* Upper-triangle copy of fp64 FULL element matrices, (A + A^T) / 2, in dtype (V-cycle
 * storage of the coarse levels; 36 of 64 entries at d = 2).  No counterpart in the
 * reference (whose coarse matrices stay full, torch_mg.py:355-370). */
int sttopt_pack_keff(int32_t dtype, int32_t d, int64_t nel, int32_t rows, const double* keff,
                     int64_t keff_stride, void* out, int64_t out_stride, void* stream);

/* The STTOPT_KEFF_NODE form, in dtype, of fp64 FULL element matrices on an nelx x nely
 * grid (each coupling summed in fp64 from (A + A^T) / 2, rounded once).  No
 * counterpart in the reference (whose coarse levels apply element matrices,
 * torch_mg.py:179-193). */
int sttopt_node_stencil(int32_t dtype, int32_t d, int32_t nelx, int32_t nely, int32_t rows,
                        const double* keff, int64_t keff_stride, void* out, int64_t out_stride,
                        void* stream);

/* Bytes of device scratch the PCG / V-cycle need for this hierarchy (per-row
 * scalars, reduction partials, flags).  levels[*].nelx/nely and rows must be set. */
int64_t sttopt_mg_scratch_bytes(const sttopt_mg* mg);

/* Galerkin coarse element matrices, fp64, entry-major.
 * Density form (fine_keff = NULL): each coarse element covers a kx x ky block of
 * fine elements of an nelx x nely grid given by density (+ reaction) per element;
 * keff_E = sum over the block of density_g G[pos g] (+ reaction_g GM[pos g]), with
 * G (and GM, may be NULL) device arrays of (kx*ky) x (4d)^2 doubles in row-major block
 * order.  With G = S^T KE S (GM = S^T (I/4) S) over one coarsening step this is
 * level 1; with G the per-position products of KE through several steps it is any
 * deeper level directly (blocks up to 4096 fine elements).
 * Explicit form: pass fine_keff (level l) and the interpolation matrices S (device,
 * (kx*ky) x (4d)^2, S[m][p][a]: child m's local dof p from the coarse element's local
 * dof a), kx, ky <= 3.
 * Replaces the einsums of torch_mg.build_hierarchy (torch_mg.py:355-370). */
int sttopt_galerkin(int32_t d, int32_t nelx, int32_t nely, int32_t kx, int32_t ky,
                    int32_t rows, const double* density, int64_t density_stride,
                    const double* reaction, int64_t reaction_stride,
                    const double* g, const double* gm, const double* fine_keff,
                    int64_t fine_keff_stride, const double* s, double* coarse_keff,
                    int64_t coarse_keff_stride, void* stream);

/* Jacobi diagonal (fp64, 1 at fixed dofs) of the nelx x nely level whose element
 * matrices are sum over the kx x ky level-0 elements g under each element of
 * rho_g G_pos(g) (kx ky <= 4): density is the level-0 field (rows, (kx nelx)(ky nely)),
 * table_diag a host array of the ky x kx positions' diagonals (4d each, position-major),
 * mask the level's free-dof bits.  Lets that level skip forming its element matrices.
 * No counterpart in the reference (whose diagonal comes from the assembled coarse
 * matrices, torch_mg.py:377-378). */
int sttopt_param_diagonal(int32_t d, int32_t nelx, int32_t nely, int32_t kx, int32_t ky,
                          int32_t rows, const double* density, int64_t density_stride,
                          const double* table_diag, const uint32_t* mask, double* out,
                          int64_t out_stride, void* stream);

/* Jacobi diagonal of a keff level: 1/diag at free dofs and 0 at fixed (inv != 0),
 * or diag with 1 at fixed dofs (inv == 0).  Replaces torch_mg.py:377-378. */
int sttopt_keff_diagonal(int32_t dtype_out, int32_t d, int32_t nelx, int32_t nely,
                         int32_t rows, const double* keff, int64_t keff_stride,
                         const uint32_t* mask, int32_t inv, void* out,
                         int64_t out_stride, void* stream);

/* Inverse of a diagonal with the fixed-dof convention above, with dtype cast. */
int sttopt_invert_diagonal(int32_t dtype_out, int64_t n, int32_t rows,
                           const double* diag, int64_t diag_stride,
                           const uint32_t* mask, void* out, int64_t out_stride,
                           void* stream);

/* Dense projected operator of the coarsest level: symmetrized, identity rows and
 * columns at fixed dofs.  out: [rows][n][n], n = ndof of the level.  Either keff
 * (a coarse level) or density with ke (a device array here) and optional rdiag
 * (an uncoarsened fine level).
 * Replaces torch_mg._dense_cholesky's assembly (torch_mg.py:269-289). */
int sttopt_dense_assemble(int32_t d, int32_t nelx, int32_t nely, int32_t rows,
                          const double* keff, int64_t keff_stride,
                          const double* density, int64_t density_stride,
                          const double* rdiag, int64_t rdiag_stride,
                          const double* ke, const uint32_t* mask, double* out,
                          int64_t out_stride, void* stream);

/* In-place inverse of a batch of dense SPD matrices (the coarsest operator) by the
 * block symmetric sweep.  a: rows matrices of n x n fp64, matrix b at a + b * a_stride;
 * work: sttopt_spd_inverse_work_count(n, rows) doubles.  Synchronizes the stream once
 * (to read the definiteness flag).  STTOPT_ENOTSPD when a pivot is not positive (a is
 * then partially overwritten).
 * Replaces torch.linalg.cholesky of the coarsest level (torch_mg.py:269-290) with the
 * factor's use by the coarse solve (torch_mg.py:448-451) folded into the inverse. */
int64_t sttopt_spd_inverse_work_count(int32_t n, int32_t rows);
int sttopt_spd_inverse(int32_t n, int32_t rows, double* a, int64_t a_stride, double* work,
                       void* stream);

/* sttopt_spd_inverse of banded matrices: every entry (i, j) with |i - j| > band must be
 * zero (a Q4 grid's assembled operator with nnx nodes per row and d dofs per node has
 * band = d (nnx + 1) + d - 1).  The sweep then skips the tiles that stay zero until it
 * reaches them.  band >= n - 1 is the dense sweep.  Same work block, result and errors
 * as sttopt_spd_inverse; the inverse itself is dense.
 * Replaces the same reference step (torch_mg.py:269-290, 448-451). */
int sttopt_spd_inverse_band(int32_t n, int32_t band, int32_t rows, double* a,
                            int64_t a_stride, double* work, void* stream);

/* sttopt_spd_inverse_band without the wait for its result: the launches are queued on
 * the stream and the device word *not_spd (device memory) receives 1 when a matrix is
 * not positive definite, else 0.  The caller reads it, or hands it to the solvers in
 * sttopt_mg.cinv_status, which report STTOPT_ENOTSPD with their final readback.  Lets the
 * host queue the rest of a hierarchy build and the solve while the inverse runs.
 * Replaces the same reference step (torch_mg.py:269-290: the factorization's error is
 * raised before the solve returns). */
int sttopt_spd_inverse_band_async(int32_t n, int32_t band, int32_t rows, double* a,
                                  int64_t a_stride, double* work, int32_t* not_spd,
                                  void* stream);

/* Grid transfers on the node grid (torch_mg._on_node_grid with _restrict_axis /
 * _interp_axis, torch_mg.py:75-109).  restrict: coarse = P_c (R fine);
 * prolong_add: fine += P_f (P coarse).  Masks may be NULL. */
int sttopt_restrict(int32_t dtype, int32_t d, int32_t cnelx, int32_t cnely,
                    int32_t kx, int32_t ky, int32_t rows, const void* fine,
                    int64_t fine_stride, const uint32_t* cmask, void* coarse,
                    int64_t coarse_stride, void* stream);
int sttopt_prolong_add(int32_t dtype, int32_t d, int32_t cnelx, int32_t cnely,
                       int32_t kx, int32_t ky, int32_t rows, const void* coarse,
                       int64_t coarse_stride, const uint32_t* fmask, void* fine,
                       int64_t fine_stride, void* stream);

/* One V-cycle z = M(r) on every active row (torch_mg.VCycle.__call__,
 * torch_mg.py:440-476).  r, z fp64 [rows][ndof0]. */
int sttopt_vcycle(const sttopt_mg* mg, const double* r, int64_t r_stride, double* z,
                  int64_t z_stride, void* stream);

/* Apply level l of the hierarchy in the V-cycle precision (torch_mg._Level.__call__). */
int sttopt_level_apply(const sttopt_mg* mg, int32_t level, const void* v,
                       int64_t v_stride, void* y, int64_t y_stride, void* stream);

/* Batched multigrid-preconditioned CG on rows of b (torch_fem.pcg with the
 * torch_mg levels, torch_fem.py:201-343).  x holds the warm start on entry (its fixed
 * dofs are ignored) and the solution on return; with cold_start != 0 the start is
 * x = 0 and x is not read (r = P b and x = 0 are written in one pass).
 * Rows retire as they converge (per-row device flags): iters[row] receives each
 * row's iteration count and resid[row] its final relative residual.  flexible != 0
 * uses the Polak-Ribiere beta (for a V-cycle in fp32).  check_every: iterations
 * between convergence checks (a mark a kernel stores in mapped pinned host memory
 * when rows are still active); each is read one check late
 * (the device keeps the next iteration queued), so the solve may launch one more
 * iteration that finds no active row.  Row b of x is at x + b * x_stride (x_stride at
 * least the level-0 dof count when rows > 1); b_stride may be 0 (one load for every
 * row).  The stream is synchronized at the start, at
 * max_iter and at the end.  Returns STTOPT_ENOCONV when a row is still above rtol
 * after max_iter iterations (x, iters, resid are still written). */
int sttopt_mgpcg(const sttopt_mg* mg, const double* b, int64_t b_stride, double* x,
                 int64_t x_stride, double rtol, int32_t max_iter, int32_t flexible,
                 int32_t check_every, int32_t cold_start, double* work, int32_t* iters,
                 double* resid, int32_t* n_iter, void* stream);

/* One hierarchy of an uneven batch for sttopt_mgpcg_batch: the hierarchy (mg, with its
 * rows), right-hand sides b, solutions x (warm start on entry unless cold_start), the
 * sttopt_mgpcg_work_count(mg) doubles of work, and per-row iteration counts and final
 * relative residuals. */
typedef struct sttopt_batch_group {
  const sttopt_mg* mg;
  const double* b;
  int64_t b_stride;
  double* x;
  int64_t x_stride;
  double* work;
  int32_t* iters;
  double* resid;
} sttopt_batch_group;

/* One batched MG-PCG over hierarchies of different grids (instances of uneven size): the
 * solve of sttopt_mgpcg for every group at once, each pass of an iteration one launch
 * over all groups (up to 4 per launch), each group with its own workgroups and
 * arguments, rows retiring on their own.  Every group needs the fp32 two-sweep cycle
 * (vdtype F32, n_smooth 2, gamma 1, flexible beta) coarsened 2 x 2 down to its
 * coarsest level, level 1 density-parameterized and deeper levels node stencils
 * (STTOPT_EINVAL otherwise), and the same dofs per node and omega.  Results equal
 * sttopt_mgpcg on each group alone.  n_iter: the slowest row's count over all groups;
 * STTOPT_ENOCONV as sttopt_mgpcg.  No counterpart in the reference, which solves one
 * grid per call (torch_fem.py:62-83 takes one element dof map). */
int sttopt_mgpcg_batch(const sttopt_batch_group* groups, int32_t ngroups, double rtol,
                       int32_t max_iter, int32_t check_every, int32_t cold_start,
                       int32_t* n_iter, void* stream);

/* Whether sttopt_mgpcg replays the second half of each iteration (the V-cycle's coarse
 * visit and post-smoothing, the beta finalizer and the direction update: the same
 * launches every iteration) from a HIP graph captured once per solve (enable 1, the
 * default) or launches it kernel by kernel (0); the results are identical.  enable < 0
 * only queries.  Returns the previous setting; process-wide.  No counterpart in the
 * reference (whose iteration is a host loop of torch ops, torch_fem.py:313-331). */
int sttopt_mgpcg_graphs(int32_t enable);

/* How many of the fp32 V-cycle's two-pass node-stencil levels above the coarsest run,
 * with the coarsest solve, as one persistent launch per cycle (0 to 3; 0: a launch per
 * pass).  The launch takes the passes' tiles as one work list, phase after phase, with
 * per-row completion counters in the sttopt_mg scratch (zeroed by sttopt_mgpcg and
 * sttopt_vcycle, left at zero by the launch) and bounded waits; the results are
 * identical.  Process-wide; STTOPT_EINVAL outside 0 to 3.  No counterpart in the
 * reference (whose cycle recurses level by level in torch ops, torch_mg.py:440-471). */
int sttopt_vcycle_bottom(int32_t max_levels);

/* Work vectors sttopt_mgpcg needs: returns the count of doubles ([rows][ndof0] x 6). */
int64_t sttopt_mgpcg_work_count(const sttopt_mg* mg);

/* ---------------------------------------------------------- adjoints -------- */

/* FemSolve.backward element terms (torch_solve.py:144-152):
 * grad_density[row][e] = -(lam_e KE) . U_e (if grad_density != NULL),
 * grad_reaction[row][e] = -sum(lam_e * U_e) / 4 (if grad_reaction != NULL). */
int sttopt_adjoint_grad(int32_t d, int32_t nelx, int32_t nely, int32_t rows,
                        const double* u, int64_t u_stride, const double* lam,
                        int64_t lam_stride, const double* ke, double* grad_density,
                        int64_t gd_stride, double* grad_reaction, int64_t gr_stride,
                        void* stream);

/* Element strain energy c_e = U_e^T KE U_e, per row (compliance.py:49-59). */
int sttopt_strain_energy(int32_t d, int32_t nelx, int32_t nely, int32_t rows,
                         const double* u, int64_t u_stride, const double* ke,
                         double* ce, int64_t ce_stride, void* stream);

/* Space-time compliance prologue (compliance.py:33-46, 137-161, 226-243):
 * row 0: density = SIMP(x), load = F_ext; row k = 1..nstage: m = H_beta(t_k - t),
 * xt = x * m, density = SIMP(xt), load_y = -fe/4 * sum over the node's elements
 * of xt, load_x = 0.  density: [1+nstage][nel], load: [1+nstage][ndof]. */
int sttopt_stage_prologue(int32_t nelx, int32_t nely, int32_t nstage,
                          const double* x, const double* t, const double* stage_times,
                          double beta, double emin, double emax, double penal,
                          double fe, const double* f_ext, double* density,
                          double* load, void* stream);

/* Space-time compliance epilogue and its hand-derived adjoint
 * (compliance.py:244-253; tests/reference/compliance.py:24-110).
 * When comp != NULL: comp[row] = sum_e SIMP(xt_row)_e * c_e(U_row), with the
 * SIMP densities read from density ([1+nstage][nel], as sttopt_stage_prologue wrote
 * them; required with comp).  When dx != NULL, also
 * dx[e] = sum_rows g[row] dc_row/dx_e and dt[e] = sum_rows g[row] dc_row/dt_e with
 * the closed forms of the self-adjoint compliance (lambda = 2U), including the
 * gravity load term -2 (C^T U_y)_e.  scratch: sttopt_stage_scratch_count doubles. */
int sttopt_stage_epilogue(int32_t nelx, int32_t nely, int32_t nstage,
                          const double* x, const double* t, const double* stage_times,
                          double beta, double emin, double emax, double penal,
                          double fe, const double* ke, const double* u,
                          const double* density, const double* g, double* comp,
                          double* dx, double* dt, double* scratch, void* stream);
int64_t sttopt_stage_scratch_count(int32_t nelx, int32_t nely, int32_t nstage);

/* sttopt_stage_epilogue without the scratch argument and without any allocation: the
 * compliance partials go to a static device array of STTOPT_NOALLOC_SCRATCH_DOUBLES
 * doubles (2 MiB, held by every GPU the library runs on until the process ends; its
 * device address is looked up once per device, on the first call with that device
 * current).  No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  STTOPT_EINVAL when comp != NULL and
 * sttopt_stage_scratch_count(nelx, nely, nstage) exceeds the array (1024 x 512 with up
 * to 127 stages fits, 2048 x 1024 with up to 31).  The stream must belong to the
 * current device.  All calls on one GPU share its array: calls that overlap on that
 * GPU (other streams or threads) race; calls in order on one stream do not. */
#define STTOPT_NOALLOC_SCRATCH_DOUBLES 262144
int sttopt_stage_epilogue_noalloc(int32_t nelx, int32_t nely, int32_t nstage,
                                  const double* x, const double* t,
                                  const double* stage_times, double beta, double emin,
                                  double emax, double penal, double fe, const double* ke,
                                  const double* u, const double* density, const double* g,
                                  double* comp, double* dx, double* dt, void* stream);

/* HIP-event timing of the outer PCG's K(rho)p launches inside sttopt_mgpcg (off by
 * default; process-wide, not thread-safe): from iteration 2 on, the apply that carries
 * the direction update (iteration 1's plain apply is not counted).  enable resets the
 * totals; read returns the summed launch durations, the algorithmic bytes (per row still
 * active in the launch's iteration 8 (5 ndof + nel) + s ndof, z in s-byte words; 24 ndof
 * per row that retired the iteration before) and the launch count.  Measurement only: no
 * counterpart in the reference. */
int sttopt_profile_outer_apply(int32_t enable);
int sttopt_profile_outer_apply_read(double* seconds, double* bytes, int64_t* launches);

/* ---- Gauss-point gradient fields of the scalar (d = 1) mesh ------------------------
 * Element fields X are [rows][nel][4][2] doubles (Gauss point g = 2 * eta_index +
 * xi_index, then x, y), 16-byte aligned with even row strides.  B: host array of
 * 4 x 2 x 4 doubles, the shape-function gradients at the Gauss points in the local node
 * order BL, BR, TR, TL (ScalarMesh.B). */

/* out[row][e][g][d] = scale * escale[row][e] * sum_i B[g][d][i] u[row][node_i(e)]
 * (escale NULL: 1).  normalize != 0: the direction field G / sqrt(|G|^2 + DBL_MIN)
 * per Gauss point instead (scale and escale unused).
 * Replaces ScalarMesh.gauss_gradients and direction_field (virtual_heat.py:290-307). */
int sttopt_gauss_gradients(int32_t nelx, int32_t nely, int32_t rows, const double* B,
                           const double* u, int64_t u_stride, const double* escale,
                           int64_t escale_stride, double scale, int32_t normalize,
                           double* out, int64_t out_stride, void* stream);

/* F[row][n] = scale * sum over the elements e at node n of
 * w[row][e] sum_{g,d} X[row][e][g][d] B[g][d][loc_n(e)] (w NULL: 1), gathered per node
 * in a fixed order (no atomics).  With scale 1/4: ScalarMesh.divergence_load
 * (virtual_heat.py:309-315); with w NULL, scale 1: the adjoint of gauss_gradients. */
int sttopt_divergence_load(int32_t nelx, int32_t nely, int32_t rows, const double* B,
                           const double* w, int64_t w_stride, const double* X,
                           int64_t X_stride, double scale, double* F, int64_t F_stride,
                           void* stream);

/* out[row][e] = scale * sum_{g,d} X[row][e][g][d] sum_i B[g][d][i] v[row][node_i(e)]:
 * the weight gradient of divergence_load.  No counterpart as a function in the
 * reference (autograd of virtual_heat.py:309-315). */
int sttopt_gauss_dot(int32_t nelx, int32_t nely, int32_t rows, const double* B,
                     const double* X, int64_t X_stride, const double* v, int64_t v_stride,
                     double scale, double* out, int64_t out_stride, void* stream);

/* Adjoint of the direction field at the Gauss points: gG = (gY - Y (Y . gY)) / n with
 * G the gradients of s, n = sqrt(|G|^2 + DBL_MIN), Y = G / n.  No counterpart as a
 * function in the reference (autograd of virtual_heat.py:294-307). */
int sttopt_direction_field_backward(int32_t nelx, int32_t nely, int32_t rows,
                                    const double* B, const double* s, int64_t s_stride,
                                    const double* gY, int64_t gY_stride, double* gG,
                                    int64_t gG_stride, void* stream);

/* ---- density filter and projection ------------------------------------------------
 * scale_mode: 0 y = H x; 1 y = (H x) / scale (per output entry); 2 y = H (x / scale)
 * (per input entry: the adjoint of mode 1 for symmetric H).  scale: device array. */

/* The density filter H[e1, e2] = max(0, rmin - |e1 - e2|) over the offsets
 * |di|, |dj| <= ceil(rmin) - 1 inside the nelx x nely element grid, applied matrix-free
 * to rows fields x ([rows][nel], row-major elements).  With scale = Hs and mode 1 this is
 * filters.apply_density_filter (filters.py:67-76) for the H, Hs of
 * filters.density_filter (filters.py:38-64). */
int sttopt_density_filter(int32_t nelx, int32_t nely, int32_t rows, double rmin,
                          const double* x, int64_t x_stride, const double* scale,
                          int32_t scale_mode, double* y, int64_t y_stride, void* stream);

/* y = A x for an n x n CSR matrix (int64 row pointers and column indices, fp64 values,
 * all device arrays), rows right-hand sides.  For a symmetric A this is
 * torch_util.symmetric_matmul (torch_util.py:157-188), forward and backward. */
int sttopt_csr_matvec(int64_t n, const int64_t* row_ptr, const int64_t* col_idx,
                      const double* values, int32_t rows, const double* x, int64_t x_stride,
                      const double* scale, int32_t scale_mode, double* y, int64_t y_stride,
                      void* stream);

/* y = H_beta,eta(x) elementwise (filters.heaviside_projection, filters.py:108-125; beta
 * = +inf: the step with 1/2 at eta) or, derivative != 0, its derivative
 * (heaviside_projection_derivative, filters.py:128-131; finite beta only), times g when
 * g != NULL. */
int sttopt_heaviside(int64_t n, double beta, double eta, int32_t derivative, const double* x,
                     const double* g, double* y, void* stream);

/* ---- MMA subproblem ------------------------------------------------------------------
 * All arrays fp64 device memory; P, Q (and dfdx) [m][n] row-major, 1 <= m <= 126.
 *
 * sttopt_mma_prepare: the asymptotes (re-initialized when iteration < 2, else adapted
 * from xold1, xold2 and clamped), the move-limited bounds alfa, beta, and the
 * subproblem's p0, q0, P, Q and b = P / (upp - x) + Q / (x - low) - fval
 * (mma.mmasub, mma.py:150-276).  asyinit, asymin, asymax, move: n-vectors or scalars
 * (stride 0).  work: sttopt_mma_prepare_work_count doubles.
 *
 * sttopt_mma_subsolv: one step of the primal-dual interior-point method of mma.subsolv
 * (mma.py:346-557) on a device state kept in work (sttopt_mma_work_count doubles):
 * op 0 the initial point; op 1 the residual (its 2-norm and max-abs written to the
 * state's output pair: the last two doubles of the state block returned by
 * sttopt_mma_state_offsets) at the current point (jj < 0) or at the line-search
 * candidate steg / 2^jj; op 2 the Newton direction and step bound steg at barrier epsi
 * (m < n: the (m+1) system in lam and z; m >= n: the (n+1) system in x and z, needing
 * n <= 126); op 3 the move to candidate jj.  The host drives the barrier and Newton
 * loops (the reference's while loops), reading the residual pair after each op 1. */
int64_t sttopt_mma_prepare_work_count(int32_t m, int32_t n);
int sttopt_mma_prepare(int32_t m, int32_t n, int32_t iteration, const double* xval,
                       const double* xmin, const double* xmax, const double* xold1,
                       const double* xold2, const double* df0dx, const double* fval,
                       const double* dfdx, double* low, double* upp, const double* asyinit,
                       int64_t asyinit_stride, const double* asymin, int64_t asymin_stride,
                       const double* asymax, int64_t asymax_stride, const double* move,
                       int64_t move_stride, double raa0, double albefa, double asyincr,
                       double asydecr, double* alfa, double* beta, double* p0, double* q0,
                       double* P, double* Q, double* b, double* work, void* stream);
int64_t sttopt_mma_work_count(int32_t m, int32_t n);
/* Offsets (in doubles, into work) of x, xsi, eta (n each), y, lam, mu, s (m each), the
 * scalars z, zet, and the residual pair (norm, max): offsets[0..9]. */
int sttopt_mma_state_offsets(int32_t m, int32_t n, int64_t* offsets);
int sttopt_mma_subsolv(int32_t m, int32_t n, const double* low, const double* upp,
                       const double* alfa, const double* beta, const double* p0,
                       const double* q0, const double* P, const double* Q, double a0,
                       const double* a, const double* b, const double* c, const double* d,
                       double* work, int32_t op, double epsi, int32_t jj, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* STTOPT_HIP_H */
